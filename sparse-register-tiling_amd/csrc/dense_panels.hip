// Dense 16-row panels of a hybrid plan on FP32 matrix cores.
//
// The reference pads column patterns up to denser kernel patterns by its ILP
// cost model (sbench/ilp_pad/nano_solver.py:76-116) and calls a tile dense
// past a density threshold (sbench/padding_reordering.py:36, 0.30).  A
// hybrid plan takes the 16-row panels at or above its threshold out of the
// register-tiled packing and stores them dense for v_mfma_f32_16x16x4_f32:
// per panel and 16-column K group, lane l holds the four k-steps it feeds,
//   A[row = l % 16][k = 16 g + 4 j + l / 16],  j = 0..3,
// as one 16-byte word (one load per lane per K group).
//
// A workgroup (4 waves) owns 4 dense panels x 64 columns of C: each K group's
// B tile (16 rows x 64 columns) is staged in LDS (double-buffered) and every
// wave runs 4 k-steps x 4 column tiles of MFMAs for its panel.  Accumulation
// is FP32 in the matrix core, k ascending in steps of 4.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "srt_internal.h"

namespace srt {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kRows = 16;      // panel rows (MFMA M)
constexpr int kGroup = 16;     // K columns per staged group (4 k-steps of 4)
constexpr int kNT = 64;        // C columns per workgroup (4 MFMA N tiles)
constexpr int kWaves = 4;      // panels per workgroup

__global__ void __launch_bounds__(64 * kWaves)
    dense_panels_kernel(int ndense, int G, const f32x4 *__restrict__ apack, const int32_t *__restrict__ panel_row,
                        const float *__restrict__ B, int64_t ldb, float *__restrict__ C, int64_t ldc, int M, int N,
                        int K) {
    __shared__ float bt[2][kGroup][kNT];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int d = blockIdx.x * kWaves + wave;  // dense panel of this wave (may be past the list)
    const bool live = d < ndense;
    const int n0 = blockIdx.y * kNT;
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    // stage group g of B into bt[buf]: 16 rows x 64 columns, 4 per thread,
    // zeros past K and N
    auto stage = [&](int g, int buf) {
        const int r = threadIdx.x >> 4, c = (threadIdx.x & 15) * 4;
        const int k = g * kGroup + r;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int n = n0 + c + e;
            bt[buf][r][c + e] = (k < K && n < N) ? B[(int64_t)k * ldb + n] : 0.0f;
        }
    };
    const f32x4 *ap = apack + (int64_t)(live ? d : 0) * G * 64 + lane;
    stage(0, 0);
    __syncthreads();
    for (int g = 0; g < G; ++g) {
        const int buf = g & 1;
        if (g + 1 < G) stage(g + 1, buf ^ 1);
        if (live) {
            const f32x4 a = ap[(int64_t)g * 64];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int kr = 4 * j + (lane >> 4);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float b = bt[buf][kr][16 * t + (lane & 15)];
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j], b, acc[t], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    if (!live) return;
    const int row0 = panel_row[d];
    // D layout: lane l holds rows 4 * (l / 16) + r of column l % 16.
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int n = n0 + 16 * t + (lane & 15);
        if (n >= N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = row0 + 4 * (lane >> 4) + r;
            if (row < M) C[(int64_t)row * ldc + n] = acc[t][r];
        }
    }
}

}  // namespace

hipError_t launch_dense_panels(int ndense, int G, const void *apack, const int32_t *panel_row, const float *B,
                               int64_t ldb, float *C, int64_t ldc, int M, int N, int K, hipStream_t s) {
    if (ndense <= 0 || N <= 0) return hipSuccess;
    const dim3 grid((unsigned)((ndense + kWaves - 1) / kWaves), (unsigned)((N + kNT - 1) / kNT));
    hipLaunchKernelGGL(dense_panels_kernel, grid, dim3(64 * kWaves), 0, s, ndense, G,
                       static_cast<const f32x4 *>(apack), panel_row, B, ldb, C, ldc, M, N, K);
    return hipGetLastError();
}

}  // namespace srt
