// Dense 16-row panels of a hybrid plan on FP32 matrix cores.
//
// The reference pads column patterns up to denser kernel patterns by its ILP
// cost model (sbench/ilp_pad/nano_solver.py:76-116) and calls a tile dense
// past a density threshold (sbench/padding_reordering.py:36, 0.30).  A
// hybrid plan takes the 16-row panels at or above its threshold out of the
// register-tiled packing and stores them dense for v_mfma_f32_16x16x4_f32:
// per panel and 16-column K group g, lane l holds the four k-steps it feeds,
//   A[row = l % 16][k = 16 g + 4 j + l / 16],  j = 0..3,
// as one 16-byte word (one load per lane per K group).
//
// Output tiles are 8 panels (128 rows) x 128 columns; K is split into S
// ranges (split-K) so that tiles x S fills the chip however few panels are
// dense.  A workgroup (8 waves, two per SIMD) owns one (tile, K range): wave
// w computes panels 4 (w / 4) .. + 3 of the tile against column tiles
// 2 (w % 4), 2 (w % 4) + 1 (eight 16 x 16 accumulators).  K advances in
// stages of 64 (4 groups) through two LDS buffers holding the stage's A
// words (as stored) and B block in MFMA B-operand order,
//   lb[group q][column tile t][lane l] = B[16 q + 4 j + l / 16][16 t + l % 16], j = 0..3.
// Per stage: the next stage's global loads are issued first, the MFMAs of
// group q run while group q + 1's operands are read from LDS, the next
// stage's words are written to the other buffer after group 1, and the one
// barrier sits before group 3, whose MFMAs hide the next stage's first
// reads.  With S > 1 each workgroup stores its accumulators (lane order) to
// a workspace and dense_tiles_reduce adds the S partials in split order and
// writes C, so results repeat bit for bit; with S = 1 the workgroup writes C.
//
// Workgroups map to (split, tile) split-major, blocks b, b + 8, ... on one
// XCD taking consecutive logical ids: with S = 8 each XCD owns one K range,
// whose A and B slices fit its L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "srt_internal.h"

namespace srt {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTP = 8;              // panels per tile (128 rows)
constexpr int kTN = 128;            // columns per tile
constexpr int kNTiles = kTN / 16;   // MFMA column tiles per tile
constexpr int kSG = 4;              // 16-column K groups per stage (64 k)
constexpr int kWaves = 8;
constexpr int kThreads = 64 * kWaves;
constexpr int kXcds = 8;
constexpr int kAW = kTP * kSG * 64 / kThreads;  // A words per thread per stage (4)
constexpr int kMaxSplit = 32;
// Workspace words (f32x4) of one (split, tile): every wave's 8 accumulators.
constexpr int64_t kTileWords = (int64_t)kWaves * 8 * 64;

template <bool kVecB>
__global__ void __launch_bounds__(kThreads)
    dense_tiles_kernel(int ndense, int G, int PT, int NT, int S, int SPS, const f32x4 *__restrict__ apack,
                       uint32_t a_bytes, const int32_t *__restrict__ panel_row, const float *__restrict__ B,
                       int64_t ldb, uint32_t b_bytes, float *__restrict__ C, int64_t ldc, f32x4 *__restrict__ ws,
                       int M, int N, int K) {
    __shared__ f32x4 la[2][kTP][kSG][64];
    __shared__ f32x4 lb[2][kSG][kNTiles][64];
    const int T = PT * NT, total = T * S;
    const int b = blockIdx.x;
    const int L = (total % kXcds == 0) ? (b % kXcds) * (total / kXcds) + b / kXcds : b;
    const int split = L / T, tile = L % T;
    const int pt = tile / NT, nt = tile % NT;
    const int stages = (G + kSG - 1) / kSG;
    const int s0 = split * SPS, s1 = min(stages, s0 + SPS);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int mw = wave >> 2, nw = wave & 3;

    // Buffer loads: an offset past the resource's extent reads 0, so groups
    // past G, rows past K, columns past N and panels past the list need no
    // branch (the stage after the last reads zeros, written but never used).
    const __amdgpu_buffer_rsrc_t brsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(B), (short)0, (int)b_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t arsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<f32x4 *>(apack), (short)0, (int)a_bytes, 0x00020000);

    // A staging: word e = tid + 512 r of the stage's [8 panels][4 groups][64 lanes]
    f32x4 ra[kAW], rb[4];
    auto load_a = [&](int s, bool ok) {
#pragma unroll
        for (int r = 0; r < kAW; ++r) {
            const int e = threadIdx.x + kThreads * r;
            const int p = e / (kSG * 64), q = (e >> 6) % kSG, l = e & 63;
            const int d = pt * kTP + p, g = s * kSG + q;
            const uint32_t off = (ok & (d < ndense) & (g < G))
                                     ? ((uint32_t)(d * G + g) * 64u + (uint32_t)l) * 16u
                                     : 0xffffffffu;
            ra[r] = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(arsrc, off, 0, 0));
        }
    };
    auto write_a = [&](int buf) {
#pragma unroll
        for (int r = 0; r < kAW; ++r) {
            const int e = threadIdx.x + kThreads * r;
            la[buf][e / (kSG * 64)][(e >> 6) % kSG][e & 63] = ra[r];
        }
    };
    // B staging: thread owns rows 16 q + 4 j + kr (j = 0..3) x columns
    // 4 c4 .. 4 c4 + 3 of the tile (512 threads = 16 (q, kr) x 32 c4)
    const int c4 = threadIdx.x & 31, qk = threadIdx.x >> 5;
    const int bq = qk >> 2, bkr = qk & 3;
    const int bn = nt * kTN + 4 * c4;
    auto load_b = [&](int s, bool ok) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = s * (16 * kSG) + 16 * bq + 4 * j + bkr;
            const bool rowok = ok & (k < K);
            if constexpr (kVecB) {
                const uint32_t at = ((uint32_t)k * (uint32_t)ldb + (uint32_t)bn) * 4u;
                const uint32_t off = (rowok & (bn < N)) ? at : 0xffffffffu;
                rb[j] = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(brsrc, off, 0, 0));
            } else {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const uint32_t at = ((uint32_t)k * (uint32_t)ldb + (uint32_t)(bn + c)) * 4u;
                    const uint32_t off = (rowok & (bn + c < N)) ? at : 0xffffffffu;
                    rb[j][c] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(brsrc, off, 0, 0));
                }
            }
        }
    };
    auto write_b = [&](int buf) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int col = 4 * c4 + c;
            lb[buf][bq][col >> 4][bkr * 16 + (col & 15)] = f32x4{rb[0][c], rb[1][c], rb[2][c], rb[3][c]};
        }
    };

    f32x4 acc[4][2];
#pragma unroll
    for (int p = 0; p < 4; ++p) acc[p][0] = acc[p][1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 a[2][4], bw[2][2];  // operands of two groups: one in use, one being read
    auto rd = [&](int buf, int q, int h) {
#pragma unroll
        for (int p = 0; p < 4; ++p) a[h][p] = la[buf][4 * mw + p][q][lane];
        bw[h][0] = lb[buf][q][2 * nw][lane];
        bw[h][1] = lb[buf][q][2 * nw + 1][lane];
    };
    auto mm = [&](int h) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    acc[p][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[h][p][j], bw[h][t][j], acc[p][t], 0, 0, 0);
    };

    load_a(s0, true);
    load_b(s0, true);
    write_a(0);
    write_b(0);
    __syncthreads();
    rd(0, 0, 0);
    for (int s = s0; s < s1; ++s) {
        const int buf = (s - s0) & 1;
        const bool more = s + 1 < s1;
        load_a(s + 1, more);
        load_b(s + 1, more);
        rd(buf, 1, 1);
        mm(0);
        rd(buf, 2, 0);
        mm(1);
        write_a(buf ^ 1);
        write_b(buf ^ 1);
        rd(buf, 3, 1);
        mm(0);
        // that order, interleaved: global loads first; each group's next
        // operands after its first two MFMAs; the other buffer's writes in
        // the middle of group 1
        __builtin_amdgcn_sched_group_barrier(0x020, 2 * kAW, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 30, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 14, 0);
        __builtin_amdgcn_sched_group_barrier(0x200, kAW + 4, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 16, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 30, 0);
        __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        rd(buf ^ 1, 0, 0);  // the next stage's group 0 (after the last stage: unused words)
        mm(1);
        __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 32, 0);
    }
    if (S == 1) {
        // D layout: lane l holds rows 4 (l / 16) .. + 3 of column l % 16
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int d = pt * kTP + 4 * mw + p;
            if (d >= ndense) continue;
            const int row0 = panel_row[d];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int n = nt * kTN + 16 * (2 * nw + t) + (lane & 15);
                if (n >= N) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = row0 + 4 * (lane >> 4) + r;
                    if (row < M) C[(int64_t)row * ldc + n] = acc[p][t][r];
                }
            }
        }
    } else {
        f32x4 *out = ws + ((int64_t)(split * T + tile) * kWaves + wave) * 8 * 64 + lane;
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int t = 0; t < 2; ++t) out[(p * 2 + t) * 64] = acc[p][t];
    }
}

// One thread per (tile, wave, accumulator, lane): the S partials added in
// split order, then written to C.
__global__ void __launch_bounds__(256)
    dense_tiles_reduce(int ndense, int PT, int NT, int S, const f32x4 *__restrict__ ws,
                       const int32_t *__restrict__ panel_row, float *__restrict__ C, int64_t ldc, int M, int N) {
    const int T = PT * NT;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)T * kTileWords) return;
    const int lane = (int)(i & 63), pt2 = (int)((i >> 6) & 7), wave = (int)((i >> 9) & 7);
    const int tile = (int)(i >> 12);
    const int pt = tile / NT, nt = tile % NT;
    const int mw = wave >> 2, nw = wave & 3, p = pt2 >> 1, t = pt2 & 1;
    const int d = pt * kTP + 4 * mw + p;
    const int n = nt * kTN + 16 * (2 * nw + t) + (lane & 15);
    if (d >= ndense || n >= N) return;
    const int64_t stride = (int64_t)T * kTileWords;
    f32x4 v = ws[i];
    int s = 1;
    for (; s + 8 <= S; s += 8) {  // eight loads in flight, added in split order
        f32x4 w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = ws[i + (s + u) * stride];
#pragma unroll
        for (int u = 0; u < 8; ++u) v += w[u];
    }
    for (; s < S; ++s) v += ws[i + s * stride];
    const int row0 = panel_row[d];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = row0 + 4 * (lane >> 4) + r;
        if (row < M) C[(int64_t)row * ldc + n] = v[r];
    }
}

// Tiles, K split and stages per split of a launch.
struct DenseGrid {
    int PT, NT, S, SPS;
};

DenseGrid dense_grid(int ndense, int G, int N, size_t ws_words) {
    DenseGrid g;
    g.PT = (ndense + kTP - 1) / kTP;
    g.NT = (N + kTN - 1) / kTN;
    const int T = g.PT * g.NT;
    const int stages = std::max(1, (G + kSG - 1) / kSG);
    int S = std::max(1, std::min(kMaxSplit, 256 / std::max(T, 1)));
    S = std::min(S, stages);
    while (S > 1 && (size_t)S * T * kTileWords > ws_words) --S;  // what the workspace holds
    g.SPS = (stages + S - 1) / S;
    g.S = (stages + g.SPS - 1) / g.SPS;  // no empty K range
    return g;
}

}  // namespace

size_t dense_panels_workspace_bytes(int ndense, int G, int n) {
    if (ndense <= 0 || n <= 0) return 0;
    const DenseGrid g = dense_grid(ndense, G, n, ~(size_t)0);
    return g.S > 1 ? (size_t)g.S * g.PT * g.NT * kTileWords * sizeof(f32x4) : 0;
}

hipError_t launch_dense_panels(int ndense, int G, const void *apack, const int32_t *panel_row, const float *B,
                               int64_t ldb, float *C, int64_t ldc, int M, int N, int K, void *ws, size_t ws_bytes,
                               hipStream_t s) {
    if (ndense <= 0 || N <= 0) return hipSuccess;
    // byte offsets are 32-bit (buffer loads); the extent must stay below the
    // 0xffffffff offset that marks a zero read
    const int64_t a_bytes = (int64_t)ndense * G * 64 * 16, b_bytes = (int64_t)K * ldb * 4;
    if (a_bytes >= 0xffffffffLL || b_bytes >= 0xffffffffLL) return hipErrorInvalidValue;
    const DenseGrid g = dense_grid(ndense, G, N, ws ? ws_bytes / sizeof(f32x4) : 0);
    const int T = g.PT * g.NT;
    const bool vec = N % 4 == 0 && ldb % 4 == 0 && (uintptr_t)B % 16 == 0;
    if (vec)
        hipLaunchKernelGGL(dense_tiles_kernel<true>, dim3((unsigned)(T * g.S)), dim3(kThreads), 0, s, ndense, G,
                           g.PT, g.NT, g.S, g.SPS, static_cast<const f32x4 *>(apack), (uint32_t)a_bytes, panel_row,
                           B, ldb, (uint32_t)b_bytes, C, ldc, static_cast<f32x4 *>(ws), M, N, K);
    else
        hipLaunchKernelGGL(dense_tiles_kernel<false>, dim3((unsigned)(T * g.S)), dim3(kThreads), 0, s, ndense, G,
                           g.PT, g.NT, g.S, g.SPS, static_cast<const f32x4 *>(apack), (uint32_t)a_bytes, panel_row,
                           B, ldb, (uint32_t)b_bytes, C, ldc, static_cast<f32x4 *>(ws), M, N, K);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || g.S == 1) return e;
    const int64_t threads = (int64_t)T * kTileWords;
    hipLaunchKernelGGL(dense_tiles_reduce, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, ndense, g.PT,
                       g.NT, g.S, static_cast<const f32x4 *>(ws), panel_row, C, ldc, M, N);
    return hipGetLastError();
}

}  // namespace srt
