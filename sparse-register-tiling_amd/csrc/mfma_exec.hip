// Register tiles on matrix cores: the M_r = 4 full-pattern packing executed
// with v_mfma_f32_4x4x1_16b_f32 outer products.
//
// Under the mapping whose only kernel pattern is 0b1111 (every non-empty
// 4-row column pattern padded to the full column, the reference's SOPOrig
// with merge pattern 0b1111 and min_merge = 1, sbench/SOP/sop_hand_configs.py:
// 7-65), every packed column of a panel carries its four values (zero where
// the panel row has no entry), in ascending column order K tile by K tile.
// The register tile's nano-kernel for that pattern is four fused multiply-adds
// per lane column; here one MFMA instruction does it for 64 columns at once:
// v_mfma_f32_4x4x1_16b_f32 runs 16 independent 4x4x1 outer products, block b
// (lanes 4b .. 4b+3) on C columns n0 + 4b .. n0 + 4b + 3:
//   A: lane 4b + i holds row i of the block's column (4 values),
//   B: lane 4b + j holds B[k][n0 + 4b + j], i.e. lane l holds B[k][n0 + l],
//   D: lane 4b + j, register i holds C[row i][n0 + 4b + j].
// With CBSZ = 4 / ABID = q every block takes block q's A, so one 64-value
// load (16 packed columns x 4 rows, lane l = value l) feeds 16 MFMAs.  Per
// output element each instruction is one multiply-add with a single rounding
// (the same bits as fmaf, checked on gfx950), applied in packed order, so the
// result is the register tile's sequential fmaf chain over the same packed
// data, bit for bit.
//
// B: a workgroup owns one 64-column slice of B and C and first copies rows
// 0..K-1 of the slice into LDS ([K][64] fp32, K <= 640); its 16 waves then
// walk the panels of its range (wave w: panels lo + w, lo + w + 16, ...), a
// lane reading its B entry of a packed column with one ds_read_b32 (a wave
// reads one contiguous 256-byte row: no bank conflicts).  Panel ranges are
// contiguous and balanced by packed columns; the host sizes them so that
// ranges x slices is one round of the chip's CUs (one workgroup per CU: the
// tile takes the LDS).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "exec_kernel.h"
#include "srt_internal.h"

namespace srt {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 16;
constexpr int kThreads = 64 * kWaves;
constexpr int kCols = 64;   // C columns per workgroup (16 blocks x 4)
constexpr int kBlock = 16;  // packed columns per A load (one per MFMA block)

// 16 MFMAs of one A word: column q of the word is block q's A (ABID = q).
template <int Q>
__device__ __forceinline__ void mfma_block(f32x4 &acc, float a, const float (&b)[kBlock]) {
    if constexpr (Q < kBlock) {
        acc = __builtin_amdgcn_mfma_f32_4x4x1f32(a, b[Q], acc, 4, Q, 0);
        mfma_block<Q + 1>(acc, a, b);
    }
}

template <int Q>
__device__ __forceinline__ void mfma_tail(f32x4 &acc, float a, const float (&b)[kBlock], int rem) {
    if constexpr (Q < kBlock) {
        if (Q < rem) {
            acc = __builtin_amdgcn_mfma_f32_4x4x1f32(a, b[Q], acc, 4, Q, 0);
            mfma_tail<Q + 1>(acc, a, b, rem);
        }
    }
}

__global__ void __launch_bounds__(kThreads)
    mfma4_exec_kernel(const int64_t *__restrict__ col_ptr, const int32_t *__restrict__ cols,
                      const float *__restrict__ vals, const int32_t *__restrict__ ranges, int S,
                      const float *__restrict__ B, uint32_t b_bytes, uint32_t ldb, float *__restrict__ C,
                      int64_t ldc, int M, int N, int K) {
    extern __shared__ float bs[];  // [K][64] B slice
    // Logical workgroup (range r, slice s), range-major: the blocks one XCD
    // receives share their panel ranges' packed data in that XCD's L2.
    const int L = xcd_swizzle(blockIdx.x, gridDim.x);
    const int r = L / S, s = L - r * S;
    const int n0 = s * kCols;

    // Copy B[0..K)[n0 .. n0 + 64) into LDS, 16-byte chunks, eight loads in
    // flight per lane.  Chunks past column N read the next row (or 0 past the
    // end of B): only lanes whose C column is past N use them.
    {
        const __amdgpu_buffer_rsrc_t rsrc =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(B), (short)0, (int)b_bytes, 0x00020000);
        constexpr int BATCH = 8;
        const int total = K * (kCols / 4);
        for (int i0 = threadIdx.x; i0 < total; i0 += BATCH * kThreads) {
            float4 v[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int i = i0 + u * kThreads;
                if (i < total) {
                    const int k = i >> 4, c = i & 15;
                    v[u] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(
                                                          rsrc, (uint32_t)k * ldb * 4u + (uint32_t)(n0 + 4 * c) * 4u,
                                                          0, 0));
                }
            }
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int i = i0 + u * kThreads;
                if (i < total) *reinterpret_cast<float4 *>(bs + (i >> 4) * kCols + 4 * (i & 15)) = v[u];
            }
        }
    }
    __syncthreads();

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int n = n0 + lane;
    const bool active = n < N;
    const float *brow = bs + lane;
    const int lo = ranges[r], hi = ranges[r + 1];
    for (int p = lo + wave; p < hi; p += kWaves) {
        const int64_t c0 = col_ptr[p], c1 = col_ptr[p + 1];
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        // A block = 16 packed columns: one A word (lane l = value l) and the
        // 16 column indices (lane l holds index l % 16), both vector loads so
        // the next block's loads stay in flight (in-order vector counter)
        // while this block's LDS reads and MFMAs run.  Words and indices past
        // the panel's last column belong to the next panel or to the arrays'
        // zero padding: loaded, never used.
        float a_next = vals[4 * c0 + lane];
        int cv_next = cols[c0 + (lane & (kBlock - 1))];
        for (int64_t c = c0; c < c1; c += kBlock) {
            const float a = a_next;
            const int cv = cv_next;
            if (c + kBlock < c1) {
                a_next = vals[4 * (c + kBlock) + lane];
                cv_next = cols[c + kBlock + (lane & (kBlock - 1))];
            }
            float b[kBlock];
#pragma unroll
            for (int q = 0; q < kBlock; ++q) b[q] = brow[__builtin_amdgcn_readlane(cv, q) * kCols];
            const int64_t rem = c1 - c;
            if (rem >= kBlock) {
                mfma_block<0>(acc, a, b);
            } else {
                mfma_tail<0>(acc, a, b, (int)rem);
            }
        }
        if (active) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = 4 * p + i;
                if (row < M) __builtin_nontemporal_store(acc[i], C + (int64_t)row * ldc + n);
            }
        }
    }
}

}  // namespace

int mfma4_max_k() { return (int)(lds_limit() / (kCols * sizeof(float))); }

hipError_t launch_mfma4_exec(const int64_t *col_ptr, const int32_t *cols, const float *vals,
                             const int32_t *ranges, int R, const float *B, int64_t ldb, float *C, int64_t ldc,
                             int M, int N, int K, hipStream_t s) {
    if (N <= 0 || M <= 0 || R <= 0) return hipSuccess;
    const int S = (N + kCols - 1) / kCols;
    const int64_t b_bytes = (int64_t)K * ldb * 4;
    if (K > mfma4_max_k() || b_bytes >= 0xffffffffLL || (int64_t)R * S > 0x7fffffff) return hipErrorInvalidValue;
    // once; a function-local static is initialised exactly once even when
    // several threads launch at the same time
    static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void *>(mfma4_exec_kernel),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit());
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL(mfma4_exec_kernel, dim3((unsigned)(R * S)), dim3(kThreads),
                       (size_t)K * kCols * sizeof(float), s, col_ptr, cols, vals, ranges, S, B,
                       (uint32_t)b_bytes, (uint32_t)ldb, C, ldc, M, N, K);
    return hipGetLastError();
}

}  // namespace srt
