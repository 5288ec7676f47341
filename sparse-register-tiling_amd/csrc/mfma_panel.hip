// Precision form: BF16 and FP8-e4m3 SpMM on MFMA over 16-row panels.
//
// A is cut into 16-row panels and 32-column K blocks, each stored dense (zeros
// filled: the panel is padded to a dense 16-row block, the extreme of the
// reference's ILP padding, sbench/ilp_pad/nano_solver.py:76-116, where each
// column pattern is padded up to all rows).  This form is meant for panels
// dense enough for matrix cores (the 50-70 % sparse precision shapes hold a
// nonzero in practically every block); sparser matrices take the register-
// tiled executors.  Block (p, kb) sits at slot p * KB + kb, laid out exactly as the
// v_mfma_f32_16x16x32_{bf16,fp8_fp8} A operand wants it: lane l holds row
// l % 16, columns 8 * (l / 16) .. + 8, so one lane reads one 16-byte (BF16) or
// 8-byte (FP8) word per block.
//
// A workgroup owns WP panels x NT columns of C and walks the K blocks of its
// K range: the B tile (32 rows x NT columns) is copied from row-major B into
// LDS transposed to [column][k] (the MFMA B operand wants eight consecutive k
// of one column per lane), then each wave runs NT / 16 MFMAs for its panel if
// the panel stores that block.  Accumulation is FP32.  FP8: A rows carry FP32
// scales s_i = max_k |A_ik| / 448 and B columns t_j = max_k |B_kj| / 448, both
// computed in FP64; C_ij = s_i * t_j * sum_k q(A)_ik q(B)_kj.
//
// A K range split over KS workgroups writes partial tiles to a workspace that
// a second kernel adds in split order (deterministic).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "srt_internal.h"

namespace srt {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kPanelRows = 16;
constexpr int kBlockK = 32;
constexpr int kBlockElems = kPanelRows * kBlockK;  // 512

// ------------------------------------------------------------ conversions
// float -> bfloat16, round to nearest even (NaN stays a quiet NaN).
__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// float -> FP8 e4m3 (OCP "fn": bias 7, no infinities, max 448), round to
// nearest even, saturating at +-448; NaN -> 0x7f.
__device__ __forceinline__ uint8_t to_e4m3(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    const uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    const uint32_t a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return 0x7f;
    if (a >= 0x43e00000u) return sign | 0x7e;  // >= 448
    const int e = (int)(a >> 23) - 127;
    if (e < -6) {
        // subnormal target: multiples of 2^-9; a * 512 is exact, rint is RNE
        const float q = __builtin_rintf(__builtin_bit_cast(float, a) * 512.0f);
        return sign | (uint8_t)q;  // q in [0, 8]; 8 is the smallest normal
    }
    uint32_t m = a & 0x7fffffu;
    uint32_t r = m >> 20;
    const uint32_t rem = m & 0xfffffu;
    uint32_t ex = (uint32_t)(e + 7);
    if (rem > 0x80000u || (rem == 0x80000u && (r & 1u))) {
        if (++r == 8u) {
            r = 0;
            ++ex;
        }
    }
    if (ex > 15u || (ex == 15u && r == 7u)) return sign | 0x7e;  // rounded past 448
    return sign | (uint8_t)((ex << 3) | r);
}

// ------------------------------------------------------------ packing
// One wave per CSR row: mark the K blocks holding a nonzero.
template <typename I>
__global__ void blk_mark_kernel(int m, int k, int KB, const I *__restrict__ Lp,
                                const I *__restrict__ Li, uint8_t *__restrict__ present,
                                int *__restrict__ bad) {
    const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= m) return;
    const int lane = threadIdx.x & 63;
    const int p = row / kPanelRows;
    for (int64_t q = (int64_t)Lp[row] + lane; q < (int64_t)Lp[row + 1]; q += 64) {
        const int64_t c = (int64_t)Li[q];
        if (c < 0 || c >= k) {
            atomicOr(bad, 1);
            continue;
        }
        present[(int64_t)p * KB + c / kBlockK] = 1;
    }
}

// One thread per panel: number of stored blocks.
__global__ void blk_count_kernel(int P, int KB, const uint8_t *__restrict__ present,
                                 int64_t *__restrict__ nblk) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int64_t n = 0;
    for (int b = 0; b < KB; ++b) n += present[(int64_t)p * KB + b];
    nblk[p] = n;
}

// One thread per panel: only blocks holding a nonzero are stored, the
// panel's from blk_ptr[p] on in K order; slot[p * KB + kb] is the stored
// index of block (p, kb) or -1.  With every block present the slots are
// p * KB + kb and the executors address blocks without the table.
__global__ void blk_slot_kernel(int P, int KB, const uint8_t *__restrict__ present,
                                const int64_t *__restrict__ blk_ptr, int32_t *__restrict__ slot,
                                int32_t *__restrict__ blk_kb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int64_t at = blk_ptr[p];
    for (int b = 0; b < KB; ++b) {
        const int64_t s = (int64_t)p * KB + b;
        if (present[s]) {
            slot[s] = (int32_t)at;
            blk_kb[at] = b;
            ++at;
        } else {
            slot[s] = -1;
        }
    }
}

// One wave per row: FP8 row scale s = max|A_i| / 448 computed in FP64.
template <typename I>
__global__ void row_scale_kernel(int m, const I *__restrict__ Lp, const float *__restrict__ Lx,
                                 float *__restrict__ scale) {
    const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= m) return;
    const int lane = threadIdx.x & 63;
    double mx = 0.0;
    for (int64_t q = (int64_t)Lp[row] + lane; q < (int64_t)Lp[row + 1]; q += 64)
        mx = fmax(mx, fabs((double)Lx[q]));
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    if (lane == 0) scale[row] = mx > 0.0 ? (float)(mx / 448.0) : 1.0f;
}

__device__ __forceinline__ int64_t operand_offset(int64_t slot, int r, int kk) {
    return slot * kBlockElems + (int64_t)(((kk >> 3) * kPanelRows + r) * 8 + (kk & 7));
}

// One wave per row: every nonzero into its block, converted.
template <typename I, bool FP8>
__global__ void blk_fill_kernel(int m, int k, int KB, const I *__restrict__ Lp,
                                const I *__restrict__ Li, const float *__restrict__ Lx,
                                const int32_t *__restrict__ slot, const float *__restrict__ rscale,
                                void *__restrict__ blocks) {
    const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= m) return;
    const int lane = threadIdx.x & 63;
    const int p = row / kPanelRows, r = row % kPanelRows;
    const double s = FP8 ? (double)rscale[row] : 1.0;
    for (int64_t q = (int64_t)Lp[row] + lane; q < (int64_t)Lp[row + 1]; q += 64) {
        const int64_t c = (int64_t)Li[q];
        if (c < 0 || c >= k) continue;
        const int32_t sl = slot[(int64_t)p * KB + c / kBlockK];
        const int64_t off = operand_offset(sl, r, (int)(c % kBlockK));
        if constexpr (FP8) {
            static_cast<uint8_t *>(blocks)[off] = to_e4m3((float)((double)Lx[q] / s));
        } else {
            static_cast<uint16_t *>(blocks)[off] = to_bf16(Lx[q]);
        }
    }
}

// ------------------------------------------------------------ B operand prep
// One block of 256 threads per 256 columns: FP8 column scales in FP64.
__global__ void col_scale_kernel(int64_t rows, int cols, const float *__restrict__ src, int64_t ld,
                                 float *__restrict__ scale) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    double mx = 0.0;
    for (int64_t r = 0; r < rows; ++r) mx = fmax(mx, fabs((double)src[r * ld + c]));
    scale[c] = mx > 0.0 ? (float)(mx / 448.0) : 1.0f;
}

// fp32 B (K x N, row-major) -> Bt (N x ldk, k contiguous) in BF16 or e4m3,
// k in [K, ldk) zero: the MFMA B operand wants eight consecutive k of one
// column per lane.  32 x 32 tiles go through LDS so that both the reads of B
// and the writes of Bt are coalesced.
template <bool FP8>
__global__ void b_transpose_kernel(int K, int N, const float *__restrict__ src, int64_t ld_src,
                                   const float *__restrict__ cscale, typename std::conditional<FP8, uint8_t, uint16_t>::type *__restrict__ dst,
                                   int64_t ldk) {
    __shared__ float t[32][33];
    const int k0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int i = ty; i < 32; i += 8) {
        const int k = k0 + i, n = n0 + tx;
        t[i][tx] = (k < K && n < N) ? src[(int64_t)k * ld_src + n] : 0.0f;
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int n = n0 + i, k = k0 + tx;
        if (n < N && k < ldk) {
            const float v = t[tx][i];
            if constexpr (FP8)
                dst[(int64_t)n * ldk + k] = to_e4m3((float)((double)v / (double)cscale[n]));
            else
                dst[(int64_t)n * ldk + k] = to_bf16(v);
        }
    }
}

// ------------------------------------------------------------ executor
template <bool FP8>
struct Elt {
    typedef uint16_t T;
    typedef uint4 AOp;                   // A operand of one lane: 8 BF16
    static constexpr int kPerChunk = 8;  // elements per 16-byte chunk
    static constexpr int kLdsK = 40;     // LDS row: 32 k + 8 padding (80 bytes)
};
template <>
struct Elt<true> {
    typedef uint8_t T;
    typedef long AOp;  // 8 e4m3
    static constexpr int kPerChunk = 16;
    static constexpr int kLdsK = 48;  // 32 k + 16 padding (48 bytes)
};

// Workgroup: WP waves x PW panels each (16 * WP * PW rows) x NT columns.
// B arrives transposed (Bt: N x ldb, ldb >= KB * 32, k past K zero), so a B
// tile is NT rows of 32 contiguous k: a straight 16-byte copy into LDS, and
// each MFMA B operand is one 16-byte (8-byte) LDS read.  Software pipeline,
// one barrier per kStageBlocks K blocks: the B chunks and A operands of later
// blocks are loaded into registers while the MFMAs of the current stage run
// from the other LDS buffer (the ring below keeps several blocks in flight).
constexpr int kStageBlocks = 2;
template <bool FP8, int NT>
using TileBuf = typename Elt<FP8>::T[2][kStageBlocks * NT * Elt<FP8>::kLdsK];

// One output tile: panels (pg * WP + wave) * PW .. + PW, columns nt * NT ..,
// K range z of ks.  Every wave of the workgroup calls it (barriers inside).
// CMP: blocks are stored compactly and found through slot[p * KB + kb]
// (-1 = no nonzero: the A operand is zero); otherwise block (p, kb) sits at
// p * KB + kb.
template <bool FP8, int NT, int WP, int PW, bool CMP>
__device__ __forceinline__ void mfma_tile(TileBuf<FP8, NT> &tile, int M, int N, int KB, int P,
                                          const typename Elt<FP8>::T *__restrict__ Bt, int64_t ldb,
                                          const void *__restrict__ blocks, const int32_t *__restrict__ slot,
                                          const float *__restrict__ rscale,
                                          const float *__restrict__ cscale, float *__restrict__ C,
                                          int64_t ldc, int ks, float *__restrict__ work, int vec_ok, int pg,
                                          int nt, int z) {
    typedef typename Elt<FP8>::T T;
    typedef typename Elt<FP8>::AOp AOp;
    constexpr int PER = Elt<FP8>::kPerChunk;
    constexpr int LDK = Elt<FP8>::kLdsK;
    constexpr int CPR = kBlockK / PER;  // 16-byte chunks per tile row
    constexpr int CHUNKS = NT * CPR;
    constexpr int NTHR = 64 * WP;
    constexpr int CPT = (CHUNKS + NTHR - 1) / NTHR;
    constexpr int TN = NT / 16;

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int p0 = (pg * WP + wave) * PW;
    const int n0 = nt * NT;
    const int kb0 = (int)((int64_t)z * KB / ks), kb1 = (int)((int64_t)(z + 1) * KB / ks);
    const int col = lane & 15, kg = lane >> 4;

    f32x4 acc[PW][TN];
#pragma unroll
    for (int i = 0; i < PW; ++i)
#pragma unroll
        for (int t = 0; t < TN; ++t) acc[i][t] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Register ring of DEPTH K blocks in flight (B chunks and A operands):
    // HBM latency spans several blocks' worth of MFMAs.  The loop is unrolled
    // by DEPTH so every ring index is a compile-time constant.
    constexpr int DEPTH = 4;
    uint4 breg[DEPTH][CPT];
    AOp areg[DEPTH][PW];
    // CMP: slots ride one ring ahead of the A operands (sreg[j] holds the
    // slots of the blocks areg[j] loads next), so an A load never waits on a
    // slot issued just before it; aok[j] says whether areg[j]'s block exists.
    // Slots are read with buffer loads (vector memory, completed in order
    // with the B and A loads).
    int sreg[CMP ? DEPTH : 1][PW];
    bool aok[CMP ? DEPTH : 1][PW];
    const __amdgpu_buffer_rsrc_t srsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int32_t *>(slot), (short)0, CMP ? (int)((int64_t)P * KB * 4) : 0, 0x00020000);
    auto load_slot = [&](int (&dst)[PW], int kb) {
#pragma unroll
        for (int i = 0; i < PW; ++i) {
            const int p = p0 + i < P ? p0 + i : P - 1;
            dst[i] = __builtin_amdgcn_raw_buffer_load_b32(srsrc, (int)(((int64_t)p * KB + kb) * 4), 0, 0);
        }
    };
    // Branch-free (the host guarantees 16-byte aligned rows): a lane past
    // column N loads row 0 (valid) and its chunk is zeroed when stored, so
    // no load sits in a branch or feeds a select (either would force a wait
    // right after it and defeat the ring).
    (void)vec_ok;
    auto load_b = [&](uint4 (&dst)[CPT], int kb) {
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const int ch = threadIdx.x + c * NTHR;
            const int n = ch / CPR, part = ch % CPR;
            const bool in = ch < CHUNKS && n0 + n < N;
            const T *src = Bt + (int64_t)(in ? n0 + n : 0) * ldb + (int64_t)kb * kBlockK + part * PER;
            dst[c] = *reinterpret_cast<const uint4 *>(src);
        }
    };
    // Loads are unconditional (panel and block indices clamped to valid
    // ones; results past the panel count or the K range are never stored),
    // so every path issues the same loads and the waits stay partial.
    auto load_a = [&](AOp (&dst)[PW], int kb) {
#pragma unroll
        for (int i = 0; i < PW; ++i) {
            const int p = p0 + i < P ? p0 + i : P - 1;
            dst[i] = reinterpret_cast<const AOp *>(blocks)[((int64_t)p * KB + kb) * 64 + lane];
        }
    };
    // CMP: the A operands of the blocks whose slots are in sl; an absent
    // block loads stored block 0 (the plan keeps at least one block's bytes)
    // and is zeroed where it is used.
    auto load_a_slots = [&](AOp (&dst)[PW], bool (&ok)[PW], const int (&sl)[PW]) {
#pragma unroll
        for (int i = 0; i < PW; ++i) {
            ok[i] = sl[i] >= 0;
            dst[i] = reinterpret_cast<const AOp *>(blocks)[(int64_t)(ok[i] ? sl[i] : 0) * 64 + lane];
        }
    };
    const bool any = kb0 < kb1;  // uniform; an empty K range still stores zeros
    if constexpr (CMP) {
        // slots of the first 2 DEPTH blocks, A operands of the first DEPTH
#pragma unroll
        for (int j = 0; j < DEPTH && any; ++j) {
            const int kb = kb0 + j < kb1 ? kb0 + j : kb1 - 1;
            load_b(breg[j], kb);
            load_slot(sreg[j], kb);
        }
#pragma unroll
        for (int j = 0; j < DEPTH && any; ++j) {
            load_a_slots(areg[j], aok[j], sreg[j]);
            const int kn = kb0 + DEPTH + j < kb1 ? kb0 + DEPTH + j : kb1 - 1;
            load_slot(sreg[j], kn);
        }
    } else {
#pragma unroll
        for (int j = 0; j < DEPTH && any; ++j) {
            const int kb = kb0 + j < kb1 ? kb0 + j : kb1 - 1;
            load_b(breg[j], kb);
            load_a(areg[j], kb);
        }
    }
    int buf = 0;
    static_assert(DEPTH % kStageBlocks == 0, "the ring holds whole stages");
    constexpr int SUB = NT * LDK;  // one K block's tile inside a stage buffer
    for (int base = kb0; base < kb1; base += DEPTH) {
#pragma unroll
        for (int j0 = 0; j0 < DEPTH; j0 += kStageBlocks) {
#pragma unroll
            for (int h = 0; h < kStageBlocks; ++h) {
#pragma unroll
                for (int c = 0; c < CPT; ++c) {
                    const int ch = threadIdx.x + c * NTHR;
                    if (ch < CHUNKS) {
                        const bool in = n0 + ch / CPR < N;
                        *reinterpret_cast<uint4 *>(&tile[buf][h * SUB + (ch / CPR) * LDK + (ch % CPR) * PER]) =
                            in ? breg[j0 + h][c] : uint4{0u, 0u, 0u, 0u};
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < kStageBlocks; ++h) {
                const int j = j0 + h;
                const int kb = base + j;
                const bool run = kb < kb1;  // uniform; every wave takes every barrier
                AOp a[PW];
#pragma unroll
                for (int i = 0; i < PW; ++i) {
                    a[i] = areg[j][i];
                    if constexpr (CMP) a[i] = aok[j][i] ? a[i] : AOp{};
                }
                {
                    const int kn = kb + DEPTH < kb1 ? kb + DEPTH : kb1 - 1;
                    load_b(breg[j], kn);
                    if constexpr (CMP) {
                        load_a_slots(areg[j], aok[j], sreg[j]);
                        const int kn2 = kb + 2 * DEPTH < kb1 ? kb + 2 * DEPTH : kb1 - 1;
                        load_slot(sreg[j], kn2);
                    } else {
                        load_a(areg[j], kn);
                    }
                }
#pragma unroll
                for (int t = 0; t < TN && run; ++t) {
                    const T *bp = &tile[buf][h * SUB + (t * 16 + col) * LDK + kg * 8];
                    if constexpr (FP8) {
                        const long bw = *reinterpret_cast<const long *>(bp);
#pragma unroll
                        for (int i = 0; i < PW; ++i)
                            acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a[i], bw, acc[i][t], 0, 0, 0);
                    } else {
                        const bf16x8 bw = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(bp));
#pragma unroll
                        for (int i = 0; i < PW; ++i)
                            acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[i]),
                                                                                bw, acc[i][t], 0, 0, 0);
                    }
                }
            }
            buf ^= 1;
        }
    }
    // D layout: lane l holds rows 4 * (l / 16) .. + 4 of column l % 16.
    const int rg = lane >> 4;
#pragma unroll
    for (int i = 0; i < PW; ++i) {
        const int p = p0 + i;
        if (p >= P) continue;
#pragma unroll
        for (int t = 0; t < TN; ++t) {
            const int n = n0 + t * 16 + col;
            if (n >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = p * kPanelRows + rg * 4 + r;
                if (row >= M) continue;
                float v = acc[i][t][r];
                if (ks > 1) {
                    work[((int64_t)z * M + row) * N + n] = v;
                } else {
                    if constexpr (FP8) v = v * rscale[row] * cscale[n];
                    C[(int64_t)row * ldc + n] = v;
                }
            }
        }
    }
}

template <bool FP8, int NT, int WP, int PW, bool CMP>
__global__ void __launch_bounds__(64 * WP)
    mfma_exec_kernel(int M, int N, int KB, int P, const typename Elt<FP8>::T *__restrict__ Bt, int64_t ldb,
                     const void *__restrict__ blocks, const int32_t *__restrict__ slot,
                     const float *__restrict__ rscale, const float *__restrict__ cscale, float *__restrict__ C,
                     int64_t ldc, int ks, float *__restrict__ work, int vec_ok) {
    __shared__ __attribute__((aligned(16))) TileBuf<FP8, NT> tile;
    mfma_tile<FP8, NT, WP, PW, CMP>(tile, M, N, KB, P, Bt, ldb, blocks, slot, rscale, cscale, C, ldc, ks, work,
                                    vec_ok, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Grouped, persistent: the tiles of every problem in the table, problem q
// owning global tiles [tile_begin[q], tile_begin[q + 1]); each workgroup
// takes tiles blockIdx.x, blockIdx.x + gridDim.x, ...  One launch runs a
// whole batch of independent SpMMs (requests, layers).  The problem count and
// tile total come from device memory (hdr[0], hdr[1]) next to the table.
//
// The tables are read through the constant address space: pointers loaded
// from there are known to address global memory (from a generic table every
// access through them would be a flat access).
typedef const __attribute__((address_space(4))) MfmaProblem *ConstMfmaProblems;
typedef const __attribute__((address_space(4))) BPrepProblem *ConstBPrepProblems;

template <bool FP8, int NT, int WP, int PW>
__global__ void __launch_bounds__(64 * WP)
    mfma_grouped_kernel(ConstMfmaProblems probs, const int *__restrict__ hdr) {
    __shared__ __attribute__((aligned(16))) TileBuf<FP8, NT> tile;
    const int nprob = hdr[0], total_tiles = hdr[1];  // read from memory: a captured graph replays any batch
    int q = 0;
    for (int t = blockIdx.x; t < total_tiles; t += gridDim.x) {
        while (q + 1 < nprob && probs[q + 1].tile_begin <= t) ++q;  // tiles ascend per workgroup
        const int local = t - probs[q].tile_begin;
        const int N = probs[q].N;
        const int ntiles = (N + NT - 1) / NT;
        const auto *Bt = static_cast<const typename Elt<FP8>::T *>(probs[q].Bt);
        if (probs[q].slot)  // uniform: compactly stored blocks
            mfma_tile<FP8, NT, WP, PW, true>(tile, probs[q].M, N, probs[q].KB, probs[q].P, Bt, probs[q].ldb,
                                             probs[q].blocks, probs[q].slot, probs[q].rscale, probs[q].cscale,
                                             probs[q].C, probs[q].ldc, 1, nullptr, probs[q].vec_ok, local / ntiles,
                                             local % ntiles, 0);
        else
            mfma_tile<FP8, NT, WP, PW, false>(tile, probs[q].M, N, probs[q].KB, probs[q].P, Bt, probs[q].ldb,
                                              probs[q].blocks, nullptr, probs[q].rscale, probs[q].cscale,
                                              probs[q].C, probs[q].ldc, 1, nullptr, probs[q].vec_ok,
                                              local / ntiles, local % ntiles, 0);
        __syncthreads();  // the next tile rewrites the LDS buffers
    }
}


// Adds the ks partial tiles in split order (and applies the FP8 scales).
template <bool FP8>
__global__ void split_reduce_kernel(int M, int N, int ks, const float *__restrict__ work,
                                    const float *__restrict__ rscale, const float *__restrict__ cscale,
                                    float *__restrict__ C, int64_t ldc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)M * N) return;
    const int64_t row = i / N, n = i % N;
    float v = work[i];
    for (int z = 1; z < ks; ++z) v = v + work[(int64_t)z * M * N + i];
    if constexpr (FP8) v = v * rscale[row] * cscale[n];
    C[row * ldc + n] = v;
}

}  // namespace

// ------------------------------------------------------------ launchers
hipError_t launch_mfma_pack(int m, int k, const void *Lp, const void *Li, int index_bytes,
                            const float *Lx, bool fp8, int P, int KB, uint8_t *present,
                            int64_t *nblk, int *bad, hipStream_t s) {
    const dim3 grid((m + 3) / 4), block(256);
    if (m <= 0) return hipSuccess;
    if (index_bytes == 4)
        hipLaunchKernelGGL(blk_mark_kernel<int32_t>, grid, block, 0, s, m, k, KB,
                           (const int32_t *)Lp, (const int32_t *)Li, present, bad);
    else
        hipLaunchKernelGGL(blk_mark_kernel<int64_t>, grid, block, 0, s, m, k, KB,
                           (const int64_t *)Lp, (const int64_t *)Li, present, bad);
    hipLaunchKernelGGL(blk_count_kernel, dim3((P + 255) / 256), dim3(256), 0, s, P, KB, present, nblk);
    (void)Lx;
    (void)fp8;
    return hipGetLastError();
}

hipError_t launch_mfma_slots(int P, int KB, const uint8_t *present, const int64_t *blk_ptr,
                             int32_t *slot, int32_t *blk_kb, hipStream_t s) {
    if (P <= 0) return hipSuccess;
    hipLaunchKernelGGL(blk_slot_kernel, dim3((P + 255) / 256), dim3(256), 0, s, P, KB, present, blk_ptr,
                       slot, blk_kb);
    return hipGetLastError();
}

hipError_t launch_mfma_fill(int m, int k, int KB, const void *Lp, const void *Li, int index_bytes,
                            const float *Lx, bool fp8, const int32_t *slot, float *rscale,
                            void *blocks, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    const dim3 grid((m + 3) / 4), block(256);
    if (index_bytes == 4) {
        if (fp8) {
            hipLaunchKernelGGL(row_scale_kernel<int32_t>, grid, block, 0, s, m, (const int32_t *)Lp, Lx, rscale);
            hipLaunchKernelGGL((blk_fill_kernel<int32_t, true>), grid, block, 0, s, m, k, KB,
                               (const int32_t *)Lp, (const int32_t *)Li, Lx, slot, rscale, blocks);
        } else {
            hipLaunchKernelGGL((blk_fill_kernel<int32_t, false>), grid, block, 0, s, m, k, KB,
                               (const int32_t *)Lp, (const int32_t *)Li, Lx, slot, rscale, blocks);
        }
    } else {
        if (fp8) {
            hipLaunchKernelGGL(row_scale_kernel<int64_t>, grid, block, 0, s, m, (const int64_t *)Lp, Lx, rscale);
            hipLaunchKernelGGL((blk_fill_kernel<int64_t, true>), grid, block, 0, s, m, k, KB,
                               (const int64_t *)Lp, (const int64_t *)Li, Lx, slot, rscale, blocks);
        } else {
            hipLaunchKernelGGL((blk_fill_kernel<int64_t, false>), grid, block, 0, s, m, k, KB,
                               (const int64_t *)Lp, (const int64_t *)Li, Lx, slot, rscale, blocks);
        }
    }
    return hipGetLastError();
}

hipError_t launch_b_convert(int64_t rows, int cols, const float *src, int64_t ld_src, bool fp8,
                            float *cscale, void *dst, int64_t ldk, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    const dim3 grid((unsigned)((ldk + 31) / 32), (unsigned)((cols + 31) / 32));
    if (fp8) {
        hipLaunchKernelGGL(col_scale_kernel, dim3((cols + 255) / 256), dim3(256), 0, s, rows, cols, src, ld_src,
                           cscale);
        hipLaunchKernelGGL(b_transpose_kernel<true>, grid, dim3(256), 0, s, (int)rows, cols, src, ld_src, cscale,
                           (uint8_t *)dst, ldk);
    } else {
        hipLaunchKernelGGL(b_transpose_kernel<false>, grid, dim3(256), 0, s, (int)rows, cols, src, ld_src,
                           (const float *)nullptr, (uint16_t *)dst, ldk);
    }
    return hipGetLastError();
}

template <bool FP8, int NT, int WP, int PW>
static hipError_t launch_exec_t(const MfmaLaunch &a, hipStream_t s) {
    const dim3 grid((a.P + WP * PW - 1) / (WP * PW), (a.N + NT - 1) / NT, a.ks);
    const int vec_ok =
        ((uintptr_t)a.B % 16 == 0 && (a.ldb * (int64_t)sizeof(typename Elt<FP8>::T)) % 16 == 0) ? 1 : 0;
    if (a.slot)
        hipLaunchKernelGGL((mfma_exec_kernel<FP8, NT, WP, PW, true>), grid, dim3(64 * WP), 0, s, a.M, a.N, a.KB,
                           a.P, static_cast<const typename Elt<FP8>::T *>(a.B), a.ldb, a.blocks, a.slot, a.rscale,
                           a.cscale, a.C, a.ldc, a.ks, a.work, vec_ok);
    else
        hipLaunchKernelGGL((mfma_exec_kernel<FP8, NT, WP, PW, false>), grid, dim3(64 * WP), 0, s, a.M, a.N, a.KB,
                           a.P, static_cast<const typename Elt<FP8>::T *>(a.B), a.ldb, a.blocks,
                           (const int32_t *)nullptr, a.rscale, a.cscale, a.C, a.ldc, a.ks, a.work, vec_ok);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || a.ks <= 1) return e;
    const int64_t n = (int64_t)a.M * a.N;
    hipLaunchKernelGGL((split_reduce_kernel<FP8>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a.M,
                       a.N, a.ks, a.work, a.rscale, a.cscale, a.C, a.ldc);
    return hipGetLastError();
}

// Launch shapes: (N tile, waves per workgroup, panels per wave).
constexpr int kShapes[][3] = {{64, 4, 2}, {128, 4, 2}, {64, 4, 1}, {128, 2, 2}, {32, 4, 1}, {128, 4, 1},
                              {64, 8, 1}, {128, 8, 1}, {64, 16, 1}, {32, 8, 1}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

template <bool FP8>
static hipError_t launch_exec_dt(const MfmaLaunch &a, hipStream_t s) {
    switch (a.shape) {
        case 0: return launch_exec_t<FP8, 64, 4, 2>(a, s);
        case 1: return launch_exec_t<FP8, 128, 4, 2>(a, s);
        case 2: return launch_exec_t<FP8, 64, 4, 1>(a, s);
        case 3: return launch_exec_t<FP8, 128, 2, 2>(a, s);
        case 4: return launch_exec_t<FP8, 32, 4, 1>(a, s);
        case 5: return launch_exec_t<FP8, 128, 4, 1>(a, s);
        case 6: return launch_exec_t<FP8, 64, 8, 1>(a, s);
        case 7: return launch_exec_t<FP8, 128, 8, 1>(a, s);
        case 8: return launch_exec_t<FP8, 64, 16, 1>(a, s);
        case 9: return launch_exec_t<FP8, 32, 8, 1>(a, s);
        default: return hipErrorInvalidValue;
    }
}

template <bool FP8, int NT, int WP, int PW>
static hipError_t launch_grouped_t(const MfmaProblem *probs, const int *hdr, int grid, hipStream_t s) {
    hipLaunchKernelGGL((mfma_grouped_kernel<FP8, NT, WP, PW>), dim3(grid), dim3(64 * WP), 0, s,
                       (ConstMfmaProblems)probs, hdr);
    return hipGetLastError();
}

template <bool FP8>
static hipError_t launch_grouped_dt(int shape, const MfmaProblem *probs, const int *hdr, int grid, hipStream_t s) {
    switch (shape) {
        case 0: return launch_grouped_t<FP8, 64, 4, 2>(probs, hdr, grid, s);
        case 1: return launch_grouped_t<FP8, 128, 4, 2>(probs, hdr, grid, s);
        case 2: return launch_grouped_t<FP8, 64, 4, 1>(probs, hdr, grid, s);
        case 3: return launch_grouped_t<FP8, 128, 2, 2>(probs, hdr, grid, s);
        case 4: return launch_grouped_t<FP8, 32, 4, 1>(probs, hdr, grid, s);
        case 5: return launch_grouped_t<FP8, 128, 4, 1>(probs, hdr, grid, s);
        case 6: return launch_grouped_t<FP8, 64, 8, 1>(probs, hdr, grid, s);
        case 7: return launch_grouped_t<FP8, 128, 8, 1>(probs, hdr, grid, s);
        case 8: return launch_grouped_t<FP8, 64, 16, 1>(probs, hdr, grid, s);
        case 9: return launch_grouped_t<FP8, 32, 8, 1>(probs, hdr, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_mfma_grouped(bool fp8, int shape, const MfmaProblem *probs, const int *hdr, int grid,
                               hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    return fp8 ? launch_grouped_dt<true>(shape, probs, hdr, grid, s)
               : launch_grouped_dt<false>(shape, probs, hdr, grid, s);
}

// Grouped B preparation: every problem's fp32 B (K x N, row-major) into its
// transposed BF16 / e4m3 operand, 32 x 32 tiles, problem q owning tiles
// [tile_begin, ...) of the table.
template <bool FP8>
__global__ void grouped_transpose_kernel(ConstBPrepProblems probs, const int *__restrict__ hdr) {
    __shared__ float tbuf[32][33];
    const int nprob = hdr[0], total_tiles = hdr[1];
    int q = 0;
    for (int t = blockIdx.x; t < total_tiles; t += gridDim.x) {
        while (q + 1 < nprob && probs[q + 1].tile_begin <= t) ++q;
        const void *src = probs[q].src;
        void *dst = probs[q].dst;
        const float *cscale = probs[q].cscale;
        const int64_t ld_src = probs[q].ld_src, ldk = probs[q].ldk;
        const int K = probs[q].K, N = probs[q].N, src_bf16 = probs[q].src_bf16;
        const int local = t - probs[q].tile_begin;
        const int ktiles = (int)((ldk + 31) / 32);
        const int k0 = (local % ktiles) * 32, n0 = (local / ktiles) * 32;
        const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
        for (int i = ty; i < 32; i += 8) {
            const int k = k0 + i, n = n0 + tx;
            float v = 0.0f;
            if (k < K && n < N) {
                if (src_bf16)  // BF16 source: widen (exact)
                    v = __builtin_bit_cast(float, (uint32_t)static_cast<const uint16_t *>(src)[(int64_t)k * ld_src + n] << 16);
                else
                    v = static_cast<const float *>(src)[(int64_t)k * ld_src + n];
            }
            tbuf[i][tx] = v;
        }
        __syncthreads();
        for (int i = ty; i < 32; i += 8) {
            const int n = n0 + i, k = k0 + tx;
            if (n < N && k < ldk) {
                const float v = tbuf[tx][i];
                if constexpr (FP8)
                    static_cast<uint8_t *>(dst)[(int64_t)n * ldk + k] =
                        to_e4m3((float)((double)v / (double)cscale[n]));
                else
                    static_cast<uint16_t *>(dst)[(int64_t)n * ldk + k] = to_bf16(v);
            }
        }
        __syncthreads();
    }
}

// Grouped widening for the register-tile serving form: every problem's BF16
// B (K x N, row-major, ld_src) into FP32 rows of stride ldk (dst), columns
// [N, w) zero, w = pad_to (ldk when 0); columns from w on are not touched, so
// problems may write side by side into one destination.  The K x w output is
// walked flat, 1024 elements per tile (4 per thread, 256 apart): problem q
// owns ceil(K * w / 1024) tiles from tile_begin.  BF16 -> FP32 is exact.
__global__ void __launch_bounds__(256) grouped_widen_kernel(ConstBPrepProblems probs, const int *__restrict__ hdr) {
    const int nprob = hdr[0], total_tiles = hdr[1];
    int q = 0;
    for (int t = blockIdx.x; t < total_tiles; t += gridDim.x) {
        while (q + 1 < nprob && probs[q + 1].tile_begin <= t) ++q;
        const int64_t ldk = probs[q].ldk;
        const int64_t w = probs[q].pad_to > 0 ? probs[q].pad_to : ldk;
        const int64_t total = (int64_t)probs[q].K * w;
        const int64_t e0 = (int64_t)(t - probs[q].tile_begin) * 1024 + threadIdx.x;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t e = e0 + 256 * u;
            if (e < total) {
                const int64_t k = e / w, n = e - k * w;
                float v = 0.0f;
                if (n < probs[q].N)
                    v = __builtin_bit_cast(
                        float, (uint32_t)static_cast<const uint16_t *>(probs[q].src)[k * probs[q].ld_src + n] << 16);
                static_cast<float *>(probs[q].dst)[k * ldk + n] = v;
            }
        }
    }
}

hipError_t launch_grouped_b_widen(const BPrepProblem *probs, const int *hdr, int grid, hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(grouped_widen_kernel, dim3(grid), dim3(256), 0, s, (ConstBPrepProblems)probs, hdr);
    return hipGetLastError();
}

hipError_t launch_grouped_b_prep(bool fp8, const BPrepProblem *probs, const int *hdr, int grid, hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    if (fp8)
        hipLaunchKernelGGL(grouped_transpose_kernel<true>, dim3(grid), dim3(256), 0, s, (ConstBPrepProblems)probs, hdr);
    else
        hipLaunchKernelGGL(grouped_transpose_kernel<false>, dim3(grid), dim3(256), 0, s, (ConstBPrepProblems)probs,
                           hdr);
    return hipGetLastError();
}

hipError_t launch_mfma_exec(const MfmaLaunch &a, hipStream_t s) {
    if (a.M <= 0 || a.N <= 0) return hipSuccess;
    return a.fp8 ? launch_exec_dt<true>(a, s) : launch_exec_dt<false>(a, s);
}

int mfma_shape_tile(int shape, int *nt, int *wp, int *pw) {
    if (shape < 0 || shape >= kNumShapes) return -1;
    *nt = kShapes[shape][0];
    *wp = kShapes[shape][1];
    *pw = kShapes[shape][2];
    return 0;
}

}  // namespace srt
