// Nano-kernel executor template for CDNA4 (gfx950).
//
// One wavefront owns one M_r-row panel of A and a 64*V-column slice of B/C:
// lane l holds C[panel rows, n0 + l*V .. n0 + l*V + V) in VGPRs (acc[M_r][V]).
// The panel's packed columns are grouped by kernel id; the panel's list of
// non-empty groups (id, count) is walked in id order and a generated switch
// jumps to the group's unrolled body.  Inside a group every packed column
//   - reads its B row slice with one buffer load per lane (scalar row offset
//     col * ldb * 4, vector lane offset),
//   - reads its values through the scalar unit (the panel is wave-uniform, so
//     column indices and values are SGPR operands),
//   - issues one v_fmac_f32 per set bit of the group's pattern and per column
//     of the lane (bit positions are template constants: static accumulator
//     indices, no divergence).
// B loads are double-buffered U columns deep and the buffer carries over
// group boundaries: the last batch of a group issues the loads of the next
// group's first columns, so the load stream never drains between groups.
//
// This is the reference's per-pattern loop (tools/codegen/
// gen_vector_specialized.py:266-296, driver template:281-297) laid out for
// 64-wide wavefronts: the AVX-512 register tile M_r x N_r*16 becomes
// M_r x (64*V) spread over the lanes.  Accumulation order per output element
// is the reference's (packed columns in (id, column) order, one fused
// multiply-add each), so results are bit-identical to a sequential fmaf chain
// over the same packed data.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "srt_internal.h"

namespace srt {

template <int V>
struct VecOf {
    typedef float __attribute__((ext_vector_type(V))) type;
};
template <>
struct VecOf<1> {
    typedef float type;
};

template <int V>
__device__ __forceinline__ typename VecOf<V>::type load_vec(const float *p) {
    return *reinterpret_cast<const typename VecOf<V>::type *>(p);
}

template <int V>
__device__ __forceinline__ float lane_elem(const typename VecOf<V>::type &v, int j) {
    return v[j];
}
template <>
__device__ __forceinline__ float lane_elem<1>(const float &v, int) {
    return v;
}

// Buffer-resource view of B: a row is addressed by a scalar byte offset
// (col * ldb * 4) and a lane by a vector byte offset.
template <int V>
struct BRows {
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t voff;       // lane byte offset inside a row
    uint32_t row_bytes;  // ldb * 4

    __device__ __forceinline__ typename VecOf<V>::type load(int col) const {
        const uint32_t so = (uint32_t)col * row_bytes;
        if constexpr (V == 1) {
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, voff, so, 0));
        } else if constexpr (V == 2) {
            return __builtin_bit_cast(typename VecOf<2>::type,
                                      __builtin_amdgcn_raw_buffer_load_b64(rsrc, voff, so, 0));
        } else {
            return __builtin_bit_cast(typename VecOf<4>::type,
                                      __builtin_amdgcn_raw_buffer_load_b128(rsrc, voff, so, 0));
        }
    }
};

// LDS view of a staged B tile: row k of the workgroup's 64*V columns sits at
// lds[k * 64 * V ..), lane l reads its V columns at offset l * V.  One wave
// reads one contiguous 256*V-byte row: no bank conflicts.
template <int V>
struct LdsRows {
    const float *base;  // staged tile + lane * V

    __device__ __forceinline__ typename VecOf<V>::type load(int col) const {
        return *reinterpret_cast<const typename VecOf<V>::type *>(base + col * (64 * V));
    }
};

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X.  This maps
// hardware block b to a logical index so that XCD x receives one contiguous
// run of logical indices (and hence, with N-tile-major logical order, one or
// two N slices of B that stay in that XCD's L2).
__device__ __forceinline__ int xcd_swizzle(int b, int nb) {
    constexpr int NX = 8;
    const int q = nb / NX, r = nb % NX;
    const int x = b % NX, i = b / NX;
    return x * q + (x < r ? x : r) + i;
}

// Reads N consecutive wave-uniform words as wide vector values.  Element-wise
// loads would be sunk by LLVM into the per-column branches that use them, so
// each column would wait for its own scalar load; a vector load used by
// several branches stays in front of them (one s_load_dwordx{2..16}, one wait).
template <typename T, int N>
__device__ __forceinline__ void load_uniform(const T *__restrict__ p, T (&out)[N]) {
    int i = 0;
#pragma unroll
    for (; i + 8 <= N; i += 8) {
        T __attribute__((ext_vector_type(8))) v;
        __builtin_memcpy(&v, p + i, sizeof(v));
#pragma unroll
        for (int j = 0; j < 8; ++j) out[i + j] = v[j];
    }
    if constexpr (N % 8 >= 4) {
        T __attribute__((ext_vector_type(4))) v;
        __builtin_memcpy(&v, p + i, sizeof(v));
#pragma unroll
        for (int j = 0; j < 4; ++j) out[i + j] = v[j];
        i += 4;
    }
    if constexpr (N % 4 >= 2) {
        T __attribute__((ext_vector_type(2))) v;
        __builtin_memcpy(&v, p + i, sizeof(v));
        out[i] = v[0];
        out[i + 1] = v[1];
        i += 2;
    }
    if constexpr (N % 2 == 1) out[i] = p[i];
}

// acc[r][:] += av[k] * b for the k-th set bit r of PAT.
template <int PAT, int MR, int V>
__device__ __forceinline__ void fma_pattern(float (&acc)[MR][V], const float *__restrict__ av,
                                            const typename VecOf<V>::type &b) {
    int k = 0;
#pragma unroll
    for (int r = 0; r < MR; ++r) {
        if ((PAT >> r) & 1) {
            const float a = av[k];
#pragma unroll
            for (int j = 0; j < V; ++j) acc[r][j] = __builtin_fmaf(a, lane_elem<V>(b, j), acc[r][j]);
            ++k;
        }
    }
}

// The packed columns of a wave's chunk form a stream of batches: each group
// (kernel id, count) is cut into batches of U columns, the last one partial.
// A batch is run by one case of the generated switch, keyed id * U + n - 1
// for n columns: n unrolled columns of PC = popcount(pattern) fused
// multiply-adds each, accumulator rows fixed at compile time.  The batches
// are precomputed per wave slot as records (key, first value, U B rows; see
// rec_stride), so the loop spends no scalar instructions on walking groups.
//
// Software pipeline.  Records and values are read through the scalar unit
// (wave-uniform SGPR operands).  Scalar loads and LDS reads share one wait
// counter and scalar loads may return out of order, so a wait for a scalar
// load also waits for every LDS read in flight.  Each iteration therefore
// first consumes what the previous one loaded (one wait), then issues
// everything the next iterations need, then runs its FMAs on registers that
// have already arrived:
//
//   iteration i:  rows'  <- B rows of batch i+1   (its record loaded by iteration i-1)
//                 rec''  <- record of batch i+2
//                 vals'  <- first W values of batch i+1
//                 FMAs of batch i from rows and vals
//
// All U rows of a batch are loaded whatever its size (record slots past n
// repeat a valid row), so the loads carry no branches.  A batch that needs
// more than W values reads the rest itself.  Every run ends in a terminator
// record (key -1, valid row and value indices) and the record array carries
// two records of padding, the value array SRT_VAL_PAD zero entries, so every
// look-ahead read stays in bounds; what is read past the run is never used.
constexpr int kValWindow = 16;

// FMAs of one batch of NCOL columns of pattern PAT.
template <int PAT, int MR, int V, int U, int NCOL>
__device__ __forceinline__ void batch_fmas(float (&acc)[MR][V], const typename VecOf<V>::type (&rows)[U],
                                           const float (&av)[kValWindow], const float *__restrict__ vp) {
    constexpr int PC = __builtin_popcount(PAT);
    constexpr int UW = kValWindow / PC < NCOL ? kValWindow / PC : NCOL;  // columns in the window
#pragma unroll
    for (int u = 0; u < UW; ++u) fma_pattern<PAT, MR, V>(acc, av + u * PC, rows[u]);
    if constexpr (UW < NCOL) {
        float rest[(NCOL - UW) * PC];
        load_uniform(vp + UW * PC, rest);
#pragma unroll
        for (int u = UW; u < NCOL; ++u) fma_pattern<PAT, MR, V>(acc, rest + (u - UW) * PC, rows[u]);
    }
}

// vals + i for a record's value index (< 2^31): a 32-bit byte offset, which
// scalar loads take as their offset operand (no 64-bit address arithmetic).
__device__ __forceinline__ const float *value_at(const float *vals, int i) {
    return reinterpret_cast<const float *>(reinterpret_cast<const char *>(vals) + ((uint32_t)i << 2));
}

// Pipeline registers of one iteration.
template <int V, int U>
struct Stage {
    typename VecOf<V>::type rows[U];  // B rows of the batch
    float vals[kValWindow];           // its first values
    int rec[2 + U];                   // record of the following batch (its used words:
                                      // unused destination words of a wide scalar load
                                      // would be reused early and force a wait)
};

// One pipeline iteration: batch (key, first value vo) in `in`, whose
// in.rec is the following batch's record; that batch is prepared in `out`
// (rows, values, and the record after it from rp).
template <int MR, int V, int U, class D, class R>
__device__ __forceinline__ void pipe_step(float (&acc)[MR][V], const Stage<V, U> &in, Stage<V, U> &out, int key,
                                          int vo, const int32_t *__restrict__ prog, uint32_t roff,
                                          const float *__restrict__ vals, const R &bb) {
#pragma unroll
    for (int u = 0; u < U; ++u) out.rows[u] = bb.load(in.rec[2 + u]);
    load_uniform(reinterpret_cast<const int32_t *>(reinterpret_cast<const char *>(prog) + roff), out.rec);
    load_uniform(value_at(vals, in.rec[1]), out.vals);
    D::template fmas<V, U>(key, acc, in.rows, in.vals, value_at(vals, vo));
}

// D::fmas<V, U>(key, acc, rows, vals, vp) is the generated switch over
// (kernel id, batch size); R reads B rows (BRows from global memory, LdsRows
// from the staged tile).
//
// Work decomposition (load balanced, like the reference's KD_*LoadBalanced
// kernel descriptors): panel p is cut into k_p contiguous chunks of its packed
// columns, chunk j = [floor(j*n_p/k_p), floor((j+1)*n_p/k_p)), one wavefront
// each; the chunks of a panel sit in consecutive wave slots of one workgroup
// and light panels share workgroups.  Each wave reads its descriptor
//   d0 = {panel, group of chunk start, first record of its run, first column}
//   d1 = {first value, end column, k_p, j}
//   then (U <= 6) a copy of its first batch record
// (panel -1 = idle slot); the executor uses the panel, the record run, k_p
// and j.  Waves j > 0 park their partial accumulators in LDS
// and wave j = 0 adds them in chunk order:
//   C = ((chain_0 + chain_1) + chain_2) + ...
// chain_0 starting from C (load_c) or 0, the others from 0.  With every
// k_p = 1 this is exactly the reference's single chain per panel.
//
// STAGE = true: the workgroup first copies its whole B tile (K rows x 64*V
// columns) into LDS with coalesced 16-byte loads, so every B row is fetched
// from L2 once per workgroup instead of once per packed column of every
// panel; all packed-column reads then hit LDS.  The partial tiles of the chunk
// reduction reuse the same LDS after a barrier.  Every wave, idle slots
// included, takes part in the copy and reaches every barrier.
// The work of one (logical) workgroup b of one executor launch; shared by
// the single-problem kernel and the grouped one.
template <int MR, int V, int U, bool STAGE, class D>
__device__ __forceinline__ void exec_body(float *lds, const int32_t *__restrict__ prog,
                                          const int4 *__restrict__ desc, const float *__restrict__ vals,
                                          const float *__restrict__ B,
                                          uint32_t b_bytes, uint32_t ldb, float *__restrict__ C, int64_t ldc,
                                          int M, int N, int K, int n_wblocks, int reduce, int load_c, int b) {
    typedef typename VecOf<V>::type vec;
    const int waves = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int ntile = b / n_wblocks;
    const int sb = b - ntile * n_wblocks;
    const int4 *__restrict__ dw = desc + ((int64_t)sb * waves + wave) * 4;
    int dv[16];  // descriptor and (for 2 + U <= 8) the first batch record: one scalar load
    load_uniform(reinterpret_cast<const int *>(dw), dv);
    const int4 d0 = make_int4(dv[0], dv[1], dv[2], dv[3]);
    const int4 d1 = make_int4(dv[4], dv[5], dv[6], dv[7]);
    const int p = d0.x;
    const bool live = p >= 0;  // wave-uniform
    if (!STAGE && !reduce && !live) return;  // no barrier below
    const int n = ntile * (64 * V) + lane * V;
    const bool active = n < N;
    const int nl = active ? n : 0;  // inactive lanes read columns 0..V-1 (in bounds)
    const int kp = d1.z, j = d1.w;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(B), (short)0, (int)b_bytes, 0x00020000);

    if constexpr (STAGE) {
        // Row k, 16-byte chunk c of the tile <- B[k][n0 + 4c .. n0 + 4c + 4).
        // Chunks past column N of the last tile read the next row (or 0 past
        // the end of the buffer); only inactive lanes see them.
        // Eight loads in flight per lane before their LDS writes.
        constexpr int CH = 16 * V;
        constexpr int BATCH = 8;
        const uint32_t n0_bytes = (uint32_t)(ntile * (64 * V)) * 4u;
        const int total = K * CH;
        const int nthr = blockDim.x;
        for (int i0 = threadIdx.x; i0 < total; i0 += BATCH * nthr) {
            float4 v[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int i = i0 + u * nthr;
                if (i < total) {
                    // The row offset differs between lanes: it goes into the
                    // per-lane offset (a scalar offset would be waterfalled).
                    const int k = i / CH, c = i % CH;
                    v[u] = __builtin_bit_cast(
                        float4, __builtin_amdgcn_raw_buffer_load_b128(
                                    rsrc, (uint32_t)k * ldb * 4u + n0_bytes + (uint32_t)c * 16u, 0, 0));
                }
            }
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int i = i0 + u * nthr;
                if (i < total) {
                    const int k = i / CH, c = i % CH;
                    *reinterpret_cast<float4 *>(lds + k * (64 * V) + c * 4) = v[u];
                }
            }
        }
        __syncthreads();
    }

    float acc[MR][V];
#pragma unroll
    for (int r = 0; r < MR; ++r)
#pragma unroll
        for (int jj = 0; jj < V; ++jj) acc[r][jj] = 0.0f;

    if (live) {
        if (load_c && j == 0) {
#pragma unroll
            for (int r = 0; r < MR; ++r) {
                const int row = p * MR + r;
                if (row < M && active) {
                    const vec c = load_vec<V>(C + (int64_t)row * ldc + nl);
#pragma unroll
                    for (int jj = 0; jj < V; ++jj) acc[r][jj] = lane_elem<V>(c, jj);
                }
            }
        }
        // Record run of this slot (its start replaces the descriptor's d0.z);
        // records are addressed by a 32-bit byte offset (the host keeps the
        // array below 4 GiB), the scalar loads' offset operand.
        constexpr uint32_t RB = rec_stride(U) * 4u;
        uint32_t roff = (uint32_t)d0.z * RB;
        int r0[2 + U];
        if constexpr (2 + U <= 8) {
#pragma unroll
            for (int k = 0; k < 2 + U; ++k) r0[k] = dv[8 + k];
        } else {
            load_uniform(reinterpret_cast<const int32_t *>(reinterpret_cast<const char *>(prog) + roff), r0);
        }
        if (r0[0] >= 0) {
            typedef typename std::conditional<STAGE, LdsRows<V>, BRows<V>>::type Rows;
            Rows bb;
            if constexpr (STAGE) {
                bb.base = lds + lane * V;
            } else {
                bb.rsrc = rsrc;
                bb.voff = (uint32_t)nl * 4u;
                bb.row_bytes = ldb * 4u;
            }
            Stage<V, U> s0, s1;
#pragma unroll
            for (int u = 0; u < U; ++u) s0.rows[u] = bb.load(r0[2 + u]);
            load_uniform(reinterpret_cast<const int32_t *>(reinterpret_cast<const char *>(prog) + roff + RB), s0.rec);
            load_uniform(value_at(vals, r0[1]), s0.vals);
            int key = r0[0], vo = r0[1];
            roff += 2 * RB;
            // Two iterations per trip so the stages alternate without copies.
            while (true) {
                const int key1 = s0.rec[0], vo1 = s0.rec[1];
                pipe_step<MR, V, U, D, Rows>(acc, s0, s1, key, vo, prog, roff, vals, bb);
                roff += RB;
                if (key1 < 0) break;
                const int key2 = s1.rec[0], vo2 = s1.rec[1];
                pipe_step<MR, V, U, D, Rows>(acc, s1, s0, key1, vo1, prog, roff, vals, bb);
                roff += RB;
                if (key2 < 0) break;
                key = key2;
                vo = vo2;
            }
        }
    }

    if (reduce) {
        if constexpr (STAGE) __syncthreads();  // every wave is done with the staged tile
        if (live && j > 0) {
            float *mine = lds + (int64_t)wave * MR * V * 64;
#pragma unroll
            for (int r = 0; r < MR; ++r)
#pragma unroll
                for (int jj = 0; jj < V; ++jj) mine[(r * V + jj) * 64 + lane] = acc[r][jj];
        }
        __syncthreads();
        if (!live || j > 0) return;
        for (int w = 1; w < kp; ++w) {
            const float *part = lds + (int64_t)(wave + w) * MR * V * 64;
#pragma unroll
            for (int r = 0; r < MR; ++r)
#pragma unroll
                for (int jj = 0; jj < V; ++jj) acc[r][jj] = acc[r][jj] + part[(r * V + jj) * 64 + lane];
        }
    }
    if (!live) return;

    if (active) {
#pragma unroll
        for (int r = 0; r < MR; ++r) {
            const int row = p * MR + r;
            if (row < M) {
                vec c;
                if constexpr (V == 1) {
                    c = acc[r][0];
                } else {
#pragma unroll
                    for (int jj = 0; jj < V; ++jj) c[jj] = acc[r][jj];
                }
                __builtin_nontemporal_store(c, reinterpret_cast<vec *>(C + (int64_t)row * ldc + n));
            }
        }
    }
}

template <int MR, int V, int U, bool STAGE, class D>
__global__ void __launch_bounds__(1024)
    exec_kernel(const int32_t *__restrict__ prog, const int4 *__restrict__ desc,
                const float *__restrict__ vals, const float *__restrict__ B, uint32_t b_bytes,
                uint32_t ldb, float *__restrict__ C, int64_t ldc, int M, int N, int K, int n_wblocks,
                int reduce, int load_c, int xcd_remap) {
    extern __shared__ float lds[];  // STAGE: B tile [K][64*V]; reduction: [waves][MR][V][64]
    int b = blockIdx.x;
    if (xcd_remap) b = xcd_swizzle(b, gridDim.x);
    exec_body<MR, V, U, STAGE, D>(lds, prog, desc, vals, B, b_bytes, ldb, C, ldc, M, N, K, n_wblocks, reduce,
                                  load_c, b);
}

// Grouped: several independent packed matrices (the layers of a network,
// a batch of requests) in one launch.  Problem q owns logical workgroups
// [block_begin[q], block_begin[q + 1]); all problems share the executor, V,
// U and the workgroup size.
// The table is read through the constant address space: pointers loaded
// from there are known to address global memory, so the uniform metadata
// reads of exec_body stay scalar loads (from a generic table they would
// become vector loads with per-lane results).
typedef const __attribute__((address_space(4))) ExecProblem *ConstProblems;

template <int MR, int V, int U, bool STAGE, class D>
__global__ void __launch_bounds__(1024)
    exec_grouped_kernel(ConstProblems probs, int nprob, int xcd_remap) {
    extern __shared__ float lds[];
    int b = blockIdx.x;
    if (xcd_remap) b = xcd_swizzle(b, gridDim.x);
    int lo = 0, hi = nprob - 1;  // last problem with block_begin <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (probs[mid].block_begin <= b) lo = mid; else hi = mid - 1;
    }
    exec_body<MR, V, U, STAGE, D>(lds, probs[lo].prog, reinterpret_cast<const int4 *>(probs[lo].desc),
                                  probs[lo].vals, probs[lo].B, probs[lo].b_bytes, probs[lo].ldb, probs[lo].C,
                                  probs[lo].ldc, probs[lo].M,
                                  probs[lo].N, probs[lo].K, probs[lo].n_wblocks, probs[lo].reduce,
                                  probs[lo].load_c, b - probs[lo].block_begin);
}

// Grouped and persistent: the problem count and the logical workgroup total
// are read from device memory (hdr[0], hdr[1]) next to the problem table, and
// a fixed grid walks logical workgroups blockIdx.x, blockIdx.x + gridDim.x, ...
// so one captured graph replays any batch written into the same buffers (the
// serving form).  Each logical workgroup runs exec_body exactly as the
// one-launch-per-problem kernel does, so every problem's result is the same
// bits alone or in any batch.
template <int MR, int V, int U, bool STAGE, class D>
__global__ void __launch_bounds__(1024)
    exec_persistent_kernel(ConstProblems probs, const int *__restrict__ hdr) {
    extern __shared__ float lds[];
    const int nprob = hdr[0], total = hdr[1];
    int q = 0;
    for (int b = blockIdx.x; b < total; b += gridDim.x) {
        while (q + 1 < nprob && probs[q + 1].block_begin <= b) ++q;  // b ascends per workgroup
        exec_body<MR, V, U, STAGE, D>(lds, probs[q].prog, reinterpret_cast<const int4 *>(probs[q].desc),
                                      probs[q].vals, probs[q].B, probs[q].b_bytes, probs[q].ldb, probs[q].C,
                                      probs[q].ldc, probs[q].M, probs[q].N, probs[q].K, probs[q].n_wblocks,
                                      probs[q].reduce, probs[q].load_c, b - probs[q].block_begin);
        __syncthreads();  // the next logical workgroup reuses the LDS
    }
}

// LDS bytes of the chunk reduction: one partial tile per wave slot.
inline size_t reduce_lds_bytes(int MR, int V, int waves, int reduce) {
    return reduce ? (size_t)waves * MR * V * 64 * sizeof(float) : 0;
}
// LDS bytes of one launch: the staged B tile (STAGE) and the reduction share it.
inline size_t exec_lds_bytes(int MR, int V, int waves, int reduce, bool stage, int K) {
    const size_t red = reduce_lds_bytes(MR, V, waves, reduce);
    const size_t tile = stage ? (size_t)K * 64 * V * sizeof(float) : 0;
    return red > tile ? red : tile;
}

template <int MR, int V, int U, bool STAGE, class D>
hipError_t launch_exec(const ExecLaunch &a, hipStream_t s) {
    const int n_tiles = (a.N + 64 * V - 1) / (64 * V);
    const long long nblocks = (long long)a.n_wblocks * n_tiles;
    if (nblocks == 0) return hipSuccess;
    if (nblocks > 0x7fffffffLL || a.waves_per_block < 1 || a.waves_per_block > 16)
        return hipErrorInvalidConfiguration;
    const size_t lds = exec_lds_bytes(MR, V, a.waves_per_block, a.reduce, STAGE, a.K);
    if (lds > 65536) {
        // once per instantiation; a function-local static is initialised
        // exactly once even when several threads launch at the same time
        static const hipError_t raised =
            hipFuncSetAttribute(reinterpret_cast<const void *>(&exec_kernel<MR, V, U, STAGE, D>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit());
        if (raised != hipSuccess) return raised;
    }
    hipLaunchKernelGGL((exec_kernel<MR, V, U, STAGE, D>), dim3((unsigned)nblocks),
                       dim3(64 * a.waves_per_block), lds, s, a.prog, reinterpret_cast<const int4 *>(a.desc),
                       a.vals, a.B, a.b_bytes, (uint32_t)a.ldb, a.C, a.ldc, a.M, a.N, a.K, a.n_wblocks,
                       a.reduce, a.load_c, a.xcd_remap);
    return hipGetLastError();
}

template <int MR, int V, int U, bool STAGE, class D>
hipError_t launch_exec_grouped(const ExecProblem *d_probs, int nprob, int total_blocks, int waves, size_t lds,
                               int xcd_remap, hipStream_t s) {
    if (total_blocks <= 0 || nprob <= 0) return hipSuccess;
    if (waves < 1 || waves > 16) return hipErrorInvalidConfiguration;
    if (lds > 65536) {
        static const hipError_t raised =
            hipFuncSetAttribute(reinterpret_cast<const void *>(&exec_grouped_kernel<MR, V, U, STAGE, D>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit());
        if (raised != hipSuccess) return raised;
    }
    hipLaunchKernelGGL((exec_grouped_kernel<MR, V, U, STAGE, D>), dim3((unsigned)total_blocks), dim3(64 * waves),
                       lds, s, (ConstProblems)d_probs, nprob, xcd_remap);
    return hipGetLastError();
}

template <int MR, int V, int U, bool STAGE, class D>
hipError_t launch_exec_persistent(const ExecProblem *d_probs, const int *d_hdr, int grid, int waves, size_t lds,
                                  hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    if (waves < 1 || waves > 16) return hipErrorInvalidConfiguration;
    if (lds > 65536) {
        static const hipError_t raised =
            hipFuncSetAttribute(reinterpret_cast<const void *>(&exec_persistent_kernel<MR, V, U, STAGE, D>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit());
        if (raised != hipSuccess) return raised;
    }
    hipLaunchKernelGGL((exec_persistent_kernel<MR, V, U, STAGE, D>), dim3((unsigned)grid), dim3(64 * waves), lds, s,
                       (ConstProblems)d_probs, d_hdr);
    return hipGetLastError();
}

}  // namespace srt
