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

// All `count` (>= 1) packed columns of one kernel pattern.
// On entry cur[0 .. min(U, count)) holds the B rows of the group's first
// columns (cp[0 ..)).  On exit cp/vp are advanced past the group and
// cur[0 .. min(U, next_count)) holds the rows of the next group's first
// columns.
//
// Scalar metadata is read in whole batches, outside any branch, so the
// compiler issues wide s_load_dwordx{4,8,16} and waits once per batch: the
// column indices of the next U columns, then the values of the current
// columns UG at a time (UG * popcount <= 32 SGPRs).  These batch reads may
// run past the panel's last column; the packed arrays carry SRT_COL_PAD /
// SRT_VAL_PAD zero entries at their end so they stay in bounds, and entries
// past the group are never used.
//
// One batch: the FMAs of up to U columns held in `src` while the loads of the
// following batch go into `dst`.  Returns true when the group is finished
// (then `dst` holds the next group's first columns).
template <int PAT, int MR, int V, int U>
__device__ __forceinline__ bool group_batch(float (&acc)[MR][V], typename VecOf<V>::type (&src)[U],
                                            typename VecOf<V>::type (&dst)[U], int &left,
                                            int next_count, const int32_t *__restrict__ &cp,
                                            const float *__restrict__ &vp, const BRows<V> &bb) {
    constexpr int PC = __builtin_popcount(PAT);
    constexpr int UG = PC * U <= 32 ? U : (32 / PC > 0 ? 32 / PC : 1);
    const int vc = left < U ? left : U;
    left -= vc;
    const int want = left > 0 ? left : next_count;
    const int vn = want < U ? want : U;
    int col[U];
    load_uniform(cp + vc, col);
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (u < vn) dst[u] = bb.load(col[u]);
#pragma unroll
    for (int s = 0; s < U; s += UG) {
        if (s < vc) {
            float av[UG * PC];
            load_uniform(vp + s * PC, av);
#pragma unroll
            for (int u = 0; u < UG; ++u)
                if (s + u < U && s + u < vc) fma_pattern<PAT, MR, V>(acc, av + u * PC, src[s + u]);
        }
    }
    cp += vc;
    vp += vc * PC;
    return left <= 0;
}

// Batches alternate between two register buffers, so no batch copies its
// prefetched rows; a group that ends on the second buffer copies it once.
template <int PAT, int MR, int V, int U>
__device__ __forceinline__ void run_group(float (&acc)[MR][V], typename VecOf<V>::type (&cur)[U],
                                          int count, int next_count,
                                          const int32_t *__restrict__ &cp,
                                          const float *__restrict__ &vp, const BRows<V> &bb) {
    typedef typename VecOf<V>::type vec;
    vec alt[U];
    int left = count;
    while (true) {
        if (group_batch<PAT, MR, V, U>(acc, cur, alt, left, next_count, cp, vp, bb)) {
#pragma unroll
            for (int u = 0; u < U; ++u) cur[u] = alt[u];
            return;
        }
        if (group_batch<PAT, MR, V, U>(acc, alt, cur, left, next_count, cp, vp, bb)) return;
    }
}

// D::run<V, U>(id, acc, cur, count, next_count, cp, vp, bb) is the generated
// switch over the pattern set's ids.
//
// Work decomposition (load balanced, like the reference's KD_*LoadBalanced
// kernel descriptors): panel p is cut into k_p contiguous chunks of its packed
// columns, chunk j = [floor(j*n_p/k_p), floor((j+1)*n_p/k_p)), one wavefront
// each; the chunks of a panel sit in consecutive wave slots of one workgroup
// and light panels share workgroups.  Each wave reads its descriptor
//   d0 = {panel, group of chunk start, skip in that group, first column}
//   d1 = {first value, end column, k_p, j}
// (panel -1 = idle slot).  Waves j > 0 park their partial accumulators in LDS
// and wave j = 0 adds them in chunk order:
//   C = ((chain_0 + chain_1) + chain_2) + ...
// chain_0 starting from C (load_c) or 0, the others from 0.  With every
// k_p = 1 this is exactly the reference's single chain per panel.
template <int MR, int V, int U, class D>
__global__ void __launch_bounds__(1024)
    exec_kernel(const int2 *__restrict__ groups, const int64_t *__restrict__ grp_ptr,
                const int4 *__restrict__ desc, const int64_t *__restrict__ col_ptr,
                const int64_t *__restrict__ val_ptr, const int32_t *__restrict__ cols,
                const float *__restrict__ vals, const float *__restrict__ B, uint32_t b_bytes,
                uint32_t ldb, float *__restrict__ C, int64_t ldc, int M, int N, int n_wblocks,
                int reduce, int load_c, int xcd_remap) {
    typedef typename VecOf<V>::type vec;
    extern __shared__ float red[];  // [waves][MR][V][64]
    const int waves = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    int b = blockIdx.x;
    if (xcd_remap) b = xcd_swizzle(b, gridDim.x);
    const int ntile = b / n_wblocks;
    const int sb = b - ntile * n_wblocks;
    const int4 *__restrict__ dw = desc + ((int64_t)sb * waves + wave) * 2;
    const int4 d0 = dw[0];
    const int4 d1 = dw[1];
    const int p = d0.x;
    const bool live = p >= 0;  // wave-uniform
    if (!reduce && !live) return;  // no barrier below when nothing is reduced
    const int n = ntile * (64 * V) + lane * V;
    const bool active = n < N;
    const int nl = active ? n : 0;  // inactive lanes read columns 0..V-1 (in bounds)
    const int kp = d1.z, j = d1.w;

    float acc[MR][V];
#pragma unroll
    for (int r = 0; r < MR; ++r)
#pragma unroll
        for (int jj = 0; jj < V; ++jj) acc[r][jj] = 0.0f;

    if (live) {
        BRows<V> bb;
        bb.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(B), (short)0, (int)b_bytes,
                                                    0x00020000);
        bb.voff = (uint32_t)nl * 4u;
        bb.row_bytes = ldb * 4u;
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
        int remaining = d1.y - d0.w;
        if (remaining > 0) {
            const int2 *__restrict__ gl = groups + grp_ptr[p] + d0.y;
            const int32_t *__restrict__ cp = cols + col_ptr[p] + d0.w;
            const float *__restrict__ vp = vals + val_ptr[p] + d1.x;
            int2 e = gl[0];
            int cnt = e.y - d0.z < remaining ? e.y - d0.z : remaining;
            vec cur[U];
            {
                const int c0 = cnt < U ? cnt : U;
                int col[U];
                load_uniform(cp, col);  // padded: in bounds
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (u < c0) cur[u] = bb.load(col[u]);
            }
            while (true) {
                remaining -= cnt;
                int2 en = make_int2(0, 0);
                if (remaining > 0) en = gl[1];
                const int ncnt = en.y < remaining ? en.y : remaining;
                D::template run<V, U>(e.x, acc, cur, cnt, ncnt, cp, vp, bb);
                if (remaining <= 0) break;
                e = en;
                cnt = ncnt;
                ++gl;
            }
        }
    }

    if (reduce) {
        if (live && j > 0) {
            float *mine = red + (int64_t)wave * MR * V * 64;
#pragma unroll
            for (int r = 0; r < MR; ++r)
#pragma unroll
                for (int jj = 0; jj < V; ++jj) mine[(r * V + jj) * 64 + lane] = acc[r][jj];
        }
        __syncthreads();
        if (!live || j > 0) return;
        for (int w = 1; w < kp; ++w) {
            const float *part = red + (int64_t)(wave + w) * MR * V * 64;
#pragma unroll
            for (int r = 0; r < MR; ++r)
#pragma unroll
                for (int jj = 0; jj < V; ++jj) acc[r][jj] = acc[r][jj] + part[(r * V + jj) * 64 + lane];
        }
    }

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

// LDS bytes the chunk reduction needs: one partial tile per wave slot.
inline size_t exec_lds_bytes(int MR, int V, int waves, int reduce) {
    return reduce ? (size_t)waves * MR * V * 64 * sizeof(float) : 0;
}

template <int MR, int V, int U, class D>
hipError_t launch_exec(const ExecLaunch &a, hipStream_t s) {
    const int n_tiles = (a.N + 64 * V - 1) / (64 * V);
    const long long nblocks = (long long)a.n_wblocks * n_tiles;
    if (nblocks == 0) return hipSuccess;
    if (nblocks > 0x7fffffffLL || a.waves_per_block < 1 || a.waves_per_block > 16)
        return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((exec_kernel<MR, V, U, D>), dim3((unsigned)nblocks), dim3(64 * a.waves_per_block),
                       exec_lds_bytes(MR, V, a.waves_per_block, a.reduce), s,
                       reinterpret_cast<const int2 *>(a.groups), a.grp_ptr,
                       reinterpret_cast<const int4 *>(a.desc), a.col_ptr, a.val_ptr, a.cols, a.vals,
                       a.B, a.b_bytes, (uint32_t)a.ldb, a.C, a.ldc, a.M, a.N, a.n_wblocks, a.reduce,
                       a.load_c, a.xcd_remap);
    return hipGetLastError();
}

}  // namespace srt
