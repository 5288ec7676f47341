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
template <int PAT, int MR, int V, int U>
__device__ __forceinline__ void run_group(float (&acc)[MR][V], typename VecOf<V>::type (&cur)[U],
                                          int count, int next_count,
                                          const int32_t *__restrict__ &cp,
                                          const float *__restrict__ &vp, const BRows<V> &bb) {
    constexpr int PC = __builtin_popcount(PAT);
    typedef typename VecOf<V>::type vec;
    int left = count;
    while (true) {
        const int vc = left < U ? left : U;
        left -= vc;
        const int want = left > 0 ? left : next_count;
        const int vn = want < U ? want : U;
        vec nxt[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (u < vn) nxt[u] = bb.load(cp[vc + u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (u < vc) fma_pattern<PAT, MR, V>(acc, vp + u * PC, cur[u]);
        cp += vc;
        vp += vc * PC;
#pragma unroll
        for (int u = 0; u < U; ++u) cur[u] = nxt[u];
        if (left <= 0) break;
    }
}

// D::run<V, U>(id, acc, cur, count, next_count, cp, vp, bb) is the generated
// switch over the pattern set's ids.
template <int MR, int V, int U, class D>
__global__ void __launch_bounds__(512)
    exec_kernel(const int2 *__restrict__ groups, const int64_t *__restrict__ grp_ptr,
                const int64_t *__restrict__ col_ptr, const int64_t *__restrict__ val_ptr,
                const int32_t *__restrict__ cols, const float *__restrict__ vals,
                const float *__restrict__ B, uint32_t b_bytes, uint32_t ldb,
                float *__restrict__ C, int64_t ldc, int M, int N, int num_panels, int p_groups,
                int load_c, int xcd_remap) {
    typedef typename VecOf<V>::type vec;
    const int waves = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    int b = blockIdx.x;
    if (xcd_remap) b = xcd_swizzle(b, gridDim.x);
    const int ntile = b / p_groups;
    const int pg = b - ntile * p_groups;
    const int p = pg * waves + wave;
    if (p >= num_panels) return;
    const int n = ntile * (64 * V) + lane * V;
    const bool active = n < N;
    const int nl = active ? n : 0;  // inactive lanes read columns 0..V-1 (in bounds)

    BRows<V> bb;
    bb.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(B), (short)0, (int)b_bytes,
                                                0x00020000);
    bb.voff = (uint32_t)nl * 4u;
    bb.row_bytes = ldb * 4u;

    float acc[MR][V];
#pragma unroll
    for (int r = 0; r < MR; ++r)
#pragma unroll
        for (int j = 0; j < V; ++j) acc[r][j] = 0.0f;
    if (load_c) {
#pragma unroll
        for (int r = 0; r < MR; ++r) {
            const int row = p * MR + r;
            if (row < M && active) {
                const vec c = load_vec<V>(C + (int64_t)row * ldc + nl);
#pragma unroll
                for (int j = 0; j < V; ++j) acc[r][j] = lane_elem<V>(c, j);
            }
        }
    }

    const int64_t g0 = grp_ptr[p];
    const int ng = (int)(grp_ptr[p + 1] - g0);
    const int2 *__restrict__ gl = groups + g0;
    const int32_t *__restrict__ cp = cols + col_ptr[p];
    const float *__restrict__ vp = vals + val_ptr[p];
    vec cur[U];
    if (ng > 0) {
        const int c0 = gl[0].y < U ? gl[0].y : U;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (u < c0) cur[u] = bb.load(cp[u]);
    }
    for (int g = 0; g < ng; ++g) {
        const int2 e = gl[g];
        const int next = g + 1 < ng ? gl[g + 1].y : 0;
        D::template run<V, U>(e.x, acc, cur, e.y, next, cp, vp, bb);
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
                    for (int j = 0; j < V; ++j) c[j] = acc[r][j];
                }
                __builtin_nontemporal_store(c, reinterpret_cast<vec *>(C + (int64_t)row * ldc + n));
            }
        }
    }
}

template <int MR, int V, int U, class D>
hipError_t launch_exec(const ExecLaunch &a, hipStream_t s) {
    const int waves = a.waves_per_block;
    const int p_groups = (a.num_panels + waves - 1) / waves;
    const int n_tiles = (a.N + 64 * V - 1) / (64 * V);
    const long long nblocks = (long long)p_groups * n_tiles;
    if (nblocks == 0) return hipSuccess;
    if (nblocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((exec_kernel<MR, V, U, D>), dim3((unsigned)nblocks), dim3(64 * waves), 0, s,
                       reinterpret_cast<const int2 *>(a.groups), a.grp_ptr, a.col_ptr, a.val_ptr,
                       a.cols, a.vals, a.B, a.b_bytes, (uint32_t)a.ldb, a.C, a.ldc, a.M, a.N,
                       a.num_panels, p_groups, a.load_c, a.xcd_remap);
    return hipGetLastError();
}

}  // namespace srt
