// HIP CSR SpMM baseline: the reference's AVX-512 row-times-panel kernel
// (sbench/SOP/_C/pytorch_csr_wrapper.cpp:70-102, dispatch :140-191) laid out
// for 64-wide wavefronts.  One wavefront owns one row of A and a 64*V-column
// slice of B/C; the row's column indices and values are wave-uniform (scalar
// loads), each lane reads V consecutive B columns per stored entry and keeps
// V fp32 accumulators.  Per output element the accumulation is one fused
// multiply-add per stored entry in storage order, like the reference.
#include <hip/hip_runtime.h>

#include "exec_kernel.h"
#include "srt_internal.h"

namespace srt {

namespace {

template <typename I, int V, int U>
__global__ void __launch_bounds__(256)
    csr_kernel(int m, int k, int n, I nnz, const I *__restrict__ rowptr,
               const I *__restrict__ colind, const float *__restrict__ vals,
               const float *__restrict__ B, int64_t ldb, float *__restrict__ C, int64_t ldc,
               int row_groups) {
    typedef typename VecOf<V>::type vec;
    const int waves = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int b = xcd_swizzle(blockIdx.x, gridDim.x);
    const int ntile = b / row_groups;
    const int rg = b - ntile * row_groups;
    const int row = rg * waves + wave;
    if (row >= m) return;
    const int col = ntile * (64 * V) + lane * V;
    const bool active = col < n;
    const float *__restrict__ Bp = B + (active ? col : 0);
    float acc[V];
#pragma unroll
    for (int j = 0; j < V; ++j) acc[j] = 0.0f;
    // Entries outside [0, nnz) and column indices outside [0, k) are never
    // dereferenced: rows are clamped to the arrays, bad columns are skipped.
    I q = rowptr[row];
    I end = rowptr[row + 1];
    q = q < 0 ? 0 : (q > nnz ? nnz : q);
    end = end > nnz ? nnz : (end < q ? q : end);
    for (; q + U <= end; q += U) {
        vec bv[U];
        float a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const I c = colind[q + u];
            const bool ok = c >= 0 && c < (I)k;
            bv[u] = load_vec<V>(Bp + (int64_t)(ok ? c : 0) * ldb);
            a[u] = ok ? vals[q + u] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int j = 0; j < V; ++j) acc[j] = __builtin_fmaf(a[u], lane_elem<V>(bv[u], j), acc[j]);
    }
    for (; q < end; ++q) {
        const I c = colind[q];
        const bool ok = c >= 0 && c < (I)k;
        const vec bv = load_vec<V>(Bp + (int64_t)(ok ? c : 0) * ldb);
        const float a = ok ? vals[q] : 0.0f;
#pragma unroll
        for (int j = 0; j < V; ++j) acc[j] = __builtin_fmaf(a, lane_elem<V>(bv, j), acc[j]);
    }
    if (active) {
        vec c;
        if constexpr (V == 1) {
            c = acc[0];
        } else {
#pragma unroll
            for (int j = 0; j < V; ++j) c[j] = acc[j];
        }
        __builtin_nontemporal_store(c, reinterpret_cast<vec *>(C + (int64_t)row * ldc + col));
    }
}

template <typename I, int V>
hipError_t launch_csr_t(int m, int k, int n, int64_t nnz, const I *rowptr, const I *colind,
                        const float *vals, const float *B, int64_t ldb, float *C, int64_t ldc,
                        hipStream_t s) {
    const int waves = 4;
    const int row_groups = (m + waves - 1) / waves;
    const int n_tiles = (n + 64 * V - 1) / (64 * V);
    const long long nblocks = (long long)row_groups * n_tiles;
    if (nblocks == 0) return hipSuccess;
    if (nblocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((csr_kernel<I, V, 4>), dim3((unsigned)nblocks), dim3(64 * waves), 0, s, m,
                       k, n, (I)nnz, rowptr, colind, vals, B, ldb, C, ldc, row_groups);
    return hipGetLastError();
}

template <typename I>
hipError_t launch_csr_i(int m, int k, int n, int64_t nnz, const void *rowptr, const void *colind,
                        const float *vals, const float *B, int64_t ldb, float *C, int64_t ldc, int V,
                        hipStream_t s) {
    const I *rp = (const I *)rowptr;
    const I *ci = (const I *)colind;
    switch (V) {
        case 4: return launch_csr_t<I, 4>(m, k, n, nnz, rp, ci, vals, B, ldb, C, ldc, s);
        case 2: return launch_csr_t<I, 2>(m, k, n, nnz, rp, ci, vals, B, ldb, C, ldc, s);
        default: return launch_csr_t<I, 1>(m, k, n, nnz, rp, ci, vals, B, ldb, C, ldc, s);
    }
}

}  // namespace

hipError_t launch_csr(int m, int k, int n, int64_t nnz, const void *rowptr, const void *colind,
                      int index_bytes, const float *vals, const float *B, int64_t ldb, float *C,
                      int64_t ldc, int vec_width, hipStream_t s) {
    if (index_bytes == 8)
        return launch_csr_i<int64_t>(m, k, n, nnz, rowptr, colind, vals, B, ldb, C, ldc, vec_width, s);
    return launch_csr_i<int32_t>(m, k, n, nnz, rowptr, colind, vals, B, ldb, C, ldc, vec_width, s);
}

}  // namespace srt
