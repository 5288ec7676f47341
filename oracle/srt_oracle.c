/*
 * C restatement of the reference's sparse-register-tiling packer and
 * executors.  TEST INFRASTRUCTURE ONLY: loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py through oracle/cbind.py, never by the product.
 *
 * Packed layout produced here (flat arrays over all panels):
 *   counts[p*NK + id]   number of packed columns of kernel id in panel p
 *                       (MicroKernelPackedData::nkern_counts)
 *   col_ptr[p]..[p+1]   range of panel p in cols[]      (num_col_indices)
 *   val_ptr[p]..[p+1]   range of panel p in vals[]      (num_nnz)
 *   cols[]              packed B-row indices, sorted by id, stable by column
 *   vals[]              one value per set bit of each packed column's kernel
 *                       pattern, rows ascending, 0.0f where the pattern pads
 *
 * Rules restated (reference file:line):
 *   pattern codes            sbench/SOP/sop_utils.py:25-30, sop_driver.py:50-57
 *   split into kernel pats   sbench/SOP/sop_driver.py:58-65 (mapping list order)
 *   stable sort by id,       sbench/SOP/_C/pytorch_wrapper_intrin.cpp.template:63-98,
 *   zero pattern dropped     :143-156
 *   value compression        template:158-217
 *   executor                 template:265-305 and the per-pattern loop of
 *                            tools/codegen/gen_vector_specialized.py:266-349:
 *                            per kernel id in id order, per packed column, one
 *                            B row, one fused multiply-add per set bit.
 *   CSR executor             sbench/SOP/_C/pytorch_csr_wrapper.cpp:70-102
 *
 * Every accumulation is a sequential chain of fmaf() in the reference's order,
 * so the result is bit-identical to any implementation that performs the same
 * chain with a correctly rounded fused multiply-add.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define SRT_ORACLE_API __attribute__((visibility("default")))

static int popc(unsigned x) { return __builtin_popcount(x); }

/* Pattern code of column c of panel p of a dense row-major matrix. */
static unsigned code_of(const float *A, long long lda, int M, int M_r, int p, int c) {
    unsigned code = 0;
    for (int r = 0; r < M_r; ++r) {
        int row = p * M_r + r;
        if (row < M && A[(long long)row * lda + c] != 0.0f) code |= 1u << r;
    }
    return code;
}

/*
 * Sizes pass: per-panel packed column and value counts.
 * lut_ptr has 2^M_r + 1 entries; lut_ids[lut_ptr[code] .. lut_ptr[code+1])
 * are the kernel ids the code is split into, in mapping-list order.
 * id_pattern[id] is the kernel pattern bitmask of id.
 */
SRT_ORACLE_API
int srt_oracle_pack_sizes(int M, int K, const float *A, long long lda, int M_r,
                          const int *lut_ptr, const int *lut_ids, int NK, const int *id_pattern,
                          long long *col_ptr, long long *val_ptr) {
    int P = (M + M_r - 1) / M_r;
    col_ptr[0] = 0;
    val_ptr[0] = 0;
    for (int p = 0; p < P; ++p) {
        long long nc = 0, nv = 0;
        for (int c = 0; c < K; ++c) {
            unsigned code = code_of(A, lda, M, M_r, p, c);
            if (!code) continue;
            for (int e = lut_ptr[code]; e < lut_ptr[code + 1]; ++e) {
                int id = lut_ids[e];
                if (id < 0 || id >= NK) return -1;
                nc += 1;
                nv += popc((unsigned)id_pattern[id]);
            }
        }
        col_ptr[p + 1] = col_ptr[p] + nc;
        val_ptr[p + 1] = val_ptr[p] + nv;
    }
    return 0;
}

/*
 * Fill pass.  Stable counting sort by id per panel (equivalent to the
 * reference's std::stable_sort with the zero pattern last, template:63-73,
 * since zero patterns produce no entry here).  Returns the number of stored
 * values that are 0.0f (the reference's padding_total, template:197-201).
 */
SRT_ORACLE_API
long long srt_oracle_pack_fill(int M, int K, const float *A, long long lda, int M_r,
                               const int *lut_ptr, const int *lut_ids, int NK, const int *id_pattern,
                               int col_offset, const long long *col_ptr, const long long *val_ptr,
                               int *counts, int *cols, float *vals) {
    int P = (M + M_r - 1) / M_r;
    long long padding = 0;
    long long *cur_c = (long long *)malloc(sizeof(long long) * (size_t)(NK > 0 ? NK : 1));
    long long *cur_v = (long long *)malloc(sizeof(long long) * (size_t)(NK > 0 ? NK : 1));
    for (int p = 0; p < P; ++p) {
        int *cnt = counts + (long long)p * NK;
        memset(cnt, 0, sizeof(int) * (size_t)NK);
        for (int c = 0; c < K; ++c) {
            unsigned code = code_of(A, lda, M, M_r, p, c);
            if (!code) continue;
            for (int e = lut_ptr[code]; e < lut_ptr[code + 1]; ++e) cnt[lut_ids[e]]++;
        }
        long long oc = col_ptr[p], ov = val_ptr[p];
        for (int id = 0; id < NK; ++id) {
            cur_c[id] = oc;
            cur_v[id] = ov;
            oc += cnt[id];
            ov += (long long)cnt[id] * popc((unsigned)id_pattern[id]);
        }
        for (int c = 0; c < K; ++c) {
            unsigned code = code_of(A, lda, M, M_r, p, c);
            if (!code) continue;
            for (int e = lut_ptr[code]; e < lut_ptr[code + 1]; ++e) {
                int id = lut_ids[e];
                cols[cur_c[id]++] = c + col_offset;
                unsigned pat = (unsigned)id_pattern[id];
                for (int r = 0; pat; ++r, pat >>= 1) {
                    if (!(pat & 1u)) continue;
                    int row = p * M_r + r;
                    float v = row < M ? A[(long long)row * lda + c] : 0.0f;
                    /* A position the pattern pads is not a stored entry of A:
                     * it holds +0.0f whatever the sign of a dense zero. */
                    if (v == 0.0f) {
                        padding++;
                        v = 0.0f;
                    }
                    vals[cur_v[id]++] = v;
                }
            }
        }
    }
    free(cur_c);
    free(cur_v);
    return padding;
}

/*
 * Executor: C[rows, 0:N] (+)= packed A . B (template:281-297).
 *
 * nsplit == 1 is the reference's order: one fused multiply-add chain per
 * output element over the panel's packed columns, starting from C (load_c)
 * or zero.  For nsplit > 1 the panel's packed columns are cut into nsplit
 * contiguous chunks, chunk w = [floor(w*n/nsplit), floor((w+1)*n/nsplit)),
 * each chunk runs its own chain (chunk 0 from C or zero, the others from
 * zero) and the chunk results are added left to right - the order the GPU
 * executor uses when it spreads one panel over nsplit wavefronts.
 * Rows of the last panel past M are computed and discarded.
 *
 * The inner loops are compiled for AVX-512, AVX2+FMA and baseline x86-64 and
 * chosen by CPU feature at first use; fmaf is exact in every variant.
 */
#define SRT_EXEC_PANEL(NAME, ATTR)                                                               \
    ATTR static void NAME(int M, int N, int M_r, int NK, int KT, const int *id_pattern,          \
                          const int *cnt, const int *cols, const float *vals, const float *B,    \
                          long long ldb, float *C, long long ldc, int load_c, int p, int nsplit, \
                          long long ncols, float *acc) {                                         \
        for (int w = 0; w < nsplit; ++w)                                                         \
            for (int r = 0; r < M_r; ++r) {                                                      \
                int row = p * M_r + r;                                                           \
                float *a = acc + ((long long)w * M_r + r) * N;                                   \
                if (w == 0 && load_c && row < M)                                                 \
                    memcpy(a, C + (long long)row * ldc, sizeof(float) * (size_t)N);              \
                else                                                                             \
                    memset(a, 0, sizeof(float) * (size_t)N);                                     \
            }                                                                                    \
        long long c = 0, v = 0;                                                                  \
        int w = 0;                                                                               \
        long long w_end = ncols / nsplit;                                                        \
        for (int t = 0; t < KT * NK; ++t) {                                                      \
            const int id = t % NK;                                                               \
            unsigned pattern = (unsigned)id_pattern[id];                                         \
            for (int i = 0; i < cnt[t]; ++i, ++c) {                                              \
                while (c >= w_end && w < nsplit - 1) {                                           \
                    ++w;                                                                         \
                    w_end = (long long)(w + 1) * ncols / nsplit;                                 \
                }                                                                                \
                const float *b = B + (long long)cols[c] * ldb;                                   \
                unsigned pat = pattern;                                                          \
                for (int r = 0; pat; ++r, pat >>= 1) {                                           \
                    if (!(pat & 1u)) continue;                                                   \
                    float av = vals[v++];                                                        \
                    float *a = acc + ((long long)w * M_r + r) * N;                               \
                    for (int j = 0; j < N; ++j) a[j] = fmaf(av, b[j], a[j]);                     \
                }                                                                                \
            }                                                                                    \
        }                                                                                        \
        for (int r = 0; r < M_r; ++r) {                                                          \
            int row = p * M_r + r;                                                               \
            float *a = acc + (long long)r * N;                                                   \
            for (int u = 1; u < nsplit; ++u) {                                                   \
                const float *q = acc + ((long long)u * M_r + r) * N;                             \
                for (int j = 0; j < N; ++j) a[j] = a[j] + q[j];                                  \
            }                                                                                    \
            if (row < M) memcpy(C + (long long)row * ldc, a, sizeof(float) * (size_t)N);         \
        }                                                                                        \
    }

#define SRT_CSR_ROW(NAME, ATTR)                                                                  \
    ATTR static void NAME(int N, const long long *rowptr, const int *colind, const float *vals,  \
                          const float *B, long long ldb, float *crow, int i) {                   \
        memset(crow, 0, sizeof(float) * (size_t)N);                                              \
        for (long long q = rowptr[i]; q < rowptr[i + 1]; ++q) {                                  \
            const float *b = B + (long long)colind[q] * ldb;                                     \
            float av = vals[q];                                                                  \
            for (int j = 0; j < N; ++j) crow[j] = fmaf(av, b[j], crow[j]);                       \
        }                                                                                        \
    }

SRT_EXEC_PANEL(exec_panel_avx512, __attribute__((target("avx512f,avx512vl,fma"))))
SRT_EXEC_PANEL(exec_panel_avx2, __attribute__((target("avx2,fma"))))
SRT_EXEC_PANEL(exec_panel_base, )
SRT_CSR_ROW(csr_row_avx512, __attribute__((target("avx512f,avx512vl,fma"))))
SRT_CSR_ROW(csr_row_avx2, __attribute__((target("avx2,fma"))))
SRT_CSR_ROW(csr_row_base, )

typedef void (*exec_panel_fn)(int, int, int, int, int, const int *, const int *, const int *,
                              const float *, const float *, long long, float *, long long, int, int,
                              int, long long, float *);
typedef void (*csr_row_fn)(int, const long long *, const int *, const float *, const float *,
                           long long, float *, int);

static int cpu_level(void) {
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512vl") &&
        __builtin_cpu_supports("fma"))
        return 2;
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return 1;
    return 0;
}

/* Which inner-loop variant runs on this host: 2 AVX-512, 1 AVX2+FMA, 0 base. */
SRT_ORACLE_API int srt_oracle_isa(void) { return cpu_level(); }

/* kp[p] = number of chunks panel p is split into (NULL: one chunk each).
 * KT = K tiles per panel: counts[(p*KT + t)*NK + id] holds tile t's counts and
 * the panel's packed columns are its tiles' columns one tile after another
 * (the K-tiled packing of count_nkerns.cpp:52-126 with load_c chaining). */
SRT_ORACLE_API
int srt_oracle_exec_kt(int M, int N, int M_r, int NK, int KT, const int *id_pattern,
                       const int *counts, const long long *col_ptr, const long long *val_ptr,
                       const int *cols, const float *vals, const float *B, long long ldb,
                       float *C, long long ldc, int load_c, const int *kp, int nthreads) {
    int P = (M + M_r - 1) / M_r;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    int kmax = 1;
    for (int p = 0; kp && p < P; ++p) {
        if (kp[p] < 1) return -2;
        if (kp[p] > kmax) kmax = kp[p];
    }
    int lvl = cpu_level();
    exec_panel_fn fn = lvl == 2 ? exec_panel_avx512 : lvl == 1 ? exec_panel_avx2 : exec_panel_base;
    int err = 0;
#pragma omp parallel num_threads(nthreads)
    {
        size_t bytes = sizeof(float) * (size_t)kmax * (size_t)M_r * (size_t)(N > 0 ? N : 1);
        float *acc = (float *)aligned_alloc(64, (bytes + 63) / 64 * 64);
        if (!acc) {
#pragma omp atomic write
            err = 1;
        } else {
#pragma omp for schedule(static)
            for (int p = 0; p < P; ++p)
                fn(M, N, M_r, NK, KT, id_pattern, counts + (long long)p * KT * NK, cols + col_ptr[p],
                   vals + val_ptr[p], B, ldb, C, ldc, load_c, p, kp ? kp[p] : 1,
                   col_ptr[p + 1] - col_ptr[p], acc);
            free(acc);
        }
    }
    return err ? -1 : 0;
}

SRT_ORACLE_API
int srt_oracle_exec(int M, int N, int M_r, int NK, const int *id_pattern,
                    const int *counts, const long long *col_ptr, const long long *val_ptr,
                    const int *cols, const float *vals, const float *B, long long ldb,
                    float *C, long long ldc, int load_c, const int *kp, int nthreads) {
    return srt_oracle_exec_kt(M, N, M_r, NK, 1, id_pattern, counts, col_ptr, val_ptr, cols, vals, B, ldb,
                              C, ldc, load_c, kp, nthreads);
}

SRT_ORACLE_API
int srt_oracle_csr(int M, int N, const long long *rowptr, const int *colind, const float *vals,
                   const float *B, long long ldb, float *C, long long ldc, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    int lvl = cpu_level();
    csr_row_fn fn = lvl == 2 ? csr_row_avx512 : lvl == 1 ? csr_row_avx2 : csr_row_base;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int i = 0; i < M; ++i) fn(N, rowptr, colind, vals, B, ldb, C + (long long)i * ldc, i);
    return 0;
}

/* Float64 dense product, for the FP64 oracle at sizes numpy handles slowly. */
SRT_ORACLE_API
int srt_oracle_dense_f64(int M, int K, int N, const float *A, long long lda, const float *B,
                         long long ldb, double *C, long long ldc, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int i = 0; i < M; ++i) {
        double *c = C + (long long)i * ldc;
        for (int j = 0; j < N; ++j) c[j] = 0.0;
        for (int k = 0; k < K; ++k) {
            double a = A[(long long)i * lda + k];
            if (a == 0.0) continue;
            const float *b = B + (long long)k * ldb;
            for (int j = 0; j < N; ++j) c[j] += a * (double)b[j];
        }
    }
    return 0;
}
