// C ABI of the precision form (16-row MFMA panels, BF16 / FP8-e4m3).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "srt_internal.h"

namespace srt {
namespace {

template <typename T>
int dalloc(T **p, size_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1));
    if (e != hipSuccess) return fail(SRT_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return SRT_OK;
}

#define MF_TRY(expr)         \
    do {                     \
        int _rc = (expr);    \
        if (_rc) return _rc; \
    } while (0)

}  // namespace
}  // namespace srt

using namespace srt;

struct srt_mfma_plan {
    int device = -1;
    int dtype = 0;  // SRT_DT_BF16 or SRT_DT_FP8
    int M = 0, K = 0, P = 0, KB = 0;
    int64_t nnz = 0, nblk = 0, nz_blocks = 0;
    bool compact = false;  // some block absent: executors look blocks up in `slot`
    uint8_t *present = nullptr;
    int32_t *slot = nullptr;
    int32_t *blk_kb = nullptr;
    void *blocks = nullptr;
    float *rscale = nullptr;
    float *work = nullptr;
    size_t work_floats = 0;
    ~srt_mfma_plan() {
        (void)hipFree(present);
        (void)hipFree(slot);
        (void)hipFree(blk_kb);
        (void)hipFree(blocks);
        (void)hipFree(rscale);
        (void)hipFree(work);
    }
};

static int elt_bytes(int dtype) { return dtype == SRT_DT_FP8 ? 1 : 2; }

// Scratch of the allocation-free entry point (srt_mfma_plan_execute_noalloc):
// the K-split partial tiles, ks * M * n floats at most SRT_NOALLOC_SCRATCH_FLOATS.
// Every GPU the library runs on holds its own copy until the process ends.
// Host code never passes `scratch` itself to a kernel: the address of this
// device's copy comes from hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) float scratch[SRT_NOALLOC_SCRATCH_FLOATS];

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
static int noalloc_scratch(float **out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static float *ptrs[kMaxDevices] = {};
    int dev = -1;
    SRT_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(SRT_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(mu);
    if (!ptrs[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(SRT_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        ptrs[dev] = static_cast<float *>(ptr);
    }
    *out = ptrs[dev];
    return SRT_OK;
}

extern "C" {

int srt_mfma_plan_create(int m, int k, const float *Lx, const void *Lp, const void *Li, int index_bytes,
                         int dtype, srt_stream_t stream, srt_mfma_plan **out) {
    if (!out) return fail(SRT_ERR_INVALID, "null out");
    *out = nullptr;
    if (m < 0 || k < 0) return fail(SRT_ERR_INVALID, "negative dimension");
    if (dtype != SRT_DT_BF16 && dtype != SRT_DT_FP8) return fail(SRT_ERR_INVALID, "dtype must be BF16 or FP8");
    if (index_bytes != 4 && index_bytes != 8) return fail(SRT_ERR_INVALID, "index_bytes must be 4 or 8");
    if (m > 0 && !Lp) return fail(SRT_ERR_INVALID, "null Lp");
    hipStream_t s = (hipStream_t)stream;
    int64_t nnz = 0;
    if (m > 0) {
        // Row pointers are checked on the host (start 0, non-decreasing), so
        // every device loop over a row stays inside the nonzero arrays.
        std::vector<int64_t> lp((size_t)m + 1);
        if (index_bytes == 4) {
            std::vector<int32_t> l32((size_t)m + 1);
            SRT_HIP_TRY(hipMemcpyAsync(l32.data(), Lp, 4 * ((size_t)m + 1), hipMemcpyDeviceToHost, s));
            SRT_HIP_TRY(hipStreamSynchronize(s));
            std::copy(l32.begin(), l32.end(), lp.begin());
        } else {
            SRT_HIP_TRY(hipMemcpyAsync(lp.data(), Lp, 8 * ((size_t)m + 1), hipMemcpyDeviceToHost, s));
            SRT_HIP_TRY(hipStreamSynchronize(s));
        }
        if (lp[0] != 0) return fail(SRT_ERR_INVALID, "Lp[0] must be 0");
        for (int i = 0; i < m; ++i)
            if (lp[i + 1] < lp[i]) return fail(SRT_ERR_INVALID, "Lp must be non-decreasing");
        nnz = lp[m];
        if (nnz > 0 && (!Li || !Lx)) return fail(SRT_ERR_INVALID, "null Li or Lx");
    }
    std::unique_ptr<srt_mfma_plan> p(new srt_mfma_plan());
    SRT_HIP_TRY(hipGetDevice(&p->device));
    p->dtype = dtype;
    p->M = m;
    p->K = k;
    p->P = (m + 15) / 16;
    p->KB = (k + 31) / 32;
    p->nnz = nnz;
    const size_t cells = (size_t)p->P * p->KB;
    MF_TRY(dalloc(&p->present, cells));
    MF_TRY(dalloc(&p->slot, cells));
    MF_TRY(dalloc(&p->rscale, (size_t)m));
    int64_t *nblk = nullptr, *blk_ptr = nullptr;
    int *bad = nullptr;
    MF_TRY(dalloc(&nblk, (size_t)p->P));
    MF_TRY(dalloc(&blk_ptr, (size_t)p->P + 1));
    MF_TRY(dalloc(&bad, 1));
    struct Free {
        void *a, *b, *c;
        ~Free() {
            (void)hipFree(a);
            (void)hipFree(b);
            (void)hipFree(c);
        }
    } fr{nblk, blk_ptr, bad};
    if (m > 0) {
        SRT_HIP_TRY(hipMemsetAsync(p->present, 0, cells, s));
        SRT_HIP_TRY(hipMemsetAsync(bad, 0, sizeof(int), s));
        SRT_HIP_TRY(launch_mfma_pack(m, k, Lp, Li, index_bytes, Lx, dtype == SRT_DT_FP8, p->P, p->KB, p->present,
                                     nblk, bad, s));
        SRT_HIP_TRY(launch_exclusive_scan(p->P, nblk, blk_ptr, s));
        int hbad = 0;
        SRT_HIP_TRY(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
        SRT_HIP_TRY(hipMemcpyAsync(&p->nz_blocks, blk_ptr + p->P, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        SRT_HIP_TRY(hipStreamSynchronize(s));
        if (hbad) return fail(SRT_ERR_INVALID, "column index outside [0, k)");
    }
    // Only the blocks holding a nonzero are stored.  At least one block's
    // bytes are allocated (and zeroed): the executor's operand load for an
    // absent block reads stored block 0 before zeroing the operand.
    p->nblk = p->nz_blocks;
    p->compact = p->nblk < (int64_t)cells;
    if (p->compact && cells * 4 > (size_t)INT32_MAX)
        return fail(SRT_ERR_UNSUPPORTED, "slot table of " + std::to_string(cells) + " blocks exceeds 2 GiB");
    MF_TRY(dalloc(&p->blk_kb, (size_t)p->nblk));
    const size_t blk_bytes = std::max<size_t>((size_t)p->nblk, 1) * 512 * elt_bytes(dtype);
    SRT_HIP_TRY(hipMalloc(&p->blocks, blk_bytes));
    if (m > 0) {
        SRT_HIP_TRY(hipMemsetAsync(p->blocks, 0, blk_bytes, s));
        SRT_HIP_TRY(launch_mfma_slots(p->P, p->KB, p->present, blk_ptr, p->slot, p->blk_kb, s));
        SRT_HIP_TRY(launch_mfma_fill(m, k, p->KB, Lp, Li, index_bytes, Lx, dtype == SRT_DT_FP8, p->slot,
                                     p->rscale, p->blocks, s));
        SRT_HIP_TRY(hipStreamSynchronize(s));
    }
    *out = p.release();
    return SRT_OK;
}

int srt_mfma_plan_reserve(srt_mfma_plan *p, int n_max, int ksplit_max) {
    if (!p) return fail(SRT_ERR_INVALID, "null plan");
    if (n_max < 0 || ksplit_max < 1) return fail(SRT_ERR_INVALID, "bad reservation");
    const size_t need = ksplit_max > 1 ? (size_t)ksplit_max * p->M * n_max : 0;
    if (need <= p->work_floats) return SRT_OK;
    (void)hipFree(p->work);
    p->work = nullptr;
    p->work_floats = 0;
    MF_TRY(dalloc(&p->work, need));
    p->work_floats = need;
    return SRT_OK;
}

int srt_mfma_convert_b(int dtype, int64_t rows, int cols, const float *B, int64_t ldb, void *out,
                       int64_t ld_out, float *col_scales, srt_stream_t stream) {
    if (dtype != SRT_DT_BF16 && dtype != SRT_DT_FP8) return fail(SRT_ERR_INVALID, "dtype must be BF16 or FP8");
    if (rows < 0 || cols < 0 || ldb < cols) return fail(SRT_ERR_INVALID, "bad shape");
    if (ld_out < (rows + 31) / 32 * 32) return fail(SRT_ERR_INVALID, "ld_out must be >= K rounded up to 32");
    if (rows * (int64_t)cols == 0) return SRT_OK;
    if (!B || !out || (dtype == SRT_DT_FP8 && !col_scales)) return fail(SRT_ERR_INVALID, "null pointer");
    SRT_HIP_TRY(launch_b_convert(rows, cols, B, ldb, dtype == SRT_DT_FP8, col_scales, out, ld_out,
                                 (hipStream_t)stream));
    return SRT_OK;
}

// Shared body of the two execute entry points: `work` holds the K-split
// partial tiles (work_floats of them).
static int mfma_execute(const srt_mfma_plan *p, float *C, int64_t ldc, const void *B, int64_t ldb,
                        const float *col_scales, int n, int shape, int ksplit, float *work, size_t work_floats,
                        const char *short_work, hipStream_t stream) {
    if (!p) return fail(SRT_ERR_INVALID, "null plan");
    if (n < 0) return fail(SRT_ERR_INVALID, "n must be >= 0");
    if (n == 0 || p->M == 0) return SRT_OK;
    if (!C || !B) return fail(SRT_ERR_INVALID, "null B or C");
    if (ldc < n) return fail(SRT_ERR_INVALID, "ldc must be >= n");
    if (p->dtype == SRT_DT_FP8 && !col_scales) return fail(SRT_ERR_INVALID, "FP8 needs the column scales of B");
    int dev = -1;
    SRT_HIP_TRY(hipGetDevice(&dev));
    if (dev != p->device) return fail(SRT_ERR_INVALID, "plan belongs to another device");
    int nt = 0, wp = 0, pw = 0;
    if (mfma_shape_tile(shape, &nt, &wp, &pw)) return fail(SRT_ERR_INVALID, "unknown launch shape");
    if (ldb < (int64_t)p->KB * 32) return fail(SRT_ERR_INVALID, "B (transposed, N x ldb) needs ldb >= K rounded up to 32");
    if ((uintptr_t)B % 16 != 0 || (ldb * elt_bytes(p->dtype)) % 16 != 0)
        return fail(SRT_ERR_INVALID, "B (transposed) rows must be 16-byte aligned (as srt_mfma_convert_b writes them)");
    const int ks = std::max(1, std::min(ksplit, std::max(p->KB, 1)));
    if (ks > 1 && (size_t)ks * p->M * n > work_floats) return fail(SRT_ERR_INVALID, short_work);
    MfmaLaunch a;
    a.fp8 = p->dtype == SRT_DT_FP8;
    a.shape = shape;
    a.M = p->M;
    a.N = n;
    a.K = p->K;
    a.KB = p->KB;
    a.P = p->P;
    a.B = B;
    a.ldb = ldb;
    a.blocks = p->blocks;
    a.slot = p->compact ? p->slot : nullptr;
    a.rscale = p->rscale;
    a.cscale = col_scales;
    a.C = C;
    a.ldc = ldc;
    a.ks = ks;
    a.work = work;
    SRT_HIP_TRY(launch_mfma_exec(a, stream));
    return SRT_OK;
}

int srt_mfma_plan_execute(const srt_mfma_plan *p, float *C, int64_t ldc, const void *B, int64_t ldb,
                          const float *col_scales, int n, int shape, int ksplit, srt_stream_t stream) {
    if (!p) return fail(SRT_ERR_INVALID, "null plan");
    return mfma_execute(p, C, ldc, B, ldb, col_scales, n, shape, ksplit, p->work, p->work_floats,
                        "K split needs a larger reservation (srt_mfma_plan_reserve)", (hipStream_t)stream);
}

int srt_mfma_plan_execute_noalloc(const srt_mfma_plan *p, float *C, int64_t ldc, const void *B, int64_t ldb,
                                  const float *col_scales, int n, int shape, int ksplit, srt_stream_t stream) {
    if (!p) return fail(SRT_ERR_INVALID, "null plan");
    if (n < 0) return fail(SRT_ERR_INVALID, "n must be >= 0");
    const int ks = std::max(1, std::min(ksplit, std::max(p->KB, 1)));
    if (ks > 1 && (size_t)ks * p->M * (size_t)n > (size_t)SRT_NOALLOC_SCRATCH_FLOATS)
        return fail(SRT_ERR_INVALID, "ksplit * m * n = " + std::to_string((size_t)ks * p->M * (size_t)n) +
                                         " exceeds the scratch maximum of " +
                                         std::to_string((size_t)SRT_NOALLOC_SCRATCH_FLOATS) + " floats");
    float *work = nullptr;
    if (ks > 1) MF_TRY(noalloc_scratch(&work));
    return mfma_execute(p, C, ldc, B, ldb, col_scales, n, shape, ksplit, work, SRT_NOALLOC_SCRATCH_FLOATS,
                        "K split exceeds the scratch maximum", (hipStream_t)stream);
}

int srt_mfma_plan_stats_get(const srt_mfma_plan *p, srt_mfma_stats *o) {
    if (!p || !o) return fail(SRT_ERR_INVALID, "null argument");
    o->M = p->M;
    o->K = p->K;
    o->num_panels = p->P;
    o->k_blocks = p->KB;
    o->dtype = p->dtype;
    o->nnz = p->nnz;
    o->stored_blocks = p->nblk;
    o->nonzero_blocks = p->nz_blocks;
    return SRT_OK;
}

int srt_mfma_plan_row_scales(const srt_mfma_plan *p, float *out) {
    if (!p || !out) return fail(SRT_ERR_INVALID, "null argument");
    if (p->M) SRT_HIP_TRY(hipMemcpy(out, p->rscale, sizeof(float) * p->M, hipMemcpyDeviceToHost));
    return SRT_OK;
}

int srt_mfma_plan_copy_blocks(const srt_mfma_plan *p, int32_t *slot, int32_t *blk_kb, void *blocks) {
    if (!p) return fail(SRT_ERR_INVALID, "null plan");
    const size_t cells = (size_t)p->P * p->KB;
    if (slot && cells) SRT_HIP_TRY(hipMemcpy(slot, p->slot, sizeof(int32_t) * cells, hipMemcpyDeviceToHost));
    if (blk_kb && p->nblk) SRT_HIP_TRY(hipMemcpy(blk_kb, p->blk_kb, sizeof(int32_t) * p->nblk, hipMemcpyDeviceToHost));
    if (blocks && p->nblk)
        SRT_HIP_TRY(hipMemcpy(blocks, p->blocks, (size_t)p->nblk * 512 * elt_bytes(p->dtype), hipMemcpyDeviceToHost));
    return SRT_OK;
}

void srt_mfma_plan_destroy(srt_mfma_plan *p) { delete p; }

static_assert(sizeof(srt_mfma_problem) == sizeof(MfmaProblem), "problem layout");
static_assert(sizeof(srt_bprep_problem) == sizeof(BPrepProblem), "problem layout");

int srt_mfma_problem_fill(const srt_mfma_plan *p, const void *Bt, int64_t ldb, const float *col_scales, float *C,
                          int64_t ldc, int n, int tile_begin, srt_mfma_problem *out) {
    if (!p || !out) return fail(SRT_ERR_INVALID, "null argument");
    if (n < 0 || ldc < n || ldb < (int64_t)p->KB * 32) return fail(SRT_ERR_INVALID, "bad problem shape");
    if (n > 0 && (!Bt || !C)) return fail(SRT_ERR_INVALID, "null B or C");
    if (p->dtype == SRT_DT_FP8 && n > 0 && !col_scales) return fail(SRT_ERR_INVALID, "FP8 needs column scales");
    if (n > 0 && ((uintptr_t)Bt % 16 != 0 || (ldb * elt_bytes(p->dtype)) % 16 != 0))
        return fail(SRT_ERR_INVALID, "B (transposed) rows must be 16-byte aligned");
    out->blocks = p->blocks;
    out->Bt = Bt;
    out->C = C;
    out->rscale = p->rscale;
    out->cscale = col_scales;
    out->ldb = ldb;
    out->ldc = ldc;
    out->M = p->M;
    out->N = n;
    out->KB = p->KB;
    out->P = p->P;
    out->tile_begin = tile_begin;
    const int eb = elt_bytes(p->dtype);
    out->vec_ok = ((uintptr_t)Bt % 16 == 0 && (ldb * eb) % 16 == 0) ? 1 : 0;
    out->slot = p->compact ? p->slot : nullptr;
    return SRT_OK;
}

int srt_mfma_problem_tiles(const srt_mfma_problem *pr, int shape) {
    int nt = 0, wp = 0, pw = 0;
    if (!pr || mfma_shape_tile(shape, &nt, &wp, &pw)) return -1;
    if (pr->N <= 0 || pr->M <= 0) return 0;
    return ((pr->P + wp * pw - 1) / (wp * pw)) * ((pr->N + nt - 1) / nt);
}

int srt_mfma_grouped_execute(int dtype, int shape, const srt_mfma_problem *d_probs, const int *d_hdr, int grid,
                             srt_stream_t stream) {
    if (dtype != SRT_DT_BF16 && dtype != SRT_DT_FP8) return fail(SRT_ERR_INVALID, "dtype must be BF16 or FP8");
    int nt = 0, wp = 0, pw = 0;
    if (mfma_shape_tile(shape, &nt, &wp, &pw)) return fail(SRT_ERR_INVALID, "unknown launch shape");
    if (grid < 1) return fail(SRT_ERR_INVALID, "grid must be >= 1");
    if (!d_probs || !d_hdr) return fail(SRT_ERR_INVALID, "null problem table");
    SRT_HIP_TRY(launch_mfma_grouped(dtype == SRT_DT_FP8, shape, reinterpret_cast<const MfmaProblem *>(d_probs),
                                    d_hdr, grid, (hipStream_t)stream));
    return SRT_OK;
}

int srt_bprep_problem_tiles(const srt_bprep_problem *pr) {
    if (!pr) return -1;
    if (pr->N <= 0 || pr->ldk <= 0) return 0;
    return (int)(((pr->ldk + 31) / 32) * ((pr->N + 31) / 32));
}

int srt_widen_problem_tiles(const srt_bprep_problem *pr) {
    if (!pr) return -1;
    const int64_t w = pr->pad_to > 0 ? pr->pad_to : pr->ldk;
    if (pr->K <= 0 || w <= 0) return 0;
    return (int)(((int64_t)pr->K * w + 1023) / 1024);
}

int srt_exec_grouped_b_widen(const srt_bprep_problem *d_probs, const int *d_hdr, int grid, srt_stream_t stream) {
    if (grid < 1) return fail(SRT_ERR_INVALID, "grid must be >= 1");
    if (!d_probs || !d_hdr) return fail(SRT_ERR_INVALID, "null table");
    SRT_HIP_TRY(launch_grouped_b_widen(reinterpret_cast<const BPrepProblem *>(d_probs), d_hdr, grid,
                                       (hipStream_t)stream));
    return SRT_OK;
}

int srt_mfma_grouped_b_prep(int dtype, const srt_bprep_problem *d_probs, const int *d_hdr, int grid,
                            srt_stream_t stream) {
    if (dtype != SRT_DT_BF16 && dtype != SRT_DT_FP8) return fail(SRT_ERR_INVALID, "dtype must be BF16 or FP8");
    if (grid < 1) return fail(SRT_ERR_INVALID, "grid must be >= 1");
    if (!d_probs || !d_hdr) return fail(SRT_ERR_INVALID, "null problem table");
    SRT_HIP_TRY(launch_grouped_b_prep(dtype == SRT_DT_FP8, reinterpret_cast<const BPrepProblem *>(d_probs), d_hdr,
                                      grid, (hipStream_t)stream));
    return SRT_OK;
}

}  // extern "C"
