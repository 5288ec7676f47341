// Internal declarations shared by the libsrt translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "srt.h"

namespace srt {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define SRT_HIP_TRY(expr)                                                              \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return ::srt::fail(SRT_ERR_HIP, std::string(#expr " failed: ") +          \
                                                hipGetErrorString(_e));                \
    } while (0)

// ------------------------------------------------------ generated executors
// Arguments of one executor launch over a packed matrix.
// Batch records (the executor's instruction stream, built per schedule):
// every wave slot owns a run of records, one per batch of U packed columns of
// one kernel id, then a terminator:
//   rec[0] = id * U + n - 1 (n columns; -1 terminates), rec[1] = index of the
//   batch's first value in vals, rec[2 .. 2 + U) = its B rows (slots past n
//   repeat a valid row; they are loaded and never used).
constexpr int rec_stride(int U) { return U <= 2 ? 4 : U <= 6 ? 8 : 16; }

struct ExecLaunch {
    const int32_t *prog;     // batch records (see rec_stride)
    const int32_t *desc;     // [n_wblocks * waves_per_block] pairs of int4 wave descriptors
    int n_wblocks;           // workgroups per N tile
    int reduce;              // some panel runs on more than one wave
    const float *vals;
    const float *B;
    uint32_t b_bytes;  // bytes of B addressable through the buffer resource
    float *C;
    int64_t ldb, ldc;
    int M, N, K, num_panels;
    int load_c;
    int vec_width;        // V
    int waves_per_block;  // panels per workgroup
    int xcd_remap;
    int unroll;
};

typedef hipError_t (*ExecLaunchFn)(const ExecLaunch &, hipStream_t);

// One problem of a grouped executor launch (device table entry).
struct ExecProblem {
    const int32_t *prog;
    const int32_t *desc;
    const float *vals;
    const float *B;
    float *C;
    int64_t ldc;
    uint32_t b_bytes, ldb;
    int M, N, K, n_wblocks, reduce, load_c;
    int block_begin;  // first logical workgroup of this problem
    int pad;
};
typedef hipError_t (*ExecGroupedFn)(const ExecProblem *d_probs, int nprob, int total_blocks, int waves,
                                    size_t lds, int xcd_remap, hipStream_t s);
// Persistent grouped launch: problem count and workgroup total read from d_hdr.
typedef hipError_t (*ExecPersistentFn)(const ExecProblem *d_probs, const int *d_hdr, int grid, int waves,
                                       size_t lds, hipStream_t s);

// One generated executor (one pattern set).  launch[s][v][u]: s = 0 reads B
// rows from global memory, s = 1 stages the B tile in LDS; V = 1,2,4 and
// unroll = 1,2,4,8 (index log2); a null entry is a variant not generated.
struct ExecutorDesc {
    const char *id;
    int M_r;
    int num_patterns;
    const int32_t *patterns;  // id order
    ExecLaunchFn launch[2][3][4];
    ExecGroupedFn grouped[3];  // global-memory B path, unroll 4, V = 1,2,4 (null: not generated)
    ExecPersistentFn persistent[3];  // the same variants, persistent over a device problem table
};

// Largest dynamic LDS allocation of one workgroup on the current device.
size_t lds_limit();

void register_executor(const ExecutorDesc *d);
const std::vector<const ExecutorDesc *> &executors();
const ExecutorDesc *find_executor(int M_r, const std::vector<int32_t> &patterns);
const ExecutorDesc *find_executor_by_id(const char *id);

// ---------------------------------------------------------------- packing
struct DeviceLut {
    int M_r = 0;
    int num_patterns = 0;
    int32_t *lut_ptr = nullptr;       // [2^M_r + 1]  code -> range in lut_ids
    int32_t *lut_ids = nullptr;       // kernel ids in mapping-list order
    int32_t *id_pattern = nullptr;    // [NK] id -> bitmask
    int32_t *pattern_id = nullptr;    // [2^M_r] bitmask -> id or -1
    int device = -1;
};

// Zero entries appended to the packed column / value arrays: the executor
// reads column indices and values in fixed-size batches that may run past the
// last panel's data (at most 2 * 8 indices and 8 * 16 + 32 values).
constexpr int SRT_COL_PAD = 32;
constexpr int SRT_VAL_PAD = 256;
// Zero (id, count) pairs after the last group: the executor reads the group
// list two entries ahead.
constexpr int SRT_GROUP_PAD = 4;

// Packed matrix in device memory (flat arrays over all panels).
struct Packed {
    int M = 0, K = 0, M_r = 0, P = 0, NK = 0;
    int KT = 1;                  // K tiles per panel (K_c tiling), each packed on its own
    int K_c = 0;                 // columns per K tile (K when untiled)
    int32_t *counts = nullptr;   // [P * KT * NK]  nkern_counts of every (panel, K tile)
    int64_t *col_ptr = nullptr;  // [P + 1]
    int64_t *val_ptr = nullptr;  // [P + 1]
    int32_t *cols = nullptr;
    float *vals = nullptr;
    int32_t *groups = nullptr;   // (id | popcount << 16, count) of non-empty ids per panel
    int64_t *grp_ptr = nullptr;  // [P + 1]
    int32_t *id_pattern = nullptr;  // [NK] kernel pattern of each id
    std::vector<int64_t> h_col_ptr;  // host copy of col_ptr (scheduling)
    std::vector<uint8_t> h_skip;     // panels the schedule leaves out (their C rows are
                                     // written by someone else: a hybrid plan's dense panels)
    int64_t total_cols = 0, total_vals = 0, total_groups = 0, padding = 0, nnz = 0;
    void release();
    size_t bytes() const;
};

// Load-balanced launch schedule over a Packed matrix: panel p runs on
// kp[p] = min(nsplit, max(1, ceil(ncols_p / chunk_cols))) waves; panels are
// laid into workgroups of waves_per_block wave slots in panel order.
struct Schedule {
    int32_t *desc = nullptr;  // device: four int4 per wave slot (see exec_kernel.h)
    int32_t *prog = nullptr;  // device: batch records of every wave slot
    int unroll = 0;           // U the records are cut for
    int n_wblocks = 0;
    int waves_per_block = 0;
    int nsplit = 0;
    int chunk_cols = 0;
    int layout = 1;           // realised tile config layout (1 panel order, 2 balanced)
    int n_tiles = 1;          // N tiles the balanced layout was sized for
    int reduce = 0;
    size_t bytes = 0;         // device bytes of desc + prog
    std::vector<int32_t> kp;  // per panel
    void release();
};

// A schedule no longer current, kept with the realised config it was built for.
struct RetiredSchedule {
    srt_tile_config cfg;
    Schedule sch;
};

// Kernel launchers (packer.hip).
// CSR packing in O(nnz) memory (packer.hip): cell keys, radix sort, cells,
// segment bounds, per-segment counts and fill.  The two hipcub-backed steps
// follow hipcub's protocol: called with temp == nullptr they only store the
// temporary storage they need in *temp_bytes.
hipError_t launch_cell_keys(int m, int64_t k, int M_r, const void *Lp, const void *Li, int index_bytes,
                            uint64_t *keys, uint32_t *idx, uint8_t *ent_row, int *bad, hipStream_t s);
hipError_t sort_cell_keys(int64_t n, int end_bit, uint64_t *keys, uint64_t *keys_alt, uint32_t *idx,
                          uint32_t *idx_alt, void *temp, size_t *temp_bytes, bool *sorted_in_alt, hipStream_t s);
hipError_t launch_cell_heads(int64_t n, const uint64_t *keys, const uint32_t *idx, const uint8_t *ent_row,
                             uint32_t *head, int *bad, hipStream_t s);
hipError_t scan_cell_heads(int64_t n, const uint32_t *head, uint32_t *cell_of, void *temp, size_t *temp_bytes,
                           hipStream_t s);
hipError_t launch_cell_fill(int64_t n, const uint64_t *keys, const uint32_t *idx, const uint8_t *ent_row,
                            const uint32_t *head, const uint32_t *cell_of, uint64_t *cell_key, uint32_t *cell_code,
                            int64_t *cell_first, hipStream_t s);
hipError_t launch_segment_bounds(int Q, int KT, int64_t k, int Kc, int64_t ncells, const uint64_t *cell_key,
                                 int64_t *seg_cells, hipStream_t s);
hipError_t launch_count_cells(int Q, const DeviceLut &lut, const int64_t *seg_cells, const uint32_t *cell_code,
                              int32_t *counts, int64_t *ncols, int64_t *nvals, hipStream_t s);
hipError_t launch_fill_cells(int Q, int64_t k, int KT, const DeviceLut &lut, const int64_t *seg_cells,
                             const uint64_t *cell_key, const uint32_t *cell_code, const int64_t *cell_first,
                             const uint32_t *ent_idx, const uint8_t *ent_row, const float *Lx,
                             const int64_t *col_ptr, const int64_t *val_ptr, const int32_t *counts,
                             int32_t *cols, float *vals, int64_t *padding, hipStream_t s);
hipError_t launch_exclusive_scan(int n, const int64_t *in, int64_t *out, hipStream_t s);
// Non-empty group lists from counts: ngroups[p], then (after a scan into
// grp_ptr) the (id | popcount(pattern) << 16, count) pairs.
hipError_t launch_group_count(int P, int NK, const int32_t *counts, int64_t *ngroups,
                              hipStream_t s);
hipError_t launch_group_fill(int P, int NK, const int32_t *counts, const int64_t *grp_ptr,
                             const int32_t *id_pattern, int32_t *groups, hipStream_t s);
// Wave descriptors from (panel, j, kp) triples (slots[3 * i]; panel -1 = idle).
hipError_t launch_wave_prog(int nslots, int U, const int32_t *groups, const int64_t *grp_ptr,
                            const int64_t *col_ptr, const int64_t *val_ptr, const int32_t *cols,
                            const int64_t *rec_start, int32_t *desc, int32_t *prog, hipStream_t s);
hipError_t launch_wave_desc(int nslots, const int32_t *slots, const int32_t *groups,
                            const int64_t *grp_ptr, const int64_t *col_ptr,
                            const int32_t *id_pattern, int32_t *desc, hipStream_t s);
// Entry-list variants (SOPTile path): one panel, explicit (pattern, column) entries.
// bad |= 1 for a pattern outside the set, |= 2 for a column outside [0, K)
// (col_indices may be null to skip the column check).
hipError_t launch_encode_entries(int n, const DeviceLut &lut, const int32_t *pat_codes,
                                 const int32_t *col_indices, int K, int32_t *ids, int32_t *bad,
                                 hipStream_t s);
hipError_t launch_count_nonzero(const float *x, int64_t n, int64_t *out, hipStream_t s);
hipError_t launch_count_entries(int n, const DeviceLut &lut, const int32_t *ids,
                                int32_t *counts, int64_t *ncols, int64_t *nvals, hipStream_t s);
hipError_t launch_fill_entries(int n, int col_offset, int K, const DeviceLut &lut,
                               const int32_t *ids, const int32_t *col_indices,
                               const float *dense_panel, const int32_t *counts, int32_t *cols,
                               float *vals, int64_t *padding, hipStream_t s);
hipError_t launch_stable_order(int n, int NK, const int32_t *ids, const int32_t *counts,
                               int32_t *encoded_sorted, int32_t *order, hipStream_t s);

// Precision form on MFMA (mfma_panel.hip): 16-row panels, 32-column blocks.
struct MfmaLaunch {
    bool fp8;
    int shape;  // (N tile, waves, panels per wave) index, see mfma_shape_tile
    int M, N, K, KB, P;
    const void *B;  // Bt: bf16 or e4m3, N x ldb (k contiguous), ldb >= KB * 32
    int64_t ldb;
    const void *blocks;     // stored blocks in MFMA A-operand layout
    const int32_t *slot;    // [P][KB] stored index of each block or -1; null: block (p, kb) at p * KB + kb
    const float *rscale;    // FP8 row scales (M)
    const float *cscale;    // FP8 column scales of B (N)
    float *C;
    int64_t ldc;
    int ks;       // K split
    float *work;  // ks * M * N partial tiles when ks > 1
};
hipError_t launch_mfma_pack(int m, int k, const void *Lp, const void *Li, int index_bytes,
                            const float *Lx, bool fp8, int P, int KB, uint8_t *present,
                            int64_t *nblk, int *bad, hipStream_t s);
hipError_t launch_mfma_slots(int P, int KB, const uint8_t *present, const int64_t *blk_ptr,
                             int32_t *slot, int32_t *blk_kb, hipStream_t s);
hipError_t launch_mfma_fill(int m, int k, int KB, const void *Lp, const void *Li, int index_bytes,
                            const float *Lx, bool fp8, const int32_t *slot, float *rscale,
                            void *blocks, hipStream_t s);
hipError_t launch_b_convert(int64_t rows, int cols, const float *src, int64_t ld_src, bool fp8,
                            float *cscale, void *dst, int64_t ldk, hipStream_t s);
hipError_t launch_mfma_exec(const MfmaLaunch &a, hipStream_t s);
// One problem of a grouped launch (layout shared with include/srt.h's
// srt_mfma_problem).
struct MfmaProblem {
    const void *blocks;
    const void *Bt;
    float *C;
    const float *rscale;
    const float *cscale;
    int64_t ldb, ldc;
    int M, N, KB, P;
    int tile_begin;  // first global tile of this problem
    int vec_ok;      // Bt rows 16-byte aligned
    const int32_t *slot;  // [P][KB] stored index of each block or -1; null: block (p, kb) at p * KB + kb
};
struct BPrepProblem {
    const void *src;  // fp32 or bf16 (src_bf16), K x N row-major, ld_src
    void *dst;
    const float *cscale;
    int64_t ld_src, ldk;
    int K, N;
    int tile_begin;
    int src_bf16;
    int pad_to;  // widening: columns written per row (0 = ldk)
    int reserved;
};
// hdr (device): {number of problems, total tiles}.
hipError_t launch_mfma_grouped(bool fp8, int shape, const MfmaProblem *probs, const int *hdr, int grid,
                               hipStream_t s);
hipError_t launch_grouped_b_prep(bool fp8, const BPrepProblem *probs, const int *hdr, int grid, hipStream_t s);
hipError_t launch_grouped_b_widen(const BPrepProblem *probs, const int *hdr, int grid, hipStream_t s);
int mfma_shape_tile(int shape, int *nt, int *wp, int *pw);

// Full-pattern 4-row packing on matrix cores (mfma_exec.hip): panel ranges
// [ranges[r], ranges[r + 1]) of r < R, 64-column B slices staged in LDS.
hipError_t launch_mfma4_exec(const int64_t *col_ptr, const int32_t *cols, const float *vals,
                             const int32_t *ranges, int R, const float *B, int64_t ldb, float *C, int64_t ldc,
                             int M, int N, int K, hipStream_t s);
int mfma4_max_k();  // B rows the LDS slice holds

// Dense 16-row panels of hybrid plans on FP32 matrix cores (dense_panels.hip).
hipError_t launch_dense_panels(int ndense, int G, const void *apack, const int32_t *panel_row, const float *B,
                               int64_t ldb, float *C, int64_t ldc, int M, int N, int K, void *ws, size_t ws_bytes,
                               hipStream_t s);
// Workspace a dense-panel launch of width n splits K with (0: no split).
size_t dense_panels_workspace_bytes(int ndense, int G, int n);

// CSR baseline (csr_spmm.hip).
hipError_t launch_csr(int m, int k, int n, int64_t nnz, const void *rowptr, const void *colind,
                      int index_bytes, const float *vals, const float *B, int64_t ldb, float *C,
                      int64_t ldc, int vec_width, hipStream_t s);

}  // namespace srt
