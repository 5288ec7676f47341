/*
 * srt.h - C ABI of the MI355X (gfx950) sparse-register-tiling SpMM library
 * (libsrt.so, built from sparse-register-tiling_amd/csrc).
 *
 * C = A_sparse * B_dense, A packed into M_r-row panels whose columns are
 * grouped by nonzero pattern ("nano-kernel" pattern), B and C dense row-major
 * fp32 in device memory.
 *
 * Conventions
 *   - Every call returns an int status: SRT_OK (0) or an SRT_ERR_* code; the
 *     message of the last failure on the calling thread is srt_last_error().
 *     Nothing in the library calls exit() (the reference exits on errors:
 *     cpp_testbed/src/spmm/nano/SpMM_SOP.h:177, template:300-303).
 *   - Pointers documented "device" are device (HBM) pointers on the current
 *     HIP device; "host" pointers are ordinary host memory.
 *   - Work is enqueued on the caller's stream (srt_stream_t is hipStream_t;
 *     NULL = the default stream).  Calls that must read sizes back (plan
 *     creation, pack_tile) synchronise that stream before returning.
 *   - A plan's packed data is immutable after creation, and concurrent
 *     executes of one plan on different streams are safe (except hybrid
 *     plans, whose dense launch splits K through a plan-owned workspace,
 *     and the allocation-free MFMA entry point, whose scratch is per GPU).  The launch
 *     schedule may change (srt_plan_allocate_executor, srt_tile_execute with
 *     another configuration), but a schedule once built stays alive until
 *     the plan / tile is destroyed, so a graph captured before the change
 *     replays unchanged.  Calls that change the schedule must not run
 *     concurrently with other calls on the same plan.
 *   - CSR inputs: row pointer and column index arrays of index_bytes = 4
 *     (int32) or 8 (int64) bytes each, zero-based, no duplicate entries.
 */
#ifndef SRT_H_
#define SRT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *srt_stream_t; /* == hipStream_t */

enum {
    SRT_OK = 0,
    SRT_ERR_INVALID = 1,     /* bad argument / shape / state */
    SRT_ERR_UNSUPPORTED = 2, /* no executor compiled for a pattern set, etc. */
    SRT_ERR_HIP = 3,         /* a HIP runtime call failed */
    SRT_ERR_IO = 4,          /* mapping file not found / unreadable */
    SRT_ERR_NOMEM = 5        /* device or host allocation failed */
};

/* Zero-pattern sentinel id (tools/codegen/gen_vector_specialized.py:42 value). */
#define SRT_ZERO_PATTERN_ID 32768

/* Maximum panel height (pattern bitmask width) the library accepts. */
#define SRT_MAX_MR 16

const char *srt_last_error(void);
int srt_version(void);

/* ------------------------------------------------------------------------
 * Tiling configuration.  Mirrors sop::TileConfig as set by
 * cpp_testbed/src/spmm/nano/SpMM_SOP.h:107-114 (fields M_c .. runtime_schedule),
 * plus the GPU realisation fields.  get_config returns the realised values.
 * ---------------------------------------------------------------------- */
typedef struct srt_tile_config {
    int M_c;               /* rows per workgroup (M_r * panels per workgroup)        */
    int K_c;               /* K tile: 0 < K_c < K packs every panel as K tiles of K_c
                              columns, one after another (count_nkerns.cpp:52-126), and
                              the chain runs tile by tile as the reference's K-tiled
                              execution with load_c does; otherwise one tile of K      */
    int N_c;               /* columns per wavefront (64 * vec_width)                  */
    int tiling_strategy;   /* reference enum: 0 manual (K_c as given), 1 CAKE (K_c sized so
                              a workgroup's B tile, K_c x 64*vec_width fp32, fits its LDS),
                              2 CAKE + TLB compensation (then K_c = max_tlb_entries *
                              tlb_page_size / (n * 4) when n x K_c fp32 needs more pages,
                              SpMM_GECSB.h:339-348); realised at plan creation          */
    float beta;            /* CAKE tiling beta (recorded; not a BLAS beta)            */
    int sparse_a;          /* recorded                                                */
    int max_tlb_entries;   /* TLB compensation (tiling_strategy 2)                    */
    int tlb_page_size;     /* TLB compensation (tiling_strategy 2), bytes             */
    int runtime_schedule;  /* reference outer schedule id (KNM=0 ...) (recorded)      */
    int vec_width;         /* V in {1,2,4}: fp32 columns per lane; 0 = choose        */
    int waves_per_block;   /* wave slots per workgroup, 1..16; a panel's waves share
                              one workgroup, light panels are packed together;
                              0 = choose                                            */
    int xcd_remap;         /* 1: group workgroups of one N tile on one XCD; -1 = choose */
    int unroll;            /* packed columns in flight per wave: 1,2,4,8; 0 = choose  */
    int nsplit;            /* most waves one panel's packed columns are split over,
                              1..waves_per_block (partial sums are added in chunk
                              order; 1 = the reference's single chain); 0 = choose */
    int chunk_cols;        /* target packed columns per wave: with layout 1 panel p
                              runs on kp = min(nsplit, max(1, ceil(ncols_p / chunk_cols)))
                              waves; with layout 2 it sizes the workgroup count (see
                              layout); 0 = choose                                   */
    int b_staging;         /* B operand path: 1 = copy the workgroup's B tile (K rows x
                              N_c columns) into LDS once and read packed columns
                              from there; 2 = read B rows from global memory per
                              packed column; 0 = choose (global memory)          */
    int layout;            /* wave slots -> workgroups: 1 = panels in order, a panel
                              on kp waves as chunk_cols says; 2 = balanced: panels
                              binned heaviest-first into the least-loaded of G
                              workgroups (G x N tiles = whole rounds of the chip's
                              CUs at about chunk_cols columns per wave), each
                              workgroup's slots then given one by one to its
                              panel with the longest chunk; 0 = choose (2)        */
    int matrix_cores;      /* 1: run the packed columns on matrix cores, one
                              v_mfma_f32_4x4x1_16b_f32 outer product per packed
                              column and 64 C columns (B slice staged in LDS).
                              Needs M_r = 4, a mapping whose only kernel pattern
                              is 0b1111 (every column padded to the full 4-row
                              column, e.g. the "full" mapping 7da09) and K <= the
                              LDS rows (640 on gfx950); the fused multiply-adds
                              run in packed order, so results are the register
                              tile's bits.  0 = register tile (VALU)            */
} srt_tile_config;

/* Default configuration (the reference's SOPConfig defaults, SpMM_SOP.h:33-54,
 * except K_c = 0: one K tile; a caller passing the reference's 256 gets K tiles). */
void srt_tile_config_default(srt_tile_config *cfg);

/* ------------------------------------------------------------------------
 * Pattern mappings (nonzero code -> list of kernel patterns).
 * ---------------------------------------------------------------------- */
typedef struct srt_mapping srt_mapping;

/* Kernel ids.  A mapping's kernel patterns get ids by the reference
 * generator's rule: the kernel list it is handed (sbench/SOP/sop_driver.py:93)
 * stable-sorted by popcount, id = position in the sorted list
 * (tools/codegen/gen_vector_specialized.py:113,130-137).  The ids fix the
 * packed column order and so the accumulation order of every output.
 *   - srt_mapping_create_ordered takes the kernel list, e.g. a strategy's
 *     all_patterns(): merge then split patterns (sbench/SOP/
 *     sop_hand_configs.py:16-17, passed at sop_bench.py:105,154).
 *   - srt_mapping_read and srt_mapping_create build it as the reference's
 *     load_mapping does for a file (sop_bench.py:72-95): every pattern added
 *     to a Python set in file (code) order, then list(set) -- CPython 3.10's
 *     iteration order for small ints, restated in C. */

/* Read mapping_<mapping_id>.txt from the first directory that has it.
 * Replaces sop::read_pattern_mapping(mapping_id, dirs)
 * (cpp_testbed/demo/count_nkerns.cpp:54-55; file format sbench/SOP/mapping.py:14-17). */
int srt_mapping_read(const char *mapping_id, const char *const *search_dirs, int num_dirs,
                     srt_mapping **out);

/* Build a mapping from arrays: code_ptr[2^M_r + 1], code_patterns[code_ptr[2^M_r]];
 * the kernel patterns of code c are code_patterns[code_ptr[c] .. code_ptr[c+1]).
 * Same content as a mapping file (sbench/SOP/mapping.py:5-18). */
int srt_mapping_create(int M_r, const int32_t *code_ptr, const int32_t *code_patterns,
                       srt_mapping **out);

/* As srt_mapping_create, with the kernel list the ids derive from
 * (kernel_list[num_kernels]; repeated patterns keep their first position;
 * listed patterns the table never uses still get ids, as in the reference's
 * generator).  Fails if the table uses a pattern the list lacks.
 * Replaces make_sop_module(acc, patterns, mapping) (sbench/SOP/sop_driver.py:88-129). */
int srt_mapping_create_ordered(int M_r, const int32_t *code_ptr, const int32_t *code_patterns,
                               const int32_t *kernel_list, int num_kernels, srt_mapping **out);

/* M_r and the number of kernel patterns (MicroKernel::num_nkern_patterns, template:42). */
int srt_mapping_info(const srt_mapping *m, int *M_r, int *num_patterns);

/* id -> kernel pattern bitmask, num_patterns entries (decode_nkern_pattern, template:192),
 * in id order (see "Kernel ids" above). */
int srt_mapping_pattern_table(const srt_mapping *m, int32_t *patterns_out);

void srt_mapping_destroy(srt_mapping *m);

/* Executor id for a mapping's pattern set (get_executor_id, count_nkerns.cpp:50;
 * nano/runtime_yaml_config.register:92-98).  Fails with SRT_ERR_UNSUPPORTED if
 * no CDNA4 executor was generated for the pattern set. */
int srt_executor_id(const srt_mapping *m, char *buf, size_t buf_len);

/* Number of generated executors and the pattern set of each (for discovery). */
int srt_executor_count(void);
int srt_executor_describe(int index, int *M_r, int *num_patterns, int32_t *patterns_out,
                          char *id_buf, size_t id_buf_len);

/* ------------------------------------------------------------------------
 * MatMulSpecialized: pack once, execute many times.
 * Replaces sop::MatMulSpecialized<KernelDesc,DataTransform>
 * (ctor / get_config / allocate_executor / operator()(C,B) / require_storage,
 * used at cpp_testbed/src/spmm/nano/SpMM_SOP.h:64,89,117-134,155-164,178).
 * ---------------------------------------------------------------------- */
typedef struct srt_plan srt_plan;

/* Pack A (m x k CSR, device arrays Lx[nnz], Lp[m+1], Li[nnz]) on the device.
 * n is the widest B the plan will see.  executor_id may be NULL (derived from
 * the mapping).  nThreads is accepted for signature parity and ignored. */
int srt_plan_create(int m, int k, int n, const float *Lx, const void *Lp, const void *Li,
                    int index_bytes, const srt_tile_config *cfg, int nThreads,
                    const char *executor_id, const srt_mapping *mapping, srt_stream_t stream,
                    srt_plan **out);

/* A hybrid plan: the 16-row panels whose density (nonzeros / (rows * k))
 * reaches dense_threshold are stored dense and run on FP32 matrix cores
 * (v_mfma_f32_16x16x4_f32) after the register-tiled launch over the other
 * panels; the rest is as srt_plan_create (K tiles, schedules, execute).
 * The dense launch follows the register-tiled one (which leaves the dense
 * panels' rows out when M_r divides 16, else writes them as zero).  It
 * splits K over up to 32 workgroups per 128 x 128 output tile when that
 * fills the chip, adding the partials (kept in a plan-owned workspace sized
 * for the widest n of create / srt_plan_allocate_executor) in split order;
 * executes of one hybrid plan overlapping in time (several host threads or
 * streams) must be serialised by the caller: they share the workspace.  The
 * reference marks a tile dense at 0.30 (sbench/padding_reordering.py:36) and
 * pads patterns by its ILP cost model (sbench/ilp_pad/nano_solver.py:76-116);
 * a threshold above 1 gives a plain plan.  Dense panels accumulate in the
 * matrix core's order (k in steps of 4), within the FP32 summation bound,
 * not bit-identical to the register-tiled chain.  Hybrid plans execute
 * alone (not in groups or persistent tables). */
int srt_plan_create_hybrid(int m, int k, int n, const float *Lx, const void *Lp, const void *Li, int index_bytes,
                           const srt_tile_config *cfg, float dense_threshold, const srt_mapping *mapping,
                           srt_stream_t stream, srt_plan **out);
/* Dense panels of a hybrid plan: their count, first rows (num_dense entries,
 * may be NULL) and device bytes. */
int srt_plan_dense_panels(const srt_plan *p, int *num_dense, int32_t *panel_rows, size_t *bytes);

int srt_plan_get_config(const srt_plan *p, srt_tile_config *out);

/* Bytes of packed A held on the device (MatMulSpecialized::require_storage). */
int srt_plan_required_storage(const srt_plan *p, size_t *bytes);

/* Device bytes of the current launch schedule (wave descriptors and batch
 * records), beside the packed data.  No counterpart in the reference. */
int srt_plan_schedule_bytes(const srt_plan *p, size_t *bytes);

/* Free the packed column indices: the batch records already carry every
 * packed column's B row, so execution no longer reads them, and a resident
 * weight (the serving form) keeps only values, records and descriptors.
 * Afterwards the schedule cannot be rebuilt (srt_plan_allocate_executor with
 * a width that needs another schedule fails) and srt_plan_copy_packed cannot
 * return the columns.  Synchronises the device.  No counterpart in the
 * reference. */
int srt_plan_release_columns(srt_plan *p);

/* Fix the launch shape for B widths up to n (MatMulSpecialized::allocate_executor). */
int srt_plan_allocate_executor(srt_plan *p, int n);

/* C[0:m, 0:n] = A * B[0:k, 0:n]  (operator()(C,B)).  Device B (ldb >= n) and
 * C (ldc >= n).  n <= the n given at allocate_executor.  Rows and columns
 * need no padding: tails are masked. */
int srt_plan_execute(const srt_plan *p, float *C, int64_t ldc, const float *B, int64_t ldb,
                     int n, srt_stream_t stream);

/* Waves each panel's packed columns are split over (num_panels entries) in
 * the current launch schedule: chunk j of kp covers packed columns
 * [floor(j*n/kp), floor((j+1)*n/kp)) and C = chain_0 + chain_1 + ... in
 * order.  The load-balanced kernel descriptors of the reference
 * (KD_IntelLoadBalanced, nano/runtime_yaml_config.register:85-87) have no
 * published partition; this one is stated so callers can reproduce the order. */
int srt_plan_split_counts(const srt_plan *p, int32_t *kp);

/* Workgroups and threads per workgroup of one srt_plan_execute at width n
 * (aligned operands; the lane width may drop for misaligned ones).  No
 * counterpart in the reference: the launch geometry the measurement harness
 * gives its launch-floor probe. */
int srt_plan_launch_shape(const srt_plan *p, int n, int *grid, int *block);

/* Launch-floor probe with the executor's geometry (measurement only): kind 0
 * an empty kernel; kind 1 every wave walking `depth` (1..4) dependent scalar
 * loads through buf (buf_bytes >= 128, zero-filled; sink: one writable int). */
int srt_probe_launch_floor(int kind, int grid, int block, const void *buf, int64_t buf_bytes, int depth,
                           void *sink, srt_stream_t stream);

typedef struct srt_plan_stats {
    int M, K, M_r, num_panels, num_patterns;
    int64_t nnz;            /* stored entries of A                           */
    int64_t packed_cols;    /* packed columns over all panels                */
    int64_t packed_values;  /* stored values incl. padding                   */
    int64_t padding;        /* stored values equal to 0.0f (template:197-201) */
    int k_tile;             /* columns per K tile (K_c; K when untiled)       */
    int k_tiles;            /* K tiles per panel                              */
} srt_plan_stats;
int srt_plan_stats_get(const srt_plan *p, srt_plan_stats *out);

/* Copy the packed arrays to host memory (for inspection / parity checks):
 * counts[num_panels*k_tiles*num_patterns] (per panel, per K tile), col_ptr[num_panels+1], val_ptr[num_panels+1],
 * cols[packed_cols], vals[packed_values].  Any pointer may be NULL. */
int srt_plan_copy_packed(const srt_plan *p, int32_t *counts, int64_t *col_ptr, int64_t *val_ptr,
                         int32_t *cols, float *vals);

void srt_plan_destroy(srt_plan *p);

/* ------------------------------------------------------------------------
 * Grouped execution: several plans (the 1x1-conv layers of a network, each
 * with its own B and C) as one kernel launch.  No counterpart in the
 * reference, whose benchmark runs one MatMulSpecialized per matrix
 * (cpp_testbed/src/spmm/nano/SpMM_SOP.h:178); every problem keeps the
 * per-plan semantics of srt_plan_execute, bit for bit.
 *
 * The plans must share one executor (mapping) and waves_per_block and live
 * on the current device.  Operands are bound at create / rebind (a device
 * problem table is written and the stream synchronised); srt_group_execute
 * only launches, so it can be captured in a graph and replayed.
 * ---------------------------------------------------------------------- */
typedef struct srt_group srt_group;

int srt_group_create(const srt_plan *const *plans, int count, const float *const *B, const int64_t *ldb,
                     float *const *C, const int64_t *ldc, const int *n, srt_stream_t stream, srt_group **out);
/* Re-bind operands (same plans; shapes may change).  Not while a launch of
 * the group is pending or a graph holding it may replay. */
int srt_group_rebind(srt_group *g, const float *const *B, const int64_t *ldb, float *const *C,
                     const int64_t *ldc, const int *n, srt_stream_t stream);
int srt_group_execute(const srt_group *g, srt_stream_t stream);
/* Lane width (floats per lane) of the launch, workgroups, LDS bytes. */
int srt_group_info(const srt_group *g, int *vec_width, int *workgroups, int *lds_bytes);
void srt_group_destroy(srt_group *g);

/* ------------------------------------------------------------------------
 * SOPTile binding primitives (sbench/SOP/_C/pytorch_wrapper_intrin.cpp.template).
 * ---------------------------------------------------------------------- */

/* sort_and_encode_patterns (template:75-98): device pat_codes[n] (kernel
 * pattern bitmasks) -> device encoded[n] (ids, sorted ascending, zero pattern
 * SRT_ZERO_PATTERN_ID last) and device order[n] (stable permutation:
 * encoded[i] = id(pat_codes[order[i]])).  Fails on a pattern outside the set. */
int srt_sort_and_encode_patterns(const srt_mapping *m, const int32_t *pat_codes, int n,
                                 int32_t *encoded, int32_t *order, srt_stream_t stream);

typedef struct srt_tile srt_tile;

/* SOPTile(rows, cols, num_panels) (template:101-114). */
int srt_tile_create(const srt_mapping *m, int64_t rows, int64_t cols, int64_t num_panels,
                    srt_tile **out);

/* SOPTile::pack_panel (template:129-217): device pat_codes[n], col_indices[n]
 * (column of dense_panel), dense_panel[K x M_r] row-major.  Synchronises. */
int srt_tile_pack_panel(srt_tile *t, int64_t panel_id, int64_t col_offset, const int32_t *pat_codes,
                        const int32_t *col_indices, int n, const float *dense_panel, int K,
                        srt_stream_t stream);

/* SOPTile::pack_tile (template:219-261): every panel must be packed. */
int srt_tile_pack_tile(srt_tile *t, srt_stream_t stream);

/* The packed tile in the reference's buffer layout (SOPTile::pack_tile,
 * template:219-261), built by srt_tile_pack_tile: one device allocation of
 * required_bytes = sum over panels of 4*num_nnz + 4*num_col_indices +
 * 4*num_nkern + 3*64 (template:222-229), holding per panel the values, the
 * column indices and the nkern_counts, each segment starting on a 64-byte
 * boundary.  used_bytes = end of the last segment (<= required_bytes).
 * panel_offsets (3 * num_panels entries, may be NULL) receives the byte
 * offsets of each panel's three segments.  The buffer belongs to the tile. */
int srt_tile_buffer(const srt_tile *t, const void **buffer, size_t *required_bytes, size_t *used_bytes,
                    int64_t *panel_offsets);

/* Copy the first `bytes` (<= required_bytes) of that buffer to host memory. */
int srt_tile_copy_buffer(const srt_tile *t, void *host_out, size_t bytes);

/* SOPTile::padding (template:125-127). */
int srt_tile_padding(const srt_tile *t, int64_t *padding);

/* The panel executor over a packed tile (sop_multiple_panel_executor,
 * template:265-305): C[0:rows, 0:N_c] (+)= tile * B[:, 0:N_c]; load_c != 0
 * accumulates into C. */
int srt_tile_execute(const srt_tile *t, int N_c, const float *B, int64_t ldb, float *C,
                     int64_t ldc, int load_c, const srt_tile_config *cfg, srt_stream_t stream);

/* Split counts (as srt_plan_split_counts) of the tile's last executed configuration. */
int srt_tile_split_counts(const srt_tile *t, int32_t *kp);

int srt_tile_copy_packed(const srt_tile *t, int32_t *counts, int64_t *col_ptr, int64_t *val_ptr,
                         int32_t *cols, float *vals, int64_t *packed_cols, int64_t *packed_values);

void srt_tile_destroy(srt_tile *t);

/* ------------------------------------------------------------------------
 * Precision form: BF16 / FP8-e4m3 A and B, FP32 accumulation on MFMA.
 * No counterpart in the reference (its executors are FP32 AVX-512 only,
 * tools/codegen/gen_vector_specialized.py:266-349); the packing follows the
 * reference's ILP padding idea (sbench/ilp_pad/nano_solver.py:76-116) at its
 * dense extreme: 16-row panels x 32-column blocks in MFMA operand layout,
 * each block holding a nonzero stored dense; blocks without one are not
 * stored (the plan's [P][KB] slot table holds -1 for them and the executors
 * read zeros), so memory follows the occupied blocks.  With every block
 * present, block (p, kb) sits at p * KB + kb and no table is read.
 * FP8: A rows scaled by s_i = max|A_i| / 448, B columns by t_j = max|B_:j| /
 * 448, both computed in FP64 on the device; C = diag(s) (qA qB) diag(t).
 * ---------------------------------------------------------------------- */
enum { SRT_DT_BF16 = 1, SRT_DT_FP8 = 2 };

typedef struct srt_mfma_plan srt_mfma_plan;

typedef struct srt_mfma_stats {
    int M, K, num_panels, k_blocks, dtype;
    int64_t nnz, stored_blocks, nonzero_blocks;
} srt_mfma_stats;

/* Pack CSR A (m x k, device Lp/Li/Lx, 4- or 8-byte indices) into panels. */
int srt_mfma_plan_create(int m, int k, const float *Lx, const void *Lp, const void *Li, int index_bytes,
                         int dtype, srt_stream_t stream, srt_mfma_plan **out);
/* Workspace for K-split launches: up to ksplit_max partial tiles of m x n_max.
 * Allocates; call it before capturing launches in a graph. */
int srt_mfma_plan_reserve(srt_mfma_plan *p, int n_max, int ksplit_max);
/* fp32 B (rows = K x cols = N, row-major, ldb) -> Bt: N x ld_out in BF16 or
 * e4m3, k contiguous (the MFMA B operand order), ld_out >= K rounded up to 32,
 * k in [K, ld_out) zero; FP8 also writes the column scales (N floats). */
int srt_mfma_convert_b(int dtype, int64_t rows, int cols, const float *B, int64_t ldb, void *out,
                       int64_t ld_out, float *col_scales, srt_stream_t stream);
/* C[:m, :n] = A . B with B given transposed as produced by srt_mfma_convert_b
 * (Bt: n x ldb in the plan's dtype, ldb >= k rounded up to 32, rows 16-byte
 * aligned; col_scales for FP8).  shape selects (N tile, waves, panels per wave): 0 (64,4,2)
 * 1 (128,4,2) 2 (64,4,1) 3 (128,2,2) 4 (32,4,1) 5 (128,4,1) 6 (64,8,1) 7 (128,8,1) 8 (64,16,1)
 * 9 (32,8,1); the ksplit K
 * ranges are added in order. */
int srt_mfma_plan_execute(const srt_mfma_plan *p, float *C, int64_t ldc, const void *B, int64_t ldb,
                          const float *col_scales, int n, int shape, int ksplit, srt_stream_t stream);
/* As srt_mfma_plan_execute, allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * K-split partial tiles live in a library-owned static device array of
 * SRT_NOALLOC_SCRATCH_FLOATS floats (one per GPU, held until the process
 * ends), so ksplit * m * n must not exceed it (SRT_ERR_INVALID otherwise; no
 * reservation is needed).  Every call on a GPU shares that array: calls that
 * overlap on one GPU (other streams, other threads) race.  Nothing in it has
 * to start at zero: each call writes every partial it reads. */
#define SRT_NOALLOC_SCRATCH_FLOATS (8u << 20) /* 32 MiB */
int srt_mfma_plan_execute_noalloc(const srt_mfma_plan *p, float *C, int64_t ldc, const void *B, int64_t ldb,
                                  const float *col_scales, int n, int shape, int ksplit, srt_stream_t stream);

int srt_mfma_plan_stats_get(const srt_mfma_plan *p, srt_mfma_stats *out);
int srt_mfma_plan_row_scales(const srt_mfma_plan *p, float *out);
/* Host copies: slot[P*KB] (block slot), blk_kb[stored], blocks (512 elements each). */
int srt_mfma_plan_copy_blocks(const srt_mfma_plan *p, int32_t *slot, int32_t *blk_kb, void *blocks);
void srt_mfma_plan_destroy(srt_mfma_plan *p);

/* Grouped, persistent execution of many independent SpMMs in one launch
 * (a batch of served requests, the layers of a network).  No counterpart in
 * the reference (it runs one MatMulSpecialized per call, SpMM_SOP.h:178).
 * The host fills one srt_mfma_problem per SpMM (srt_mfma_problem_fill),
 * assigns tile_begin as the running sum of srt_mfma_problem_tiles in table
 * order, and copies the table plus a header {number of problems, total
 * tiles} (two ints) to device memory; the launch reads both from there, so
 * a captured graph replays any batch written into the same buffers.  Every
 * workgroup of the persistent grid walks tiles blockIdx, blockIdx + grid, ...
 * All problems share one dtype and launch shape; no K split. */
typedef struct srt_mfma_problem {
    const void *blocks;   /* plan's packed panels                         */
    const void *Bt;       /* transposed B operand (n x ldb)               */
    float *C;
    const float *rscale;  /* plan's FP8 row scales                        */
    const float *cscale;  /* FP8 column scales of B                       */
    int64_t ldb, ldc;
    int M, N, KB, P;
    int tile_begin;
    int vec_ok;
    const int32_t *slot;  /* plan's block table ([P][KB] stored index or -1),
                             NULL when every block is stored (at p * KB + kb) */
} srt_mfma_problem;

int srt_mfma_problem_fill(const srt_mfma_plan *p, const void *Bt, int64_t ldb, const float *col_scales, float *C,
                          int64_t ldc, int n, int tile_begin, srt_mfma_problem *out);
int srt_mfma_problem_tiles(const srt_mfma_problem *pr, int shape);
int srt_mfma_grouped_execute(int dtype, int shape, const srt_mfma_problem *d_probs, const int *d_hdr, int grid,
                             srt_stream_t stream);

/* Grouped B preparation: fp32 or BF16 B (K x N, row-major) -> transposed
 * operand (N x ldk, zeros past K), many problems in one launch (FP8 needs
 * cscale computed beforehand).  Same table + header convention. */
typedef struct srt_bprep_problem {
    const void *src;
    void *dst;
    const float *cscale;
    int64_t ld_src, ldk;
    int K, N;
    int tile_begin;
    int src_bf16;   /* 1: src holds BF16, 0: fp32 */
    int pad_to;     /* widening only: columns written per row, N <= pad_to <= ldk
                       ([N, pad_to) zero; columns from pad_to on untouched, so
                       several problems can share one destination); 0 = ldk */
    int reserved;
} srt_bprep_problem;

int srt_bprep_problem_tiles(const srt_bprep_problem *pr);
int srt_mfma_grouped_b_prep(int dtype, const srt_bprep_problem *d_probs, const int *d_hdr, int grid,
                            srt_stream_t stream);

/* ------------------------------------------------------------------------
 * Register-tile serving: many plans' executions in one persistent launch
 * over a device problem table.  No counterpart in the reference (one
 * MatMulSpecialized per call, SpMM_SOP.h:178).  The host fills one
 * srt_exec_problem per SpMM (srt_exec_problem_fill), sets block_begin to the
 * running sum of n_wblocks * ceil(n / (64 * vec_width)) in table order, and
 * copies the table plus a header {number of problems, total workgroups} to
 * device memory; the launch reads both from there, so a captured graph
 * replays any batch written into the same buffers.  Every problem's result is
 * bit-identical to srt_plan_execute of its plan at that lane width.  All
 * plans of one table share the executor (mapping) and waves_per_block of
 * `proto` and were built with unroll 4.
 * ---------------------------------------------------------------------- */
typedef struct srt_exec_problem {
    const int32_t *prog;   /* plan's batch records   */
    const int32_t *desc;   /* plan's wave descriptors */
    const float *vals;     /* plan's packed values    */
    const float *B;        /* K x ldb fp32            */
    float *C;              /* M x ldc fp32            */
    int64_t ldc;
    uint32_t b_bytes, ldb;
    int M, N, K, n_wblocks, reduce, load_c;
    int block_begin;
    int pad;
} srt_exec_problem;

/* One table record for plan p (B, C may be NULL with n = 0 for a template
 * whose operand fields the caller fills in later; then it must keep
 * b_bytes = K * ldb * 4 and the vec_width rules: vec_width | ldb, ldc; ldb
 * and ldc >= n rounded up to vec_width; 4 * vec_width-byte aligned B and C).
 * n need not be a multiple of vec_width: a lane's last vector reaches column
 * round_up(n, vec_width) - 1, so B's columns n .. that one are read (zeros
 * keep them harmless) and C's are overwritten; columns below n are the same
 * bits at every lane width. */
int srt_exec_problem_fill(const srt_plan *p, const float *B, int64_t ldb, float *C, int64_t ldc, int n,
                          int vec_width, int block_begin, srt_exec_problem *out);
int srt_exec_persistent_execute(const srt_plan *proto, int vec_width, const srt_exec_problem *d_probs,
                                const int *d_hdr, int grid, srt_stream_t stream);

/* Grouped BF16 -> FP32 widening of the requests' B operands (same table and
 * header convention, srt_bprep_problem with ldk = the FP32 row stride,
 * columns [N, pad_to) written as zero, pad_to = 0 meaning ldk; src_bf16 must
 * be 1).  A problem's tiles are 1024-element pieces of its K x pad_to output:
 * ceil(K * pad_to / 1024). */
int srt_widen_problem_tiles(const srt_bprep_problem *pr);
int srt_exec_grouped_b_widen(const srt_bprep_problem *d_probs, const int *d_hdr, int grid, srt_stream_t stream);

/* ------------------------------------------------------------------------
 * CSR baseline (sbench/SOP/_C/pytorch_csr_wrapper.cpp:243-312 _csr_kernel,
 * torch entry csr.executor :314-353): C[m x n] = A_csr * B, device arrays
 * rowptr[m+1], colind[nnz], vals[nnz].  n_tile in {16,32,64,128,256}
 * selects columns per wavefront (values below 64 run as 64); tails masked.
 * Row ranges are clamped to [0, nnz) and column indices outside [0, k) are
 * skipped, so malformed input cannot make the kernel read out of bounds.
 * ---------------------------------------------------------------------- */
int srt_csr_spmm(int m, int k, int n, int64_t nnz, const void *rowptr, const void *colind,
                 int index_bytes, const float *vals, const float *B, int64_t ldb, float *C,
                 int64_t ldc, int n_tile, srt_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SRT_H_ */
