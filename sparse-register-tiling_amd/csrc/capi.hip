// C ABI of libsrt: mappings, executor registry, MatMulSpecialized plans,
// SOPTile packing and the CSR baseline.  Declarations and the reference
// interface each entry point replaces are in include/srt.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <functional>
#include <memory>
#include <mutex>
#include <queue>
#include <tuple>
#include <sstream>
#include <string>
#include <vector>

#include "srt.h"
#include "srt_internal.h"

namespace srt {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

// ------------------------------------------------------------------ registry
static std::vector<const ExecutorDesc *> &registry() {
    static std::vector<const ExecutorDesc *> r;
    return r;
}

void register_executor(const ExecutorDesc *d) { registry().push_back(d); }

const std::vector<const ExecutorDesc *> &executors() { return registry(); }

const ExecutorDesc *find_executor(int M_r, const std::vector<int32_t> &patterns) {
    for (const ExecutorDesc *d : registry()) {
        if (d->M_r != M_r || d->num_patterns != (int)patterns.size()) continue;
        if (std::equal(patterns.begin(), patterns.end(), d->patterns)) return d;
    }
    return nullptr;
}

const ExecutorDesc *find_executor_by_id(const char *id) {
    for (const ExecutorDesc *d : registry())
        if (std::strcmp(d->id, id) == 0) return d;
    return nullptr;
}

void Packed::release() {
    hipFree(counts);
    hipFree(col_ptr);
    hipFree(val_ptr);
    hipFree(cols);
    hipFree(vals);
    hipFree(groups);
    hipFree(grp_ptr);
    hipFree(id_pattern);
    counts = nullptr;
    col_ptr = val_ptr = grp_ptr = nullptr;
    cols = groups = id_pattern = nullptr;
    vals = nullptr;
}

size_t Packed::bytes() const {
    return sizeof(int32_t) * (size_t)P * KT * NK + 3 * sizeof(int64_t) * (size_t)(P + 1) +
           (cols ? sizeof(int32_t) * (size_t)total_cols : 0) + sizeof(float) * (size_t)total_vals +
           2 * sizeof(int32_t) * (size_t)total_groups + sizeof(int32_t) * (size_t)NK;
}

// ------------------------------------------------------------------ mapping

// Iteration order of a CPython set of small non-negative ints (hash(i) == i)
// built by adding `keys` in order.  The reference's load_mapping collects a
// mapping file's kernel patterns into a set (sbench/SOP/sop_bench.py:72-95)
// and hands list(set) to its generator (sop_driver.py:93), so this order
// fixes the kernel ids of a mapping read from a file.  Restated from the
// published algorithm of CPython 3.10's set (open addressing over a power-of-
// two table of at least 8 slots; 9 linear probes after the home slot
// hash & mask, then i = 5i + 1 + perturb with perturb >>= 5; the table grows
// to the smallest power of two above 4 * used once fill * 5 >= mask * 3,
// re-inserting in slot order).  Checked against the interpreter's own sets.
static std::vector<int32_t> python_set_order(const std::vector<int32_t> &keys) {
    constexpr size_t kProbes = 9, kShift = 5, kMinSize = 8;
    std::vector<int64_t> table(kMinSize, -1);  // -1: empty slot
    size_t mask = kMinSize - 1, used = 0;
    auto insert_clean = [](std::vector<int64_t> &t, size_t m, int64_t key) {
        size_t perturb = (size_t)key, i = (size_t)key & m;
        while (true) {
            if (t[i] < 0) {
                t[i] = key;
                return;
            }
            if (i + kProbes <= m)
                for (size_t j = 1; j <= kProbes; ++j)
                    if (t[i + j] < 0) {
                        t[i + j] = key;
                        return;
                    }
            perturb >>= kShift;
            i = (i * 5 + 1 + perturb) & m;
        }
    };
    for (int32_t k : keys) {
        const int64_t key = k;
        size_t i = (size_t)key & mask, perturb = (size_t)key;
        size_t slot = SIZE_MAX;
        bool present = false;
        if (table[i] < 0) {
            slot = i;
        } else {
            while (true) {
                if (table[i] == key) {
                    present = true;
                    break;
                }
                if (i + kProbes <= mask) {
                    for (size_t j = 1; j <= kProbes && slot == SIZE_MAX && !present; ++j) {
                        if (table[i + j] < 0) slot = i + j;
                        else if (table[i + j] == key) present = true;
                    }
                    if (present || slot != SIZE_MAX) break;
                }
                perturb >>= kShift;
                i = (i * 5 + 1 + perturb) & mask;
                if (table[i] < 0) {
                    slot = i;
                    break;
                }
            }
        }
        if (present) continue;
        table[slot] = key;
        ++used;  // no deletions: fill == used
        if (used * 5 < mask * 3) continue;
        const size_t minused = used > 50000 ? used * 2 : used * 4;
        size_t size = kMinSize;
        while (size <= minused) size <<= 1;
        std::vector<int64_t> grown(size, -1);
        for (int64_t e : table)
            if (e >= 0) insert_clean(grown, size - 1, e);
        table.swap(grown);
        mask = size - 1;
    }
    std::vector<int32_t> out;
    for (int64_t e : table)
        if (e >= 0) out.push_back((int32_t)e);
    return out;
}

struct MappingImpl {
    int M_r = 0;
    std::vector<std::vector<int32_t>> table;  // code -> kernel patterns (list order)
    std::vector<int32_t> code_order;          // codes in file / array order
    bool has_kernel_list = false;             // kernel list given (else: the file's set)
    std::vector<int32_t> kernel_list;         // the list fed to the generator
    std::vector<int32_t> patterns;            // id -> pattern
    std::vector<int32_t> pattern_id;          // pattern -> id or -1
    std::vector<int32_t> lut_ptr, lut_ids;
    std::mutex mu;
    std::deque<DeviceLut> dev;  // per device, lazily uploaded (stable addresses)

    ~MappingImpl() {
        int cur = -1;
        hipGetDevice(&cur);
        for (DeviceLut &d : dev) {
            if (d.device < 0) continue;
            hipSetDevice(d.device);
            hipFree(d.lut_ptr);
            hipFree(d.lut_ids);
            hipFree(d.id_pattern);
            hipFree(d.pattern_id);
        }
        if (cur >= 0) hipSetDevice(cur);
    }

    // Validates the table and derives the id order and the code -> ids LUT.
    // Ids: the kernel list stable-sorted by popcount
    // (tools/codegen/gen_vector_specialized.py:113,130-137); the list is the
    // one given, or, for a table read from a file, list(set) of its patterns
    // in file order (python_set_order).
    int finalize() {
        const int ncode = 1 << M_r;
        if ((int)table.size() != ncode) table.resize(ncode);
        if (code_order.empty())
            for (int c = 1; c < ncode; ++c) code_order.push_back(c);
        std::vector<int32_t> file_list;
        for (int c : code_order)
            if (c > 0 && c < ncode)
                for (int32_t p : table[c])
                    if (p > 0) file_list.push_back(p);
        std::vector<char> seen(ncode, 0);
        for (int c = 1; c < ncode; ++c) {
            unsigned uni = 0;
            if (table[c].empty())
                return fail(SRT_ERR_INVALID, "mapping has no kernel patterns for code " +
                                                 std::to_string(c));
            std::vector<int32_t> kept;
            for (int32_t p : table[c]) {
                if (p == 0) continue;  // a zero pattern stores nothing
                if (p < 0 || p >= ncode)
                    return fail(SRT_ERR_INVALID, "kernel pattern " + std::to_string(p) +
                                                     " does not fit M_r=" + std::to_string(M_r));
                if ((unsigned)p & uni)
                    return fail(SRT_ERR_INVALID, "kernel patterns of code " + std::to_string(c) +
                                                     " overlap (a value would be added twice)");
                uni |= (unsigned)p;
                kept.push_back(p);
                seen[p] = 1;
            }
            if ((uni & (unsigned)c) != (unsigned)c)
                return fail(SRT_ERR_INVALID, "kernel patterns of code " + std::to_string(c) +
                                                 " do not cover it");
            table[c] = kept;
        }
        patterns.clear();
        std::vector<char> listed(ncode, 0);
        for (int32_t p : has_kernel_list ? kernel_list : python_set_order(file_list)) {
            if (p == 0) continue;
            if (p < 0 || p >= ncode)
                return fail(SRT_ERR_INVALID, "kernel list pattern " + std::to_string(p) +
                                                 " does not fit M_r=" + std::to_string(M_r));
            if (listed[p]) continue;  // a repeated pattern keeps its first position
            listed[p] = 1;
            patterns.push_back(p);
        }
        for (int p = 1; p < ncode; ++p)
            if (seen[p] && !listed[p])
                return fail(SRT_ERR_INVALID, "the mapping uses kernel pattern " + std::to_string(p) +
                                                 ", which the kernel list lacks");
        std::stable_sort(patterns.begin(), patterns.end(), [](int32_t a, int32_t b) {
            return __builtin_popcount((unsigned)a) < __builtin_popcount((unsigned)b);
        });
        if (patterns.size() > 4096)
            return fail(SRT_ERR_UNSUPPORTED, "more than 4096 kernel patterns in one mapping");
        pattern_id.assign(ncode, -1);
        for (size_t i = 0; i < patterns.size(); ++i) pattern_id[patterns[i]] = (int32_t)i;
        lut_ptr.assign(ncode + 1, 0);
        lut_ids.clear();
        for (int c = 0; c < ncode; ++c) {
            lut_ptr[c] = (int32_t)lut_ids.size();
            if (c)
                for (int32_t p : table[c]) lut_ids.push_back(pattern_id[p]);
        }
        lut_ptr[ncode] = (int32_t)lut_ids.size();
        return SRT_OK;
    }

    int device_lut(DeviceLut **out) {
        int device = 0;
        SRT_HIP_TRY(hipGetDevice(&device));
        std::lock_guard<std::mutex> g(mu);
        for (DeviceLut &d : dev)
            if (d.device == device) {
                *out = &d;
                return SRT_OK;
            }
        DeviceLut d;
        d.M_r = M_r;
        d.num_patterns = (int)patterns.size();
        const size_t ncode = (size_t)1 << M_r;
        SRT_HIP_TRY(hipMalloc(&d.lut_ptr, sizeof(int32_t) * (ncode + 1)));
        SRT_HIP_TRY(hipMalloc(&d.lut_ids, sizeof(int32_t) * std::max<size_t>(1, lut_ids.size())));
        SRT_HIP_TRY(hipMalloc(&d.id_pattern, sizeof(int32_t) * std::max<size_t>(1, patterns.size())));
        SRT_HIP_TRY(hipMalloc(&d.pattern_id, sizeof(int32_t) * ncode));
        SRT_HIP_TRY(hipMemcpy(d.lut_ptr, lut_ptr.data(), sizeof(int32_t) * (ncode + 1),
                              hipMemcpyHostToDevice));
        if (!lut_ids.empty())
            SRT_HIP_TRY(hipMemcpy(d.lut_ids, lut_ids.data(), sizeof(int32_t) * lut_ids.size(),
                                  hipMemcpyHostToDevice));
        if (!patterns.empty())
            SRT_HIP_TRY(hipMemcpy(d.id_pattern, patterns.data(),
                                  sizeof(int32_t) * patterns.size(), hipMemcpyHostToDevice));
        SRT_HIP_TRY(hipMemcpy(d.pattern_id, pattern_id.data(), sizeof(int32_t) * ncode,
                              hipMemcpyHostToDevice));
        d.device = device;
        dev.push_back(d);
        *out = &dev.back();
        return SRT_OK;
    }
};

// Parses "M_r\n<code>: [p, p, ...]\n..." (sbench/SOP/mapping.py:14-17).
static int parse_mapping(std::istream &in, MappingImpl &m, const std::string &what) {
    std::string line;
    if (!std::getline(in, line)) return fail(SRT_ERR_IO, "empty mapping file " + what);
    m.M_r = std::atoi(line.c_str());
    if (m.M_r < 1 || m.M_r > SRT_MAX_MR)
        return fail(SRT_ERR_INVALID, "mapping " + what + ": M_r must be in [1, 16]");
    m.table.assign((size_t)1 << m.M_r, {});
    int lineno = 1;
    while (std::getline(in, line)) {
        ++lineno;
        size_t colon = line.find(':');
        if (colon == std::string::npos) {
            bool blank = std::all_of(line.begin(), line.end(), [](char ch) { return std::isspace((unsigned char)ch); });
            if (blank) continue;
            return fail(SRT_ERR_INVALID, what + ":" + std::to_string(lineno) + ": expected 'code: [..]'");
        }
        long code = std::strtol(line.c_str(), nullptr, 10);
        if (code < 1 || code >= (1L << m.M_r))
            return fail(SRT_ERR_INVALID, what + ":" + std::to_string(lineno) + ": code out of range");
        std::vector<int32_t> pats;
        const char *s = line.c_str() + colon + 1;
        while (*s) {
            if (std::isdigit((unsigned char)*s) || *s == '-') {
                char *e = nullptr;
                long v = std::strtol(s, &e, 10);
                pats.push_back((int32_t)v);
                s = e;
            } else {
                ++s;
            }
        }
        m.table[code] = pats;
        m.code_order.push_back((int)code);
    }
    return m.finalize();
}

}  // namespace srt

using namespace srt;

struct srt_mapping {
    std::shared_ptr<MappingImpl> impl;
};

// ------------------------------------------------------------------- plans
// Dense 16-row panels of a hybrid plan (srt_plan_create_hybrid), run on FP32
// matrix cores after the register-tiled part (dense_panels.hip).
struct DensePart {
    int ndense = 0, G = 0;          // dense panels, 16-column K groups
    void *apack = nullptr;          // [ndense][G][64] float4, MFMA A-operand order
    int32_t *panel_row = nullptr;   // first row of each dense panel
    std::vector<int32_t> h_panel_row;
    size_t bytes = 0;
    // K-split partials of the dense launch (sized for the widest B so far;
    // earlier ones stay alive for graphs that captured them)
    void *ws = nullptr;
    size_t ws_bytes = 0;
    std::vector<void *> retired_ws;
    int reserve(int n) {
        const size_t need = dense_panels_workspace_bytes(ndense, G, n);
        if (need <= ws_bytes) return SRT_OK;
        void *w = nullptr;
        const hipError_t e = hipMalloc(&w, need);
        if (e != hipSuccess) return fail(SRT_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        if (ws) retired_ws.push_back(ws);
        ws = w;
        ws_bytes = need;
        return SRT_OK;
    }
    void release() {
        (void)hipFree(apack);
        (void)hipFree(panel_row);
        (void)hipFree(ws);
        for (void *w : retired_ws) (void)hipFree(w);
        retired_ws.clear();
        ws = nullptr;
        ws_bytes = 0;
        apack = nullptr;
        panel_row = nullptr;
        ndense = 0;
    }
};

struct srt_plan {
    Packed pk;
    Schedule sch;
    const ExecutorDesc *exec = nullptr;
    srt_tile_config cfg_user{};  // as given (0 / -1 fields = choose)
    srt_tile_config cfg{};       // realised
    int n_max = 0;
    int n_alloc = 0;
    int device = 0;
    std::vector<RetiredSchedule> retired;  // schedules of earlier widths, alive until destroy
    DensePart dense;                       // hybrid plans: panels on matrix cores
    // matrix_cores: panel ranges for R ranges (the last entry is current;
    // earlier ones stay alive for graphs that captured them)
    std::vector<std::pair<int, int32_t *>> mfma_ranges;
};

struct PanelBuf {
    int32_t *counts = nullptr;  // NK + 1 (last = zero-pattern entries)
    int32_t *cols = nullptr;
    float *vals = nullptr;
    int64_t ncols = 0, nvals = 0, padding = 0;
    bool packed = false;
};

struct srt_tile {
    std::shared_ptr<MappingImpl> map;
    int64_t rows = 0, cols = 0, num_panels = 0;
    std::vector<PanelBuf> panels;
    bool tile_packed = false;
    Packed pk;
    const ExecutorDesc *exec = nullptr;
    int64_t padding_total = 0;
    int device = 0;
    // Schedule of the last executed configuration (rebuilt when the realised
    // nsplit / chunk_cols / waves_per_block change).
    mutable std::mutex mu;
    mutable Schedule sch;
    mutable srt_tile_config sch_cfg{};
    mutable std::vector<RetiredSchedule> retired;  // earlier schedules, alive until destroy
    // The reference's packed buffer (template:219-261): per panel
    // [values][pad64][col_indices][pad64][nkern_counts][pad64] in one
    // allocation of required_bytes; byte offsets of the three segments per panel.
    void *ref_buffer = nullptr;
    size_t ref_required = 0, ref_used = 0;
    std::vector<int64_t> ref_offsets;
};

extern "C" {

const char *srt_last_error(void) { return g_last_error.c_str(); }

int srt_version(void) { return 10000; }

void srt_tile_config_default(srt_tile_config *c) {
    if (!c) return;
    // SOPConfig defaults (cpp_testbed/src/spmm/nano/SpMM_SOP.h:33-54).
    c->M_c = 16;
    c->K_c = 0;  /* 0: one K tile (the reference's SOPConfig default is 256) */
    c->N_c = 64;
    c->tiling_strategy = 0;
    c->beta = 1.0f;
    c->sparse_a = 1;
    c->max_tlb_entries = 64;
    c->tlb_page_size = 4096;
    c->runtime_schedule = 0;
    c->vec_width = 0;
    c->waves_per_block = 0;
    c->xcd_remap = -1;
    c->unroll = 0;
    c->nsplit = 0;
    c->chunk_cols = 0;
    c->b_staging = 0;
    c->layout = 0;
    c->matrix_cores = 0;
}

// ----------------------------------------------------------------- mapping
int srt_mapping_read(const char *mapping_id, const char *const *dirs, int num_dirs,
                     srt_mapping **out) {
    if (!mapping_id || !out) return fail(SRT_ERR_INVALID, "srt_mapping_read: null argument");
    std::string name = std::string("mapping_") + mapping_id + ".txt";
    for (int i = 0; i < num_dirs; ++i) {
        if (!dirs || !dirs[i]) continue;
        std::string path = std::string(dirs[i]);
        if (!path.empty() && path.back() != '/') path += '/';
        path += name;
        std::ifstream f(path);
        if (!f) continue;
        auto impl = std::make_shared<MappingImpl>();
        int rc = parse_mapping(f, *impl, path);
        if (rc) return rc;
        *out = new srt_mapping{impl};
        return SRT_OK;
    }
    return fail(SRT_ERR_IO, "mapping " + std::string(mapping_id) + " not found in search dirs");
}

int srt_mapping_create(int M_r, const int32_t *code_ptr, const int32_t *code_patterns,
                       srt_mapping **out) {
    return srt_mapping_create_ordered(M_r, code_ptr, code_patterns, nullptr, -1, out);
}

int srt_mapping_create_ordered(int M_r, const int32_t *code_ptr, const int32_t *code_patterns,
                               const int32_t *kernel_list, int num_kernels, srt_mapping **out) {
    if (!code_ptr || !out) return fail(SRT_ERR_INVALID, "srt_mapping_create: null argument");
    if (M_r < 1 || M_r > SRT_MAX_MR) return fail(SRT_ERR_INVALID, "M_r must be in [1, 16]");
    if (num_kernels > 0 && !kernel_list) return fail(SRT_ERR_INVALID, "null kernel list");
    auto impl = std::make_shared<MappingImpl>();
    impl->M_r = M_r;
    if (num_kernels >= 0) {
        impl->has_kernel_list = true;
        impl->kernel_list.assign(kernel_list, kernel_list + num_kernels);
    }
    const int ncode = 1 << M_r;
    impl->table.assign(ncode, {});
    for (int c = 0; c < ncode; ++c) {
        if (code_ptr[c + 1] < code_ptr[c]) return fail(SRT_ERR_INVALID, "code_ptr not monotone");
        for (int e = code_ptr[c]; e < code_ptr[c + 1]; ++e) impl->table[c].push_back(code_patterns[e]);
    }
    impl->table[0].clear();
    int rc = impl->finalize();
    if (rc) return rc;
    *out = new srt_mapping{impl};
    return SRT_OK;
}

int srt_mapping_info(const srt_mapping *m, int *M_r, int *num_patterns) {
    if (!m) return fail(SRT_ERR_INVALID, "null mapping");
    if (M_r) *M_r = m->impl->M_r;
    if (num_patterns) *num_patterns = (int)m->impl->patterns.size();
    return SRT_OK;
}

int srt_mapping_pattern_table(const srt_mapping *m, int32_t *out) {
    if (!m || !out) return fail(SRT_ERR_INVALID, "null argument");
    std::copy(m->impl->patterns.begin(), m->impl->patterns.end(), out);
    return SRT_OK;
}

void srt_mapping_destroy(srt_mapping *m) { delete m; }

int srt_executor_id(const srt_mapping *m, char *buf, size_t len) {
    if (!m || !buf) return fail(SRT_ERR_INVALID, "null argument");
    const ExecutorDesc *d = find_executor(m->impl->M_r, m->impl->patterns);
    if (!d) return fail(SRT_ERR_UNSUPPORTED, "no CDNA4 executor generated for this pattern set");
    std::snprintf(buf, len, "%s", d->id);
    return SRT_OK;
}

int srt_executor_count(void) { return (int)executors().size(); }

int srt_executor_describe(int index, int *M_r, int *num_patterns, int32_t *patterns_out,
                          char *id_buf, size_t id_buf_len) {
    if (index < 0 || index >= (int)executors().size()) return fail(SRT_ERR_INVALID, "bad index");
    const ExecutorDesc *d = executors()[index];
    if (M_r) *M_r = d->M_r;
    if (num_patterns) *num_patterns = d->num_patterns;
    if (patterns_out) std::copy(d->patterns, d->patterns + d->num_patterns, patterns_out);
    if (id_buf) std::snprintf(id_buf, id_buf_len, "%s", d->id);
    return SRT_OK;
}

}  // extern "C"

// ------------------------------------------------------------ launch shape
namespace srt {

static int log2_small(int v) { return v >= 8 ? 3 : v >= 4 ? 2 : v >= 2 ? 1 : 0; }

static bool pow2_upto8(int v) { return v == 1 || v == 2 || v == 4 || v == 8; }

static constexpr int kMaxWaves = 16;        // wave slots per workgroup (launch bound 1024)
static constexpr size_t kMaxLdsL2 = 65536;  // reduction LDS of the global-memory B path

static size_t reduce_lds(int waves, int M_r, int V) { return (size_t)waves * M_r * V * 64 * sizeof(float); }
static size_t tile_lds(int K, int V) { return (size_t)K * 64 * V * sizeof(float); }

size_t lds_limit() {
    static int cached[64];  // per device, 0 = not queried
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return kMaxLdsL2;
    if (!cached[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = (int)kMaxLdsL2;
        cached[dev] = v;
    }
    return (size_t)cached[dev];
}

// Whether a launch of lane width V and `waves` slots fits the LDS budget of
// its B path (stage: staged tile and reduction share one allocation).
static bool lds_fits(bool stage, int K, int waves, int M_r, int V, bool reduce) {
    const size_t red = reduce ? reduce_lds(waves, M_r, V) : 0;
    if (!stage) return red <= kMaxLdsL2;
    return std::max(red, tile_lds(K, V)) <= lds_limit();
}

// Compute units of the current device (256 on an MI355X).
static int cu_count() {
    static int cached[64];  // per device, 0 = not queried
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cached[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        cached[dev] = v;
    }
    return cached[dev];
}

// Realise the GPU fields of a tile config for a matrix of K columns with
// `total_cols` packed columns and B width n (the choices the reference's
// heuristic JSON made per sparsity bucket are made here from the launch
// geometry; the autotuner overrides them with measured choices).
static void realise_config(srt_tile_config &c, int M_r, int K, int K_tile, int64_t total_cols, int n) {
    if (c.vec_width != 1 && c.vec_width != 2 && c.vec_width != 4)
        c.vec_width = (n % 4 == 0 && n >= 256) ? 4 : (n % 2 == 0 && n >= 128) ? 2 : 1;
    if (c.waves_per_block < 1 || c.waves_per_block > kMaxWaves) c.waves_per_block = 8;
    if (c.nsplit < 1 || c.nsplit > c.waves_per_block) c.nsplit = c.waves_per_block;
    // B path: global-memory rows unless the LDS tile is asked for and its
    // narrowest form (V = 1) fits (measured faster on the pruned-FFN shapes).
    if (c.b_staging != 1 && c.b_staging != 2) c.b_staging = 2;
    const bool can_stage = tile_lds(std::max(K, 1), 1) <= lds_limit();
    if (c.b_staging == 1 && !can_stage) c.b_staging = 2;
    const bool stage = c.b_staging == 1;
    const bool split = c.nsplit > 1;
    while (c.vec_width > 1 && !lds_fits(stage, K, c.waves_per_block, M_r, c.vec_width, split))
        c.vec_width >>= 1;
    while (c.waves_per_block > 1 && !lds_fits(stage, K, c.waves_per_block, M_r, c.vec_width, split)) {
        c.waves_per_block >>= 1;
        c.nsplit = std::min(c.nsplit, c.waves_per_block);
    }
    if (c.chunk_cols < 1) {
        // About four waves per SIMD over the whole launch: long enough chains
        // to amortise the chunk start, short enough to hide latency.
        const int64_t n_tiles = (n + 64 * c.vec_width - 1) / (64 * c.vec_width);
        const int64_t target_waves = 4096;
        c.chunk_cols = (int)std::max<int64_t>(8, (total_cols * n_tiles + target_waves - 1) / target_waves);
    }
    if (c.xcd_remap != 0 && c.xcd_remap != 1) c.xcd_remap = 1;
    if (!pow2_upto8(c.unroll)) c.unroll = 4;
    if (c.layout != 1 && c.layout != 2) c.layout = 2;
    if (c.matrix_cores != 1) c.matrix_cores = 0;
    c.M_c = M_r * c.waves_per_block;
    c.N_c = 64 * c.vec_width;
    c.K_c = K_tile;
}

static bool aligned(const void *p, int bytes) { return ((uintptr_t)p % (uintptr_t)bytes) == 0; }

static int run_executor(const Packed &pk, const Schedule &sch, const ExecutorDesc *exec,
                        const srt_tile_config &cfg, const float *B, int64_t ldb, float *C,
                        int64_t ldc, int n, int load_c, hipStream_t s) {
    if (n < 0) return fail(SRT_ERR_INVALID, "n must be >= 0");
    if (n == 0 || pk.M == 0) return SRT_OK;
    if (!B || !C) return fail(SRT_ERR_INVALID, "null B or C");
    if (ldb < n || ldc < n) return fail(SRT_ERR_INVALID, "ldb and ldc must be >= n");
    if (!sch.desc || sch.waves_per_block < 1 || sch.waves_per_block > kMaxWaves)
        return fail(SRT_ERR_INVALID, "no launch schedule");
    int V = cfg.vec_width;
    // Vector loads need V | n, V | ldb, V | ldc and 4V-byte aligned B and C.
    while (V > 1 && (n % V || ldb % V || ldc % V || !aligned(B, 4 * V) || !aligned(C, 4 * V)))
        V >>= 1;
    bool stage = cfg.b_staging == 1 && tile_lds(pk.K, 1) <= lds_limit();
    if (stage) {  // executors built at run time may hold the global-memory path only
        bool any = false;
        for (int v = 0; v < 3; ++v)
            for (int u = 0; u < 4; ++u) any = any || exec->launch[1][v][u] != nullptr;
        stage = any;
    }
    while (V > 1 && !lds_fits(stage, pk.K, sch.waves_per_block, pk.M_r, V, sch.reduce)) V >>= 1;
    if (stage && !lds_fits(true, pk.K, sch.waves_per_block, pk.M_r, V, sch.reduce)) stage = false;
    if (!lds_fits(stage, pk.K, sch.waves_per_block, pk.M_r, V, sch.reduce))
        return fail(SRT_ERR_UNSUPPORTED, "chunk reduction needs more LDS than a workgroup has");
    // The batch records are cut for sch.unroll: the launch must use that U.
    const int ui = log2_small(sch.unroll);
    const int vi = log2_small(V);
    ExecLaunchFn fn = exec->launch[stage][vi][ui];
    if (!fn && stage) {
        stage = false;
        if (!lds_fits(false, pk.K, sch.waves_per_block, pk.M_r, V, sch.reduce))
            return fail(SRT_ERR_UNSUPPORTED, "chunk reduction needs more LDS than a workgroup has");
        fn = exec->launch[0][vi][ui];
    }
    if (!fn || !sch.prog) return fail(SRT_ERR_UNSUPPORTED, "executor variant not generated");
    // B rows are addressed with a 32-bit byte offset (buffer loads).
    const int64_t b_extent = (int64_t)pk.K * ldb * (int64_t)sizeof(float);
    if (b_extent > (int64_t)0xffffffffLL)
        return fail(SRT_ERR_UNSUPPORTED, "B larger than 4 GiB (K * ldb * 4 bytes)");
    ExecLaunch a;
    a.prog = sch.prog;
    a.desc = sch.desc;
    a.n_wblocks = sch.n_wblocks;
    a.reduce = sch.reduce;
    a.vals = pk.vals;
    a.B = B;
    a.b_bytes = (uint32_t)std::max<int64_t>(b_extent, 4);
    a.C = C;
    a.ldb = ldb;
    a.ldc = ldc;
    a.M = pk.M;
    a.N = n;
    a.K = pk.K;
    a.num_panels = pk.P;
    a.load_c = load_c;
    a.vec_width = V;
    a.waves_per_block = sch.waves_per_block;
    a.xcd_remap = cfg.xcd_remap;
    a.unroll = sch.unroll;
    SRT_HIP_TRY(fn(a, s));
    return SRT_OK;
}

template <typename T>
static int dmalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess)
        return fail(SRT_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return SRT_OK;
}

struct Scratch {
    std::vector<void *> ptrs;
    ~Scratch() {
        for (void *p : ptrs) hipFree(p);
    }
    template <typename T>
    int get(T **p, size_t count) {
        int rc = dmalloc(p, count);
        if (!rc) ptrs.push_back(*p);
        return rc;
    }
};

#define SRT_TRY(expr)            \
    do {                         \
        int _rc = (expr);        \
        if (_rc) return _rc;     \
    } while (0)

// Per-panel lists of non-empty groups (id, count) from pk.counts, the id ->
// pattern table, and a host copy of the column offsets.
static int build_groups(Packed &pk, const int32_t *d_id_pattern, hipStream_t s) {
    // Group lists per (panel, K tile) segment, laid out in segment order, so
    // the list of panel p is the concatenation of its K tiles' lists; the
    // panel-level offsets are every KT-th segment offset.
    const int Q = pk.P * pk.KT;
    Scratch sc;
    int64_t *ng = nullptr, *seg_ptr = nullptr;
    SRT_TRY(sc.get(&ng, Q));
    SRT_TRY(dmalloc(&pk.grp_ptr, (size_t)pk.P + 1));
    int64_t *gp = pk.grp_ptr;
    if (pk.KT > 1) {
        SRT_TRY(sc.get(&seg_ptr, (size_t)Q + 1));
        gp = seg_ptr;
    }
    SRT_HIP_TRY(launch_group_count(Q, pk.NK, pk.counts, ng, s));
    SRT_HIP_TRY(launch_exclusive_scan(Q, ng, gp, s));
    int64_t tot = 0;
    SRT_HIP_TRY(hipMemcpyAsync(&tot, gp + Q, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SRT_HIP_TRY(hipStreamSynchronize(s));
    pk.total_groups = tot;
    SRT_TRY(dmalloc(&pk.groups, 2 * ((size_t)tot + SRT_GROUP_PAD)));
    SRT_HIP_TRY(hipMemsetAsync(pk.groups + 2 * tot, 0, sizeof(int32_t) * 2 * SRT_GROUP_PAD, s));
    SRT_HIP_TRY(launch_group_fill(Q, pk.NK, pk.counts, gp, d_id_pattern, pk.groups, s));
    if (pk.KT > 1)
        SRT_HIP_TRY(hipMemcpy2DAsync(pk.grp_ptr, sizeof(int64_t), seg_ptr, sizeof(int64_t) * pk.KT,
                                     sizeof(int64_t), (size_t)pk.P + 1, hipMemcpyDeviceToDevice, s));
    SRT_TRY(dmalloc(&pk.id_pattern, (size_t)std::max(pk.NK, 1)));
    if (pk.NK)
        SRT_HIP_TRY(hipMemcpyAsync(pk.id_pattern, d_id_pattern, sizeof(int32_t) * pk.NK,
                                   hipMemcpyDeviceToDevice, s));
    pk.h_col_ptr.assign((size_t)pk.P + 1, 0);
    SRT_HIP_TRY(hipMemcpyAsync(pk.h_col_ptr.data(), pk.col_ptr, sizeof(int64_t) * (pk.P + 1),
                               hipMemcpyDeviceToHost, s));
    SRT_HIP_TRY(hipStreamSynchronize(s));
    return SRT_OK;
}

void Schedule::release() {
    hipFree(desc);
    hipFree(prog);
    desc = nullptr;
    prog = nullptr;
    n_wblocks = 0;
    kp.clear();
}

// Load-balanced launch schedule for a realised config: panel p gets
// kp = min(nsplit, max(1, ceil(ncols_p / chunk_cols))) waves, panels are laid
// into workgroups of waves_per_block slots in panel order (a panel never
// straddles two workgroups), and the wave descriptors are computed on the
// device from the group lists.
// The unroll the batch records are cut for: the requested one when the
// executor has it for every lane width of the global-memory path, else 4
// (which every generated executor has).
static int schedule_unroll(const ExecutorDesc *exec, int want) {
    const int ui = log2_small(want);
    for (int v = 0; v < 3; ++v)
        if (!exec->launch[0][v][ui]) return 4;
    return want;
}

// N tiles of a launch of width n under a realised config.
static int config_tiles(const srt_tile_config &c, int n) {
    const int nt = 64 * std::max(1, c.vec_width);
    return std::max(1, (n + nt - 1) / nt);
}

// Layout 2: panels binned heaviest first (ties in panel order) into the
// least-loaded of G workgroups that still has a free slot, G per N tile such
// that G x n_tiles is a whole number of rounds of the chip's CUs at about
// chunk_cols packed columns per wave (so no CU runs one workgroup more than
// another); then each workgroup's free slots go one at a time to its panel
// whose chunk is longest, while that chunk is longer than 2 * kMinChunk.
// A panel adds at least one average slot's columns to its bin's load, and
// equal loads go to the bin with fewer panels: empty or light panels (rows
// of a hybrid plan's dense panels, dead rows of a pruned matrix) would
// otherwise all pile into one bin, taking the slots its heavy panel needs
// for splitting (one unsplit chain of thousands of columns).
// Fills kp and appends (p, j, kp) per slot, W slots per workgroup.
static void balanced_slots(const Packed &pk, Schedule &sch, int n_tiles, std::vector<int32_t> &slots) {
    constexpr int64_t kMinChunk = 8;
    const int W = sch.waves_per_block, P = pk.P;
    std::vector<int64_t> nc(P);
    int64_t total = 0;
    for (int p = 0; p < P; ++p) {
        nc[p] = pk.h_col_ptr[p + 1] - pk.h_col_ptr[p];
        total += nc[p];
    }
    const int64_t cus = cu_count();
    const int64_t per_round = cus * W * (int64_t)sch.chunk_cols;
    const int64_t rounds = std::max<int64_t>(1, (total * n_tiles + per_round / 2) / per_round);
    int64_t G = (rounds * cus + n_tiles - 1) / n_tiles;
    G = std::min<int64_t>(G, P);
    G = std::max<int64_t>(G, (P + W - 1) / W);
    G = std::max<int64_t>(G, 1);
    std::vector<int> order;
    order.reserve(P);
    for (int p = 0; p < P; ++p)
        if (pk.h_skip.empty() || !pk.h_skip[p]) order.push_back(p);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nc[a] > nc[b]; });
    std::vector<std::vector<int>> bins((size_t)G);
    std::vector<int64_t> load((size_t)G, 0);
    const int64_t per_slot = std::max<int64_t>(1, total / (G * W));
    typedef std::tuple<int64_t, int, int64_t> Entry;  // (load, panels, bin): the smallest pops first
    std::priority_queue<Entry, std::vector<Entry>, std::greater<Entry>> heap;
    for (int64_t b = 0; b < G; ++b) heap.push(Entry(0, 0, b));
    for (int p : order) {  // G * W >= P, so a bin with a free slot remains
        const int64_t b = std::get<2>(heap.top());
        heap.pop();
        bins[b].push_back(p);
        load[b] += std::max(nc[p], per_slot);
        if ((int)bins[b].size() < W) heap.push(Entry(load[b], (int)bins[b].size(), b));
    }
    for (int64_t b = 0; b < G; ++b) {
        const std::vector<int> &L = bins[b];
        std::vector<int> k(L.size(), 1);
        for (int free = W - (int)L.size(); free > 0; --free) {
            int best = -1;
            int64_t longest = 2 * kMinChunk;
            for (size_t i = 0; i < L.size(); ++i) {
                const int64_t len = (nc[L[i]] + k[i] - 1) / k[i];
                if (k[i] < sch.nsplit && len > longest) {
                    longest = len;
                    best = (int)i;
                }
            }
            if (best < 0) break;
            ++k[best];
        }
        int used = 0;
        for (size_t i = 0; i < L.size(); ++i) {
            sch.kp[L[i]] = k[i];
            if (k[i] > 1) sch.reduce = 1;
            for (int j = 0; j < k[i]; ++j) {
                slots.push_back(L[i]);
                slots.push_back(j);
                slots.push_back(k[i]);
            }
            used += k[i];
        }
        for (; used < W; ++used) {
            slots.push_back(-1);
            slots.push_back(0);
            slots.push_back(1);
        }
    }
}

// Builds the launch schedule of `cfg` into `out`.  The schedule `out` held
// before moves to `retired` and stays alive: a graph captured with it keeps
// replaying from valid memory until the owner is destroyed.
static int build_schedule(const Packed &pk, const srt_tile_config &cfg, const ExecutorDesc *exec, int n_tiles,
                          Schedule &out, std::vector<RetiredSchedule> &retired, const srt_tile_config &out_cfg,
                          hipStream_t s) {
    if (!pk.cols && pk.total_cols)
        return fail(SRT_ERR_INVALID, "the plan's column indices were released: its schedule cannot change");
    Schedule sch;
    sch.unroll = schedule_unroll(exec, pow2_upto8(cfg.unroll) ? cfg.unroll : 4);
    sch.waves_per_block = cfg.waves_per_block;
    sch.nsplit = std::min(cfg.nsplit, cfg.waves_per_block);
    sch.chunk_cols = std::max(1, cfg.chunk_cols);
    sch.layout = cfg.layout;
    sch.n_tiles = n_tiles;
    sch.kp.assign(pk.P, 1);
    std::vector<int32_t> slots;  // (p, j, kp) per slot
    if (cfg.layout == 2) {
        balanced_slots(pk, sch, n_tiles, slots);
    } else {
        int used = 0;
        auto close_block = [&]() {
            for (; used > 0 && used < sch.waves_per_block; ++used) {
                slots.push_back(-1);
                slots.push_back(0);
                slots.push_back(1);
            }
            used = 0;
        };
        for (int p = 0; p < pk.P; ++p) {
            if (!pk.h_skip.empty() && pk.h_skip[p]) continue;
            const int64_t nc = pk.h_col_ptr[p + 1] - pk.h_col_ptr[p];
            const int64_t want = (nc + sch.chunk_cols - 1) / sch.chunk_cols;
            const int kp = (int)std::max<int64_t>(1, std::min<int64_t>(sch.nsplit, want));
            sch.kp[p] = kp;
            if (kp > 1) sch.reduce = 1;
            if (used + kp > sch.waves_per_block) close_block();
            for (int j = 0; j < kp; ++j) {
                slots.push_back(p);
                slots.push_back(j);
                slots.push_back(kp);
            }
            used += kp;
        }
        close_block();
    }
    const int nslots = (int)(slots.size() / 3);
    sch.n_wblocks = nslots / std::max(1, sch.waves_per_block);
    // Record run of each slot: one record per batch of its chunk plus a
    // terminator, runs laid out in slot order.  A chunk's batches are its
    // overlap with each (K tile, id) group cut into U-column pieces
    // (wave_prog_kernel), counted here from the per-panel counts.
    std::vector<int32_t> hc((size_t)pk.P * pk.KT * pk.NK);
    if (!hc.empty()) {
        SRT_HIP_TRY(hipMemcpyAsync(hc.data(), pk.counts, sizeof(int32_t) * hc.size(), hipMemcpyDeviceToHost, s));
        SRT_HIP_TRY(hipStreamSynchronize(s));
    }
    const int64_t U = sch.unroll;
    std::vector<int64_t> rec_start((size_t)std::max(nslots, 1), 0);
    int64_t nrun = 0;
    for (int i = 0; i < nslots; ++i) {
        rec_start[i] = nrun;
        const int p = slots[3 * i], j = slots[3 * i + 1], kp = slots[3 * i + 2];
        if (p >= 0) {
            const int64_t nc = pk.h_col_ptr[p + 1] - pk.h_col_ptr[p];
            const int64_t c0 = (int64_t)j * nc / kp, c1 = (int64_t)(j + 1) * nc / kp;
            const int32_t *cp = hc.data() + (size_t)p * pk.KT * pk.NK;
            int64_t pos = 0;
            for (int64_t g = 0; g < (int64_t)pk.KT * pk.NK && pos < c1; ++g) {
                const int64_t lo = std::max(pos, c0), hi = std::min(pos + cp[g], c1);
                if (hi > lo) nrun += (hi - lo + U - 1) / U;
                pos += cp[g];
            }
        }
        nrun += 1;
    }
    Scratch sc;
    int32_t *d_slots = nullptr;
    int64_t *d_start = nullptr;
    SRT_TRY(sc.get(&d_slots, slots.size()));
    SRT_TRY(sc.get(&d_start, rec_start.size()));
    SRT_TRY(dmalloc(&sch.desc, 16 * (size_t)std::max(nslots, 1)));
    if (nslots) {
        SRT_HIP_TRY(hipMemcpyAsync(d_slots, slots.data(), sizeof(int32_t) * slots.size(),
                                   hipMemcpyHostToDevice, s));
        SRT_HIP_TRY(launch_wave_desc(nslots, d_slots, pk.groups, pk.grp_ptr, pk.col_ptr,
                                     pk.id_pattern, sch.desc, s));
    }
    // Batch records: one per batch of each chunk plus a terminator per slot
    // (nrun in all), and two records of look-ahead padding.
    if (pk.total_vals + SRT_VAL_PAD > (int64_t)0x7fffffffLL)
        return fail(SRT_ERR_UNSUPPORTED, "more than 2^31 packed values");
    const size_t nrec = (size_t)nrun + 2;
    const int RS = rec_stride(sch.unroll);
    if (nrec * RS * sizeof(int32_t) > (size_t)0xffffffffULL)
        return fail(SRT_ERR_UNSUPPORTED, "batch records exceed 4 GiB");
    SRT_TRY(dmalloc(&sch.prog, nrec * RS));
    sch.bytes = sizeof(int32_t) * (nrec * RS + 16 * (size_t)std::max(nslots, 1));
    SRT_HIP_TRY(hipMemsetAsync(sch.prog, 0, sizeof(int32_t) * nrec * RS, s));
    if (nslots) {
        SRT_HIP_TRY(hipMemcpyAsync(d_start, rec_start.data(), sizeof(int64_t) * nslots, hipMemcpyHostToDevice, s));
        SRT_HIP_TRY(launch_wave_prog(nslots, sch.unroll, pk.groups, pk.grp_ptr, pk.col_ptr, pk.val_ptr, pk.cols,
                                     d_start, sch.desc, sch.prog, s));
    }
    SRT_HIP_TRY(hipStreamSynchronize(s));
    if (out.desc || out.prog) retired.push_back({out_cfg, std::move(out)});
    out = std::move(sch);
    return SRT_OK;
}

// Whether a schedule built for config b (n_tiles N tiles) serves config a at
// a_tiles N tiles (only the balanced layout depends on the tile count).
static bool same_schedule(const srt_tile_config &a, int a_tiles, const srt_tile_config &b, const Schedule &sch) {
    return a.nsplit == b.nsplit && a.chunk_cols == b.chunk_cols && a.waves_per_block == b.waves_per_block &&
           a.unroll == b.unroll && a.layout == b.layout && (a.layout != 2 || a_tiles == sch.n_tiles);
}

// Makes `cur` the schedule of config c: a retired schedule of the same
// shape comes back (the current one retires in its place), else a new one
// is built.  Every schedule stays alive until its owner is destroyed.
static int switch_schedule(const Packed &pk, const srt_tile_config &c, int tiles, const ExecutorDesc *exec,
                           Schedule &cur, srt_tile_config &cur_cfg, std::vector<RetiredSchedule> &retired,
                           hipStream_t s) {
    if (cur.desc && same_schedule(c, tiles, cur_cfg, cur)) return SRT_OK;
    for (RetiredSchedule &r : retired)
        if (same_schedule(c, tiles, r.cfg, r.sch)) {
            std::swap(r.sch, cur);
            std::swap(r.cfg, cur_cfg);
            if (!r.sch.desc && !r.sch.prog) {  // nothing was current: drop the empty entry
                r = std::move(retired.back());
                retired.pop_back();
            }
            return SRT_OK;
        }
    SRT_TRY(build_schedule(pk, c, exec, tiles, cur, retired, cur_cfg, s));
    cur_cfg = c;
    return SRT_OK;
}

// Column / value arrays of pk.total_cols / pk.total_vals entries plus the
// zeroed tail padding the executor's batch reads rely on.
static int alloc_packed_arrays(Packed &pk, hipStream_t s) {
    SRT_TRY(dmalloc(&pk.cols, (size_t)pk.total_cols + SRT_COL_PAD));
    SRT_TRY(dmalloc(&pk.vals, (size_t)pk.total_vals + SRT_VAL_PAD));
    SRT_HIP_TRY(hipMemsetAsync(pk.cols + pk.total_cols, 0, sizeof(int32_t) * SRT_COL_PAD, s));
    SRT_HIP_TRY(hipMemsetAsync(pk.vals + pk.total_vals, 0, sizeof(float) * SRT_VAL_PAD, s));
    return SRT_OK;
}

// K tile of the packing for the tiling strategy (the reference's enum,
// cpp_testbed/src/spmm/GECSB/SpMM_GECSB.h:21-25):
//   0 manual  - K_c as given (0 = one tile);
//   1 CAKE    - CAKE sizes the K block so one core's B block stays in its
//               cache (get_cache_dims_4, SpMM_GECSB.h:316-335; the CAKE
//               library itself is external); here the block is a
//               workgroup's B tile, K_c rows x 64 * vec_width columns, held
//               to the LDS of one workgroup (a multiple of 32 rows);
//   2 CAKE + TLB compensation - as 1, then K_c = max_tlb_entries *
//               tlb_page_size / (n * 4) whenever the B block n x K_c fp32
//               needs more than max_tlb_entries pages (SpMM_GECSB.h:339-348).
static int tiled_k(const srt_tile_config &c, int k, int n) {
    int K_c = c.K_c;
    if (c.tiling_strategy == 1 || c.tiling_strategy == 2) {
        const int V = (c.vec_width == 1 || c.vec_width == 2 || c.vec_width == 4) ? c.vec_width : 4;
        K_c = (int)std::max<size_t>(32, lds_limit() / ((size_t)64 * V * sizeof(float)) / 32 * 32);
    }
    if (c.tiling_strategy == 2 && n > 0 && c.tlb_page_size > 0 && c.max_tlb_entries > 0) {
        const int kc = (K_c >= 1 && K_c < k) ? K_c : k;
        const int64_t bc_bytes = (int64_t)n * kc * (int64_t)sizeof(float);
        if (bc_bytes / c.tlb_page_size > c.max_tlb_entries)
            K_c = (int)std::max<int64_t>(1, (int64_t)c.max_tlb_entries * c.tlb_page_size / ((int64_t)n * sizeof(float)));
    }
    return (K_c >= 1 && K_c < k) ? K_c : 0;
}

// Packs an m x k CSR matrix (device arrays) into pk.  Mirrors the packing
// done by the MatMulSpecialized constructor (SpMM_SOP.h:117-124).
static int pack_csr(Packed &pk, MappingImpl &map, int m, int k, int K_c, const float *Lx, const void *Lp,
                    const void *Li, int index_bytes, hipStream_t s) {
    DeviceLut *lut = nullptr;
    SRT_TRY(map.device_lut(&lut));
    // The row pointer is walked by the key kernel: check it on the host
    // first (non-negative start, non-decreasing).
    int64_t rp_first = 0, rp_last = 0;
    {
        std::vector<int64_t> rp(m + 1);
        if (index_bytes == 8) {
            SRT_HIP_TRY(hipMemcpyAsync(rp.data(), Lp, 8 * (size_t)(m + 1), hipMemcpyDeviceToHost, s));
            SRT_HIP_TRY(hipStreamSynchronize(s));
        } else {
            std::vector<int32_t> rp32(m + 1);
            SRT_HIP_TRY(hipMemcpyAsync(rp32.data(), Lp, 4 * (size_t)(m + 1), hipMemcpyDeviceToHost, s));
            SRT_HIP_TRY(hipStreamSynchronize(s));
            for (int i = 0; i <= m; ++i) rp[i] = rp32[i];
        }
        if (rp[0] < 0) return fail(SRT_ERR_INVALID, "CSR row pointer starts below zero");
        for (int i = 0; i < m; ++i)
            if (rp[i + 1] < rp[i]) return fail(SRT_ERR_INVALID, "CSR row pointer decreases at row " + std::to_string(i));
        if (rp[m] > rp[0] && (!Lx || !Li)) return fail(SRT_ERR_INVALID, "null CSR value or index array");
        rp_first = rp[0];
        rp_last = rp[m];
    }
    const int M_r = map.M_r;
    const int NK = (int)map.patterns.size();
    const int P = (m + M_r - 1) / M_r;
    // K tiling (the reference's K_c tiles, count_nkerns.cpp:52-126): panel p
    // is packed as KT segments of K_c columns, segment after segment, so its
    // chain runs tile by tile exactly as K-tiled execution with load_c would.
    const int Kc = (K_c >= 1 && K_c < k) ? K_c : std::max(k, 1);
    const int KT = k > 0 ? (k + Kc - 1) / Kc : 1;
    const int Q = P * KT;
    const int64_t ld = (int64_t)KT * Kc;  // padded row of the staging arrays
    pk.M = m;
    pk.K = k;
    pk.M_r = M_r;
    pk.P = P;
    pk.NK = NK;
    pk.KT = KT;
    pk.K_c = k > 0 ? Kc : 0;
    // Memory proportional to the nonzeros: every CSR entry becomes a key
    // (its cell = panel * k + column), the keys are radix-sorted stably (a
    // cell's entries stay in row order), and the packer walks the cells.
    const int64_t nnz = rp_last - rp_first;
    if (nnz > (int64_t)0x7fffffffLL) return fail(SRT_ERR_UNSUPPORTED, "more than 2^31 - 1 nonzeros");
    Scratch sc;
    int *bad = nullptr;
    int64_t *ncols = nullptr, *nvals = nullptr, *pad = nullptr, *seg_col = nullptr, *seg_val = nullptr;
    int64_t *seg_cells = nullptr;
    SRT_TRY(sc.get(&bad, 1));
    SRT_TRY(sc.get(&ncols, Q));
    SRT_TRY(sc.get(&nvals, Q));
    SRT_TRY(sc.get(&pad, Q));
    SRT_TRY(sc.get(&seg_cells, (size_t)Q + 1));
    SRT_HIP_TRY(hipMemsetAsync(bad, 0, sizeof(int), s));
    SRT_TRY(dmalloc(&pk.counts, (size_t)Q * NK));
    SRT_TRY(dmalloc(&pk.col_ptr, (size_t)P + 1));
    SRT_TRY(dmalloc(&pk.val_ptr, (size_t)P + 1));
    if (KT > 1) {
        SRT_TRY(sc.get(&seg_col, (size_t)Q + 1));
        SRT_TRY(sc.get(&seg_val, (size_t)Q + 1));
    } else {
        seg_col = pk.col_ptr;
        seg_val = pk.val_ptr;
    }
    const size_t n = (size_t)std::max<int64_t>(nnz, 1);
    uint64_t *keys = nullptr, *keys_alt = nullptr, *cell_key = nullptr;
    uint32_t *idx = nullptr, *idx_alt = nullptr, *head = nullptr, *cell_of = nullptr, *cell_code = nullptr;
    uint8_t *ent_row = nullptr;
    int64_t *cell_first = nullptr;
    SRT_TRY(sc.get(&keys, n));
    SRT_TRY(sc.get(&keys_alt, n));
    SRT_TRY(sc.get(&idx, n));
    SRT_TRY(sc.get(&idx_alt, n));
    SRT_TRY(sc.get(&ent_row, n));
    SRT_TRY(sc.get(&head, n));
    SRT_TRY(sc.get(&cell_of, n));
    SRT_TRY(sc.get(&cell_key, n));
    SRT_TRY(sc.get(&cell_code, n));
    SRT_TRY(sc.get(&cell_first, n + 1));
    SRT_HIP_TRY(launch_cell_keys(m, k, M_r, Lp, Li, index_bytes, keys, idx, ent_row, bad, s));
    int h_bad = 0;
    SRT_HIP_TRY(hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    SRT_HIP_TRY(hipStreamSynchronize(s));
    if (h_bad) return fail(SRT_ERR_INVALID, "CSR column index out of range [0, k)");
    int64_t ncells = 0;
    if (nnz > 0) {
        int end_bit = 1;  // bits of the largest key, P * k - 1
        while (end_bit < 64 && ((uint64_t)P * (uint64_t)k - 1) >> end_bit) ++end_bit;
        size_t sort_bytes = 0, scan_bytes = 0;
        bool in_alt = false;
        SRT_HIP_TRY(sort_cell_keys(nnz, end_bit, keys, keys_alt, idx, idx_alt, nullptr, &sort_bytes, &in_alt, s));
        SRT_HIP_TRY(scan_cell_heads(nnz, head, cell_of, nullptr, &scan_bytes, s));
        char *temp = nullptr;
        SRT_TRY(sc.get(&temp, std::max<size_t>(std::max(sort_bytes, scan_bytes), 1)));
        SRT_HIP_TRY(sort_cell_keys(nnz, end_bit, keys, keys_alt, idx, idx_alt, temp, &sort_bytes, &in_alt, s));
        if (in_alt) {
            std::swap(keys, keys_alt);
            std::swap(idx, idx_alt);
        }
        SRT_HIP_TRY(launch_cell_heads(nnz, keys, idx, ent_row, head, bad, s));
        SRT_HIP_TRY(scan_cell_heads(nnz, head, cell_of, temp, &scan_bytes, s));
        SRT_HIP_TRY(launch_cell_fill(nnz, keys, idx, ent_row, head, cell_of, cell_key, cell_code, cell_first, s));
        uint32_t last_of = 0, last_head = 0;
        SRT_HIP_TRY(hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
        SRT_HIP_TRY(hipMemcpyAsync(&last_of, cell_of + nnz - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        SRT_HIP_TRY(hipMemcpyAsync(&last_head, head + nnz - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        SRT_HIP_TRY(hipStreamSynchronize(s));
        if (h_bad) return fail(SRT_ERR_INVALID, "CSR holds an entry twice (same row and column)");
        ncells = (int64_t)last_of + last_head;
    }
    SRT_HIP_TRY(launch_segment_bounds(Q, KT, k, Kc, ncells, cell_key, seg_cells, s));
    SRT_HIP_TRY(launch_count_cells(Q, *lut, seg_cells, cell_code, pk.counts, ncols, nvals, s));
    SRT_HIP_TRY(launch_exclusive_scan(Q, ncols, seg_col, s));
    SRT_HIP_TRY(launch_exclusive_scan(Q, nvals, seg_val, s));
    int64_t tot_c = 0, tot_v = 0;
    SRT_HIP_TRY(hipMemcpyAsync(&tot_c, seg_col + Q, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SRT_HIP_TRY(hipMemcpyAsync(&tot_v, seg_val + Q, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SRT_HIP_TRY(hipStreamSynchronize(s));
    pk.nnz = nnz;
    pk.total_cols = tot_c;
    pk.total_vals = tot_v;
    SRT_TRY(alloc_packed_arrays(pk, s));
    SRT_HIP_TRY(launch_fill_cells(Q, k, KT, *lut, seg_cells, cell_key, cell_code, cell_first, idx, ent_row,
                                  Lx ? Lx + rp_first : nullptr, seg_col, seg_val, pk.counts, pk.cols, pk.vals,
                                  pad, s));
    if (KT > 1) {  // panel offsets = every KT-th segment offset
        SRT_HIP_TRY(hipMemcpy2DAsync(pk.col_ptr, sizeof(int64_t), seg_col, sizeof(int64_t) * KT,
                                     sizeof(int64_t), (size_t)P + 1, hipMemcpyDeviceToDevice, s));
        SRT_HIP_TRY(hipMemcpy2DAsync(pk.val_ptr, sizeof(int64_t), seg_val, sizeof(int64_t) * KT,
                                     sizeof(int64_t), (size_t)P + 1, hipMemcpyDeviceToDevice, s));
    }
    std::vector<int64_t> h_pad(Q);
    if (Q) SRT_HIP_TRY(hipMemcpyAsync(h_pad.data(), pad, sizeof(int64_t) * Q, hipMemcpyDeviceToHost, s));
    SRT_HIP_TRY(hipStreamSynchronize(s));
    pk.padding = 0;
    for (int64_t v : h_pad) pk.padding += v;
    return build_groups(pk, lut->id_pattern, s);
}

}  // namespace srt

extern "C" {

// Matrix-core execution: the plan's mapping must be the full 4-row pattern.
static bool full_pattern(const ExecutorDesc *e) {
    return e && e->M_r == 4 && e->num_patterns == 1 && e->patterns[0] == 15;
}

// Panel ranges of the matrix-core executor for B width n: R contiguous
// ranges balanced by packed columns (+4 per panel for its C rows), R x
// (64-column slices) = one round of 256 CUs; the ranges of a width used
// before are reused, earlier arrays stay alive until the plan is destroyed.
static int mfma_ranges_for(srt_plan *p, int n) {
    const int S = (std::max(n, 1) + 63) / 64;
    const int P = p->pk.P;
    const int R = std::max(1, std::min(P, (256 + S - 1) / S));
    for (size_t i = 0; i < p->mfma_ranges.size(); ++i)
        if (p->mfma_ranges[i].first == R) {
            std::swap(p->mfma_ranges[i], p->mfma_ranges.back());
            return SRT_OK;
        }
    const std::vector<int64_t> &cp = p->pk.h_col_ptr;
    if ((int)cp.size() != P + 1) return fail(SRT_ERR_INVALID, "plan has no host column pointer");
    auto weight = [&](int q) { return (cp[q] - cp[0]) + 4 * (int64_t)q; };
    const int64_t total = weight(P);
    std::vector<int32_t> rg((size_t)R + 1, 0);
    int q = 0;
    for (int j = 1; j < R; ++j) {
        const int64_t target = total * j / R;
        while (q < P && weight(q) < target) ++q;
        rg[j] = q;
    }
    rg[R] = P;
    int32_t *d = nullptr;
    SRT_HIP_TRY(hipMalloc(&d, sizeof(int32_t) * rg.size()));
    hipError_t e = hipMemcpy(d, rg.data(), sizeof(int32_t) * rg.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        SRT_HIP_TRY(e);
    }
    p->mfma_ranges.emplace_back(R, d);
    return SRT_OK;
}

int srt_plan_create(int m, int k, int n, const float *Lx, const void *Lp, const void *Li,
                    int index_bytes, const srt_tile_config *cfg, int nThreads,
                    const char *executor_id, const srt_mapping *mapping, srt_stream_t stream,
                    srt_plan **out) {
    (void)nThreads;
    if (!out || !mapping) return fail(SRT_ERR_INVALID, "srt_plan_create: null argument");
    *out = nullptr;
    if (m < 0 || k < 0 || n < 0) return fail(SRT_ERR_INVALID, "negative dimension");
    if (index_bytes != 4 && index_bytes != 8) return fail(SRT_ERR_INVALID, "index_bytes must be 4 or 8");
    if (!Lp) return fail(SRT_ERR_INVALID, "null CSR row pointer");
    MappingImpl &map = *mapping->impl;
    const ExecutorDesc *exec = executor_id && executor_id[0] ? find_executor_by_id(executor_id)
                                                             : find_executor(map.M_r, map.patterns);
    if (!exec)
        return fail(SRT_ERR_UNSUPPORTED, executor_id && executor_id[0]
                                             ? "unknown executor id " + std::string(executor_id)
                                             : "no CDNA4 executor generated for this pattern set");
    if (exec->M_r != map.M_r || exec->num_patterns != (int)map.patterns.size() ||
        !std::equal(map.patterns.begin(), map.patterns.end(), exec->patterns))
        return fail(SRT_ERR_INVALID, "executor and mapping pattern sets differ");
    auto plan = std::make_unique<srt_plan>();
    SRT_HIP_TRY(hipGetDevice(&plan->device));
    const int K_c = cfg ? tiled_k(*cfg, k, n) : 0;
    int rc = pack_csr(plan->pk, map, m, k, K_c, Lx, Lp, Li, index_bytes, (hipStream_t)stream);
    if (rc) {
        plan->pk.release();
        return rc;
    }
    plan->exec = exec;
    if (cfg) plan->cfg_user = *cfg;
    else srt_tile_config_default(&plan->cfg_user);
    plan->n_max = n;
    plan->cfg = plan->cfg_user;
    realise_config(plan->cfg, map.M_r, k, plan->pk.K_c, plan->pk.total_cols, std::max(n, 1));
    if (plan->cfg.matrix_cores == 1) {
        if (!full_pattern(exec))
            rc = fail(SRT_ERR_UNSUPPORTED, "matrix_cores needs M_r = 4 and a mapping whose only kernel pattern is 15");
        else if (k > mfma4_max_k())
            rc = fail(SRT_ERR_UNSUPPORTED, "matrix_cores stages all K rows of a 64-column B slice in LDS: K <= " +
                                               std::to_string(mfma4_max_k()));
        else
            rc = mfma_ranges_for(plan.get(), n);
        if (rc) {
            plan->pk.release();
            for (auto &r : plan->mfma_ranges) (void)hipFree(r.second);
            return rc;
        }
    }
    rc = build_schedule(plan->pk, plan->cfg, plan->exec, config_tiles(plan->cfg, std::max(n, 1)), plan->sch,
                        plan->retired, plan->cfg, (hipStream_t)stream);
    if (rc) {
        plan->pk.release();
        for (auto &r : plan->mfma_ranges) (void)hipFree(r.second);
        return rc;
    }
    *out = plan.release();
    return SRT_OK;
}

int srt_plan_create_hybrid(int m, int k, int n, const float *Lx, const void *Lp, const void *Li, int index_bytes,
                           const srt_tile_config *cfg, float dense_threshold, const srt_mapping *mapping,
                           srt_stream_t stream, srt_plan **out) {
    if (!out || !mapping || !Lp) return fail(SRT_ERR_INVALID, "srt_plan_create_hybrid: null argument");
    *out = nullptr;
    if (m < 0 || k < 0 || n < 0) return fail(SRT_ERR_INVALID, "negative dimension");
    if (index_bytes != 4 && index_bytes != 8) return fail(SRT_ERR_INVALID, "index_bytes must be 4 or 8");
    if (!(dense_threshold >= 0.0f)) return fail(SRT_ERR_INVALID, "dense_threshold must be >= 0");
    hipStream_t s = (hipStream_t)stream;
    // The CSR on the host: row pointer (checked as pack_csr checks it), then
    // the entries.
    std::vector<int64_t> rp((size_t)m + 1);
    if (index_bytes == 8) {
        SRT_HIP_TRY(hipMemcpyAsync(rp.data(), Lp, 8 * ((size_t)m + 1), hipMemcpyDeviceToHost, s));
    } else {
        std::vector<int32_t> r32((size_t)m + 1);
        SRT_HIP_TRY(hipMemcpyAsync(r32.data(), Lp, 4 * ((size_t)m + 1), hipMemcpyDeviceToHost, s));
        SRT_HIP_TRY(hipStreamSynchronize(s));
        std::copy(r32.begin(), r32.end(), rp.begin());
    }
    SRT_HIP_TRY(hipStreamSynchronize(s));
    if (rp[0] < 0) return fail(SRT_ERR_INVALID, "CSR row pointer starts below zero");
    for (int i = 0; i < m; ++i)
        if (rp[i + 1] < rp[i]) return fail(SRT_ERR_INVALID, "CSR row pointer decreases at row " + std::to_string(i));
    const int64_t base = rp[0], nnz = rp[m] - rp[0];
    if (nnz > 0 && (!Lx || !Li)) return fail(SRT_ERR_INVALID, "null CSR value or index array");
    // Dense panels: 16-row panels whose density reaches the threshold.
    const int P16 = (m + 15) / 16;
    std::vector<int32_t> dense_rows;
    std::vector<char> is_dense((size_t)P16, 0);
    for (int p = 0; p < P16 && k > 0; ++p) {
        const int r0 = 16 * p, r1 = std::min(m, r0 + 16);
        const double density = (double)(rp[r1] - rp[r0]) / ((double)(r1 - r0) * k);
        if (density >= dense_threshold && rp[r1] > rp[r0]) {
            is_dense[p] = 1;
            dense_rows.push_back(r0);
        }
    }
    if (dense_rows.empty())
        return srt_plan_create(m, k, n, Lx, Lp, Li, index_bytes, cfg, 1, nullptr, mapping, stream, out);
    std::vector<int64_t> ci((size_t)std::max<int64_t>(nnz, 1));
    std::vector<float> cv((size_t)std::max<int64_t>(nnz, 1));
    if (nnz > 0) {
        if (index_bytes == 8) {
            SRT_HIP_TRY(hipMemcpyAsync(ci.data(), (const int64_t *)Li + base, 8 * (size_t)nnz, hipMemcpyDeviceToHost, s));
        } else {
            std::vector<int32_t> c32((size_t)nnz);
            SRT_HIP_TRY(hipMemcpyAsync(c32.data(), (const int32_t *)Li + base, 4 * (size_t)nnz, hipMemcpyDeviceToHost, s));
            SRT_HIP_TRY(hipStreamSynchronize(s));
            std::copy(c32.begin(), c32.end(), ci.begin());
        }
        SRT_HIP_TRY(hipMemcpyAsync(cv.data(), Lx + base, 4 * (size_t)nnz, hipMemcpyDeviceToHost, s));
        SRT_HIP_TRY(hipStreamSynchronize(s));
    }
    // The register-tiled part: the CSR with the dense panels' rows emptied.
    std::vector<int64_t> rp2((size_t)m + 1, 0), ci2;
    std::vector<float> cv2;
    for (int i = 0; i < m; ++i) {
        if (!is_dense[i / 16])
            for (int64_t q = rp[i] - base; q < rp[i + 1] - base; ++q) {
                ci2.push_back(ci[q]);
                cv2.push_back(cv[q]);
            }
        rp2[i + 1] = (int64_t)ci2.size();
    }
    // The dense part in MFMA A-operand order: A[row = l % 16][k = 16 g + 4 j + l / 16]
    // at ((d * G + g) * 64 + l) * 4 + j.
    const int nd = (int)dense_rows.size(), G = (k + 15) / 16;
    std::vector<float> ap((size_t)nd * G * 64 * 4, 0.0f);
    std::vector<char> seen(ap.size(), 0);
    for (int d = 0; d < nd; ++d)
        for (int r = 0; r < 16 && dense_rows[d] + r < m; ++r) {
            const int row = dense_rows[d] + r;
            for (int64_t q = rp[row] - base; q < rp[row + 1] - base; ++q) {
                const int64_t col = ci[q];
                if (col < 0 || col >= k) return fail(SRT_ERR_INVALID, "CSR column index out of range [0, k)");
                const int g = (int)(col / 16), w = (int)(col % 16);
                const size_t at = (((size_t)d * G + g) * 64 + (size_t)(r + 16 * (w % 4))) * 4 + (size_t)(w / 4);
                if (seen[at]) return fail(SRT_ERR_INVALID, "CSR holds an entry twice (same row and column)");
                seen[at] = 1;
                ap[at] = cv[q];
            }
        }
    int64_t *d_rp = nullptr, *d_ci = nullptr;
    float *d_cv = nullptr;
    Scratch sc;
    SRT_TRY(sc.get(&d_rp, rp2.size()));
    SRT_TRY(sc.get(&d_ci, std::max<size_t>(ci2.size(), 1)));
    SRT_TRY(sc.get(&d_cv, std::max<size_t>(cv2.size(), 1)));
    SRT_HIP_TRY(hipMemcpyAsync(d_rp, rp2.data(), 8 * rp2.size(), hipMemcpyHostToDevice, s));
    if (!ci2.empty()) {
        SRT_HIP_TRY(hipMemcpyAsync(d_ci, ci2.data(), 8 * ci2.size(), hipMemcpyHostToDevice, s));
        SRT_HIP_TRY(hipMemcpyAsync(d_cv, cv2.data(), 4 * cv2.size(), hipMemcpyHostToDevice, s));
    }
    srt_plan *plan = nullptr;
    SRT_TRY(srt_plan_create(m, k, n, d_cv, d_rp, d_ci, 8, cfg, 1, nullptr, mapping, stream, &plan));
    std::unique_ptr<srt_plan, void (*)(srt_plan *)> guard(plan, srt_plan_destroy);
    DensePart &dp = plan->dense;
    dp.ndense = nd;
    dp.G = G;
    dp.h_panel_row = dense_rows;
    dp.bytes = sizeof(float) * ap.size() + sizeof(int32_t) * (size_t)nd;
    SRT_HIP_TRY(hipMalloc(&dp.apack, sizeof(float) * ap.size()));
    SRT_HIP_TRY(hipMalloc((void **)&dp.panel_row, sizeof(int32_t) * (size_t)nd));
    SRT_HIP_TRY(hipMemcpyAsync(dp.apack, ap.data(), sizeof(float) * ap.size(), hipMemcpyHostToDevice, s));
    SRT_HIP_TRY(hipMemcpyAsync(dp.panel_row, dense_rows.data(), sizeof(int32_t) * (size_t)nd, hipMemcpyHostToDevice, s));
    SRT_HIP_TRY(hipStreamSynchronize(s));
    SRT_TRY(dp.reserve(n));
    // With 16-row dense panels made of whole register-tile panels (M_r
    // dividing 16) and some rows left to the register tile, the register-
    // tiled schedule is rebuilt without the dense panels' rows (nothing can
    // have captured the first one yet): the two launches write disjoint rows.
    // (Run concurrently on two streams they slowed each other down: each
    // holds whole CUs, so the dense launch stretched from 17 to 63 us and
    // the register tile from 57 to 114 us on 1024 x 4096 at 70 %.)
    const bool all_dense = (int64_t)nd * 16 >= m;
    Packed &pk = plan->pk;
    if (nd > 0 && !all_dense && 16 % pk.M_r == 0 && plan->cfg.matrix_cores != 1) {
        pk.h_skip.assign((size_t)pk.P, 0);
        for (int d = 0; d < nd; ++d)
            for (int r = dense_rows[d]; r < std::min(m, dense_rows[d] + 16); r += pk.M_r) pk.h_skip[(size_t)(r / pk.M_r)] = 1;
        Schedule first = std::move(plan->sch);
        plan->sch = Schedule();
        const int rc = build_schedule(pk, plan->cfg, plan->exec, config_tiles(plan->cfg, std::max(n, 1)), plan->sch,
                                      plan->retired, plan->cfg, s);
        first.release();
        SRT_TRY(rc);
    }
    plan->pk.nnz = nnz;  // the whole matrix's entries (the register-tiled part holds the others)
    *out = guard.release();
    return SRT_OK;
}

int srt_plan_dense_panels(const srt_plan *p, int *num_dense, int32_t *panel_rows, size_t *bytes) {
    if (!p) return fail(SRT_ERR_INVALID, "null plan");
    if (num_dense) *num_dense = p->dense.ndense;
    if (panel_rows) std::copy(p->dense.h_panel_row.begin(), p->dense.h_panel_row.end(), panel_rows);
    if (bytes) *bytes = p->dense.bytes;
    return SRT_OK;
}

int srt_plan_get_config(const srt_plan *p, srt_tile_config *out) {
    if (!p || !out) return fail(SRT_ERR_INVALID, "null argument");
    *out = p->cfg;
    return SRT_OK;
}

int srt_plan_required_storage(const srt_plan *p, size_t *bytes) {
    if (!p || !bytes) return fail(SRT_ERR_INVALID, "null argument");
    *bytes = p->pk.bytes();
    return SRT_OK;
}

int srt_plan_allocate_executor(srt_plan *p, int n) {
    if (!p) return fail(SRT_ERR_INVALID, "null plan");
    if (n < 0) return fail(SRT_ERR_INVALID, "n must be >= 0");
    p->n_alloc = n;
    if (n > p->n_max) p->n_max = n;
    if (p->dense.ndense) SRT_TRY(p->dense.reserve(n));
    // Re-realise the choices that depend on the B width and switch to that
    // schedule (schedules of other widths stay alive: graphs captured with
    // them replay unchanged).  Not concurrently with other calls on the plan.
    srt_tile_config c = p->cfg_user;
    realise_config(c, p->pk.M_r, p->pk.K, p->pk.K_c, p->pk.total_cols, std::max(n, 1));
    const int tiles = config_tiles(c, std::max(n, 1));
    srt_tile_config sch_cfg = p->cfg;
    SRT_TRY(switch_schedule(p->pk, c, tiles, p->exec, p->sch, sch_cfg, p->retired, nullptr));
    p->cfg = c;
    if (c.matrix_cores == 1) SRT_TRY(mfma_ranges_for(p, n));
    return SRT_OK;
}

int srt_plan_execute(const srt_plan *p, float *C, int64_t ldc, const float *B, int64_t ldb, int n,
                     srt_stream_t stream) {
    if (!p) return fail(SRT_ERR_INVALID, "null plan");
    int dev = -1;
    SRT_HIP_TRY(hipGetDevice(&dev));
    if (dev != p->device) return fail(SRT_ERR_INVALID, "plan belongs to another device");
    // With every 16-row panel dense the register-tiled part is empty and the
    // dense panels write every row of C.
    const bool all_dense = p->dense.ndense > 0 && (int64_t)p->dense.ndense * 16 >= p->pk.M;
    if (all_dense || p->cfg.matrix_cores == 1) {
        if (n < 0) return fail(SRT_ERR_INVALID, "n must be >= 0");
        if (n > 0 && (!B || !C)) return fail(SRT_ERR_INVALID, "null B or C");
        if (n > 0 && (ldb < n || ldc < n)) return fail(SRT_ERR_INVALID, "ldb and ldc must be >= n");
        if (!all_dense && n > 0 && p->pk.M > 0) {
            if (!p->pk.cols) return fail(SRT_ERR_INVALID, "matrix-core execution needs the column indices (released)");
            if (p->mfma_ranges.empty()) return fail(SRT_ERR_INVALID, "no matrix-core panel ranges");
            const auto &rg = p->mfma_ranges.back();
            SRT_HIP_TRY(launch_mfma4_exec(p->pk.col_ptr, p->pk.cols, p->pk.vals, rg.second, rg.first, B, ldb, C,
                                          ldc, p->pk.M, n, p->pk.K, (hipStream_t)stream));
        }
    } else
        SRT_TRY(run_executor(p->pk, p->sch, p->exec, p->cfg, B, ldb, C, ldc, n, 0, (hipStream_t)stream));
    // the dense panels' rows (left out by the register-tiled schedule, or
    // written by it as zero and overwritten here)
    if (p->dense.ndense && n > 0)
        SRT_HIP_TRY(launch_dense_panels(p->dense.ndense, p->dense.G, p->dense.apack, p->dense.panel_row, B, ldb, C,
                                        ldc, p->pk.M, n, p->pk.K, p->dense.ws, p->dense.ws_bytes,
                                        (hipStream_t)stream));
    return SRT_OK;
}

int srt_plan_release_columns(srt_plan *p) {
    if (!p) return fail(SRT_ERR_INVALID, "null plan");
    if (p->cfg.matrix_cores == 1)
        return fail(SRT_ERR_UNSUPPORTED, "matrix-core execution reads the column indices at every launch");
    SRT_HIP_TRY(hipDeviceSynchronize());  // no launch that builds a schedule may still read them
    (void)hipFree(p->pk.cols);
    p->pk.cols = nullptr;
    return SRT_OK;
}

int srt_plan_schedule_bytes(const srt_plan *p, size_t *bytes) {
    if (!p || !bytes) return fail(SRT_ERR_INVALID, "null argument");
    *bytes = p->sch.bytes;
    return SRT_OK;
}

int srt_plan_launch_shape(const srt_plan *p, int n, int *grid, int *block) {
    if (!p || n < 0) return fail(SRT_ERR_INVALID, "null plan or negative n");
    const int V = std::max(1, p->cfg.vec_width);
    const int64_t tiles = ((int64_t)n + 64 * V - 1) / (64 * V);
    if (grid) *grid = (int)std::min<int64_t>((int64_t)p->sch.n_wblocks * tiles, 0x7fffffff);
    if (block) *block = 64 * p->sch.waves_per_block;
    return SRT_OK;
}

int srt_plan_split_counts(const srt_plan *p, int32_t *kp) {
    if (!p || !kp) return fail(SRT_ERR_INVALID, "null argument");
    std::copy(p->sch.kp.begin(), p->sch.kp.end(), kp);
    return SRT_OK;
}

int srt_plan_stats_get(const srt_plan *p, srt_plan_stats *o) {
    if (!p || !o) return fail(SRT_ERR_INVALID, "null argument");
    o->M = p->pk.M;
    o->K = p->pk.K;
    o->M_r = p->pk.M_r;
    o->num_panels = p->pk.P;
    o->num_patterns = p->pk.NK;
    o->nnz = p->pk.nnz;
    o->packed_cols = p->pk.total_cols;
    o->packed_values = p->pk.total_vals;
    o->padding = p->pk.padding;
    o->k_tile = p->pk.K_c;
    o->k_tiles = p->pk.KT;
    return SRT_OK;
}

static int copy_packed(const Packed &pk, int32_t *counts, int64_t *col_ptr, int64_t *val_ptr,
                       int32_t *cols, float *vals) {
    if (counts && pk.P && pk.NK)
        SRT_HIP_TRY(hipMemcpy(counts, pk.counts, sizeof(int32_t) * pk.P * pk.KT * pk.NK, hipMemcpyDeviceToHost));
    if (col_ptr) SRT_HIP_TRY(hipMemcpy(col_ptr, pk.col_ptr, sizeof(int64_t) * (pk.P + 1), hipMemcpyDeviceToHost));
    if (val_ptr) SRT_HIP_TRY(hipMemcpy(val_ptr, pk.val_ptr, sizeof(int64_t) * (pk.P + 1), hipMemcpyDeviceToHost));
    if (cols && pk.total_cols) {
        if (!pk.cols) return fail(SRT_ERR_INVALID, "the plan's column indices were released");
        SRT_HIP_TRY(hipMemcpy(cols, pk.cols, sizeof(int32_t) * pk.total_cols, hipMemcpyDeviceToHost));
    }
    if (vals && pk.total_vals)
        SRT_HIP_TRY(hipMemcpy(vals, pk.vals, sizeof(float) * pk.total_vals, hipMemcpyDeviceToHost));
    return SRT_OK;
}

int srt_plan_copy_packed(const srt_plan *p, int32_t *counts, int64_t *col_ptr, int64_t *val_ptr,
                         int32_t *cols, float *vals) {
    if (!p) return fail(SRT_ERR_INVALID, "null plan");
    return copy_packed(p->pk, counts, col_ptr, val_ptr, cols, vals);
}

void srt_plan_destroy(srt_plan *p) {
    if (!p) return;
    p->pk.release();
    p->sch.release();
    for (RetiredSchedule &r : p->retired) r.sch.release();
    p->dense.release();
    for (auto &r : p->mfma_ranges) (void)hipFree(r.second);
    delete p;
}

// ------------------------------------------------------------ grouped plans
}  // extern "C"

struct srt_group {
    std::vector<const srt_plan *> plans;
    const ExecutorDesc *exec = nullptr;
    int waves = 0, V = 1, total_blocks = 0, xcd_remap = 1, device = 0;
    size_t lds = 0;
    ExecProblem *d_probs = nullptr;
    std::vector<ExecProblem> h_probs;
};

// Problem table of a group for the given operands: the lane width is the
// widest every problem allows (alignment, V | n, LDS of its reduction), each
// problem gets n_wblocks * ceil(n / 64V) consecutive workgroups.
static int group_bind(srt_group &g, const float *const *B, const int64_t *ldb, float *const *C, const int64_t *ldc,
                      const int *n, hipStream_t s) {
    if (!B || !ldb || !C || !ldc || !n) return fail(SRT_ERR_INVALID, "srt_group: null operand array");
    const int cnt = (int)g.plans.size();
    int V = 4;
    for (int i = 0; i < cnt; ++i) V = std::min(V, g.plans[i]->cfg.vec_width);
    for (int i = 0; i < cnt; ++i) {
        const srt_plan &p = *g.plans[i];
        if (n[i] < 0) return fail(SRT_ERR_INVALID, "srt_group: n must be >= 0");
        if (n[i] > 0 && p.pk.M > 0) {
            if (!B[i] || !C[i]) return fail(SRT_ERR_INVALID, "srt_group: null B or C");
            if (ldb[i] < n[i] || ldc[i] < n[i]) return fail(SRT_ERR_INVALID, "srt_group: ldb and ldc must be >= n");
            if ((int64_t)p.pk.K * ldb[i] * (int64_t)sizeof(float) > (int64_t)0xffffffffLL)
                return fail(SRT_ERR_UNSUPPORTED, "srt_group: B larger than 4 GiB (K * ldb * 4 bytes)");
            while (V > 1 && (n[i] % V || ldb[i] % V || ldc[i] % V || !aligned(B[i], 4 * V) || !aligned(C[i], 4 * V)))
                V >>= 1;
        }
        while (V > 1 && !lds_fits(false, p.pk.K, g.waves, p.pk.M_r, V, p.sch.reduce)) V >>= 1;
        if (!lds_fits(false, p.pk.K, g.waves, p.pk.M_r, V, p.sch.reduce))
            return fail(SRT_ERR_UNSUPPORTED, "srt_group: chunk reduction needs more LDS than a workgroup has");
    }
    std::vector<ExecProblem> t;
    int64_t total = 0;
    size_t lds = 0;
    for (int i = 0; i < cnt; ++i) {
        const srt_plan &p = *g.plans[i];
        if (n[i] == 0 || p.pk.M == 0 || p.sch.n_wblocks == 0) continue;
        ExecProblem q{};
        q.prog = p.sch.prog;
        q.desc = p.sch.desc;
        q.vals = p.pk.vals;
        q.B = B[i];
        q.C = C[i];
        q.ldc = ldc[i];
        q.b_bytes = (uint32_t)std::max<int64_t>((int64_t)p.pk.K * ldb[i] * (int64_t)sizeof(float), 4);
        q.ldb = (uint32_t)ldb[i];
        q.M = p.pk.M;
        q.N = n[i];
        q.K = p.pk.K;
        q.n_wblocks = p.sch.n_wblocks;
        q.reduce = p.sch.reduce;
        q.load_c = 0;
        q.block_begin = (int)total;
        total += (int64_t)p.sch.n_wblocks * ((n[i] + 64 * V - 1) / (64 * V));
        if (total > 0x7fffffffLL) return fail(SRT_ERR_UNSUPPORTED, "srt_group: more than 2^31 workgroups");
        if (p.sch.reduce) lds = std::max(lds, reduce_lds(g.waves, p.pk.M_r, V));
        t.push_back(q);
    }
    if (t.size() > g.h_probs.size() || !g.d_probs) {
        ExecProblem *d = nullptr;
        SRT_TRY(dmalloc(&d, std::max<size_t>(t.size(), 1)));
        hipFree(g.d_probs);
        g.d_probs = d;
    }
    if (!t.empty()) {
        SRT_HIP_TRY(hipMemcpyAsync(g.d_probs, t.data(), sizeof(ExecProblem) * t.size(), hipMemcpyHostToDevice, s));
        SRT_HIP_TRY(hipStreamSynchronize(s));
    }
    g.h_probs = std::move(t);
    g.V = V;
    g.total_blocks = (int)total;
    g.lds = lds;
    return SRT_OK;
}

extern "C" {

int srt_group_create(const srt_plan *const *plans, int count, const float *const *B, const int64_t *ldb,
                     float *const *C, const int64_t *ldc, const int *n, srt_stream_t stream, srt_group **out) {
    if (!out || !plans) return fail(SRT_ERR_INVALID, "srt_group_create: null argument");
    *out = nullptr;
    if (count < 1) return fail(SRT_ERR_INVALID, "srt_group_create: count must be >= 1");
    for (int i = 0; i < count; ++i)
        if (!plans[i]) return fail(SRT_ERR_INVALID, "srt_group_create: null plan");
    auto g = std::make_unique<srt_group>();
    SRT_HIP_TRY(hipGetDevice(&g->device));
    for (int i = 0; i < count; ++i) {
        const srt_plan *p = plans[i];
        if (p->device != g->device) return fail(SRT_ERR_INVALID, "srt_group_create: plan belongs to another device");
        if (p->dense.ndense) return fail(SRT_ERR_UNSUPPORTED, "srt_group_create: hybrid plans execute alone");
        if (i == 0) {
            g->exec = p->exec;
            g->waves = p->sch.waves_per_block;
            g->xcd_remap = p->cfg.xcd_remap;
        } else if (p->exec != g->exec) {
            return fail(SRT_ERR_INVALID, "srt_group_create: plans use different executors (mappings)");
        } else if (p->sch.waves_per_block != g->waves) {
            return fail(SRT_ERR_INVALID, "srt_group_create: plans use different waves_per_block");
        }
        if (p->sch.unroll != 4)
            return fail(SRT_ERR_UNSUPPORTED, "srt_group_create: grouped launches need unroll 4");
        g->plans.push_back(p);
    }
    if (!g->exec->grouped[0]) return fail(SRT_ERR_UNSUPPORTED, "srt_group_create: executor has no grouped variant");
    int rc = group_bind(*g, B, ldb, C, ldc, n, (hipStream_t)stream);
    if (rc) {
        hipFree(g->d_probs);
        return rc;
    }
    *out = g.release();
    return SRT_OK;
}

int srt_group_rebind(srt_group *g, const float *const *B, const int64_t *ldb, float *const *C, const int64_t *ldc,
                     const int *n, srt_stream_t stream) {
    if (!g) return fail(SRT_ERR_INVALID, "null group");
    return group_bind(*g, B, ldb, C, ldc, n, (hipStream_t)stream);
}

int srt_group_execute(const srt_group *g, srt_stream_t stream) {
    if (!g) return fail(SRT_ERR_INVALID, "null group");
    int dev = -1;
    SRT_HIP_TRY(hipGetDevice(&dev));
    if (dev != g->device) return fail(SRT_ERR_INVALID, "group belongs to another device");
    if (g->total_blocks == 0) return SRT_OK;
    ExecGroupedFn fn = g->exec->grouped[log2_small(g->V)];
    if (!fn) return fail(SRT_ERR_UNSUPPORTED, "grouped executor variant not generated");
    SRT_HIP_TRY(fn(g->d_probs, (int)g->h_probs.size(), g->total_blocks, g->waves, g->lds, g->xcd_remap,
                   (hipStream_t)stream));
    return SRT_OK;
}

// ---------------------------------------------------- persistent problem table
static_assert(sizeof(srt_exec_problem) == sizeof(ExecProblem), "srt_exec_problem mirrors ExecProblem");
static_assert(offsetof(srt_exec_problem, ldc) == offsetof(ExecProblem, ldc), "srt_exec_problem layout");
static_assert(offsetof(srt_exec_problem, block_begin) == offsetof(ExecProblem, block_begin),
              "srt_exec_problem layout");

// The persistent launchers are generated at this unroll (codegen GROUPED_UNROLL).
static constexpr int kPersistentUnroll = 4;

int srt_exec_problem_fill(const srt_plan *p, const float *B, int64_t ldb, float *C, int64_t ldc, int n,
                          int vec_width, int block_begin, srt_exec_problem *out) {
    if (!p || !out) return fail(SRT_ERR_INVALID, "srt_exec_problem_fill: null argument");
    const int V = vec_width;
    if (V != 1 && V != 2 && V != 4) return fail(SRT_ERR_INVALID, "vec_width must be 1, 2 or 4");
    if (n < 0 || ldb < n || ldc < n || block_begin < 0) return fail(SRT_ERR_INVALID, "bad n / ldb / ldc / block_begin");
    if (p->sch.unroll != kPersistentUnroll)
        return fail(SRT_ERR_UNSUPPORTED, "the plan's schedule is not cut for unroll 4 (TileConfig.unroll = 4)");
    if (p->dense.ndense) return fail(SRT_ERR_UNSUPPORTED, "hybrid plans execute alone");
    if (n > 0 && (!B || !C)) return fail(SRT_ERR_INVALID, "null B or C");
    // A lane's last vector reaches column round_up(n, V) - 1: the strides
    // must hold it (B's columns there are read, C's are written).
    const int64_t n_up = (int64_t)(n + V - 1) / V * V;
    if (ldb % V || ldc % V || ldb < n_up || ldc < n_up || !aligned(B, 4 * V) || !aligned(C, 4 * V))
        return fail(SRT_ERR_INVALID, "vec_width must divide ldb and ldc, both >= n rounded up to vec_width, and "
                                     "the operands be 4 * vec_width aligned");
    if ((int64_t)p->pk.K * ldb * (int64_t)sizeof(float) > (int64_t)0xffffffffLL)
        return fail(SRT_ERR_UNSUPPORTED, "B larger than 4 GiB (K * ldb * 4 bytes)");
    if (p->sch.reduce && !lds_fits(false, p->pk.K, p->sch.waves_per_block, p->pk.M_r, V, 1))
        return fail(SRT_ERR_UNSUPPORTED, "chunk reduction needs more LDS than a workgroup has");
    ExecProblem q{};
    q.prog = p->sch.prog;
    q.desc = p->sch.desc;
    q.vals = p->pk.vals;
    q.B = B;
    q.C = C;
    q.ldc = ldc;
    q.b_bytes = (uint32_t)std::max<int64_t>((int64_t)p->pk.K * ldb * (int64_t)sizeof(float), 4);
    q.ldb = (uint32_t)ldb;
    q.M = p->pk.M;
    q.N = n;
    q.K = p->pk.K;
    q.n_wblocks = p->sch.n_wblocks;
    q.reduce = p->sch.reduce;
    q.load_c = 0;
    q.block_begin = block_begin;
    std::memcpy(out, &q, sizeof(q));
    return SRT_OK;
}

int srt_exec_persistent_execute(const srt_plan *proto, int vec_width, const srt_exec_problem *d_probs,
                                const int *d_hdr, int grid, srt_stream_t stream) {
    if (!proto || !d_probs || !d_hdr) return fail(SRT_ERR_INVALID, "srt_exec_persistent_execute: null argument");
    const int V = vec_width;
    if (V != 1 && V != 2 && V != 4) return fail(SRT_ERR_INVALID, "vec_width must be 1, 2 or 4");
    if (grid < 1) return fail(SRT_ERR_INVALID, "grid must be >= 1");
    if (proto->sch.unroll != kPersistentUnroll)
        return fail(SRT_ERR_UNSUPPORTED, "the plan's schedule is not cut for unroll 4 (TileConfig.unroll = 4)");
    int dev = -1;
    SRT_HIP_TRY(hipGetDevice(&dev));
    if (dev != proto->device) return fail(SRT_ERR_INVALID, "plan belongs to another device");
    ExecPersistentFn fn = proto->exec->persistent[log2_small(V)];
    if (!fn) return fail(SRT_ERR_UNSUPPORTED, "persistent executor variant not generated");
    const int waves = proto->sch.waves_per_block;
    const size_t lds = reduce_lds(waves, proto->pk.M_r, V);
    if (lds > lds_limit()) return fail(SRT_ERR_UNSUPPORTED, "chunk reduction needs more LDS than a workgroup has");
    SRT_HIP_TRY(fn(reinterpret_cast<const ExecProblem *>(d_probs), d_hdr, grid, waves, lds, (hipStream_t)stream));
    return SRT_OK;
}

int srt_group_info(const srt_group *g, int *vec_width, int *workgroups, int *lds_bytes) {
    if (!g) return fail(SRT_ERR_INVALID, "null group");
    if (vec_width) *vec_width = g->V;
    if (workgroups) *workgroups = g->total_blocks;
    if (lds_bytes) *lds_bytes = (int)g->lds;
    return SRT_OK;
}

void srt_group_destroy(srt_group *g) {
    if (!g) return;
    hipFree(g->d_probs);
    delete g;
}

// ------------------------------------------------------------ SOPTile path
int srt_sort_and_encode_patterns(const srt_mapping *m, const int32_t *pat_codes, int n,
                                 int32_t *encoded, int32_t *order, srt_stream_t stream) {
    if (!m || n < 0) return fail(SRT_ERR_INVALID, "bad argument");
    if (n == 0) return SRT_OK;
    if (!pat_codes || !encoded || !order) return fail(SRT_ERR_INVALID, "null array");
    hipStream_t s = (hipStream_t)stream;
    DeviceLut *lut = nullptr;
    SRT_TRY(m->impl->device_lut(&lut));
    Scratch sc;
    int32_t *ids = nullptr, *bad = nullptr, *counts = nullptr;
    int64_t *tot = nullptr;
    SRT_TRY(sc.get(&ids, n));
    SRT_TRY(sc.get(&bad, 1));
    SRT_TRY(sc.get(&counts, lut->num_patterns + 1));
    SRT_TRY(sc.get(&tot, 2));
    SRT_HIP_TRY(hipMemsetAsync(bad, 0, sizeof(int32_t), s));
    SRT_HIP_TRY(launch_encode_entries(n, *lut, pat_codes, nullptr, 0, ids, bad, s));
    int h_bad = 0;
    SRT_HIP_TRY(hipMemcpyAsync(&h_bad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SRT_HIP_TRY(hipStreamSynchronize(s));
    if (h_bad) return fail(SRT_ERR_INVALID, "pattern not in the kernel pattern set");
    SRT_HIP_TRY(launch_count_entries(n, *lut, ids, counts, tot, tot + 1, s));
    SRT_HIP_TRY(launch_stable_order(n, lut->num_patterns, ids, counts, encoded, order, s));
    SRT_HIP_TRY(hipStreamSynchronize(s));
    return SRT_OK;
}

int srt_tile_create(const srt_mapping *m, int64_t rows, int64_t cols, int64_t num_panels,
                    srt_tile **out) {
    if (!m || !out) return fail(SRT_ERR_INVALID, "null argument");
    if (rows < 0 || cols < 0 || num_panels < 0 || num_panels > 0x7fffffff)
        return fail(SRT_ERR_INVALID, "bad tile shape");
    if (num_panels * m->impl->M_r < rows)
        return fail(SRT_ERR_INVALID, "num_panels * M_r must cover the tile rows");
    const ExecutorDesc *exec = find_executor(m->impl->M_r, m->impl->patterns);
    if (!exec) return fail(SRT_ERR_UNSUPPORTED, "no CDNA4 executor generated for this pattern set");
    auto t = std::make_unique<srt_tile>();
    t->map = m->impl;
    t->rows = rows;
    t->cols = cols;
    t->num_panels = num_panels;
    t->panels.resize(num_panels);
    t->exec = exec;
    SRT_HIP_TRY(hipGetDevice(&t->device));
    *out = t.release();
    return SRT_OK;
}

int srt_tile_pack_panel(srt_tile *t, int64_t panel_id, int64_t col_offset, const int32_t *pat_codes,
                        const int32_t *col_indices, int n, const float *dense_panel, int K,
                        srt_stream_t stream) {
    if (!t) return fail(SRT_ERR_INVALID, "null tile");
    if (panel_id < 0 || panel_id >= t->num_panels) return fail(SRT_ERR_INVALID, "panel_id out of range");
    if (t->tile_packed || t->panels[panel_id].packed) return fail(SRT_ERR_INVALID, "panel already packed");
    if (n < 0 || K < 0) return fail(SRT_ERR_INVALID, "negative size");
    if (n > 0 && (!pat_codes || !col_indices || !dense_panel)) return fail(SRT_ERR_INVALID, "null array");
    // Packed B-row indices are col + col_offset with col < K: they must stay
    // inside the tile's columns (B has exactly `cols` rows, template:366).
    if (col_offset < 0 || col_offset + (int64_t)K > t->cols)
        return fail(SRT_ERR_INVALID, "col_offset + K exceeds the tile's columns");
    hipStream_t s = (hipStream_t)stream;
    MappingImpl &map = *t->map;
    DeviceLut *lut = nullptr;
    SRT_TRY(map.device_lut(&lut));
    const int NK = lut->num_patterns;
    PanelBuf pb;
    Scratch sc;
    int32_t *ids = nullptr, *bad = nullptr;
    int64_t *tot = nullptr, *pad = nullptr, *nzc = nullptr;
    SRT_TRY(sc.get(&ids, n));
    SRT_TRY(sc.get(&bad, 1));
    SRT_TRY(sc.get(&tot, 2));
    SRT_TRY(sc.get(&pad, 1));
    SRT_TRY(sc.get(&nzc, 1));
    SRT_TRY(dmalloc(&pb.counts, NK + 1));
    SRT_HIP_TRY(hipMemsetAsync(bad, 0, sizeof(int32_t), s));
    SRT_HIP_TRY(hipMemsetAsync(pad, 0, sizeof(int64_t), s));
    SRT_HIP_TRY(launch_encode_entries(n, *lut, pat_codes, col_indices, K, ids, bad, s));
    SRT_HIP_TRY(launch_count_entries(n, *lut, ids, pb.counts, tot, tot + 1, s));
    SRT_HIP_TRY(launch_count_nonzero(dense_panel, (int64_t)K * map.M_r, nzc, s));
    int h_bad = 0;
    int64_t h_tot[2] = {0, 0};
    int64_t h_nz = 0;
    SRT_HIP_TRY(hipMemcpyAsync(&h_bad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SRT_HIP_TRY(hipMemcpyAsync(h_tot, tot, sizeof(h_tot), hipMemcpyDeviceToHost, s));
    SRT_HIP_TRY(hipMemcpyAsync(&h_nz, nzc, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SRT_HIP_TRY(hipStreamSynchronize(s));
    if (h_bad & 1) {
        hipFree(pb.counts);
        return fail(SRT_ERR_INVALID, "pattern not in the kernel pattern set");
    }
    if (h_bad & 2) {
        hipFree(pb.counts);
        return fail(SRT_ERR_INVALID, "column index out of range of dense_panel");
    }
    pb.ncols = h_tot[0];
    pb.nvals = h_tot[1];
    int rc = dmalloc(&pb.cols, (size_t)pb.ncols);
    if (!rc) rc = dmalloc(&pb.vals, (size_t)pb.nvals);
    if (rc) {
        hipFree(pb.counts);
        hipFree(pb.cols);
        return rc;
    }
    SRT_HIP_TRY(launch_fill_entries(n, (int)col_offset, K, *lut, ids, col_indices, dense_panel,
                                    pb.counts, pb.cols, pb.vals, pad, s));
    SRT_HIP_TRY(hipMemcpyAsync(&pb.padding, pad, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SRT_HIP_TRY(hipStreamSynchronize(s));
    // The reference's packing invariant (template:214).
    if (pb.nvals != pb.padding + h_nz) {
        hipFree(pb.counts);
        hipFree(pb.cols);
        hipFree(pb.vals);
        return fail(SRT_ERR_INVALID,
                    "packed values != padding + nonzeros of dense_panel: the patterns do not "
                    "cover the panel's nonzeros exactly once");
    }
    pb.packed = true;
    t->padding_total += pb.padding;
    t->panels[panel_id] = pb;
    return SRT_OK;
}

int srt_tile_pack_tile(srt_tile *t, srt_stream_t stream) {
    if (!t) return fail(SRT_ERR_INVALID, "null tile");
    if (t->tile_packed) return fail(SRT_ERR_INVALID, "tile already packed");
    for (const PanelBuf &pb : t->panels)
        if (!pb.packed) return fail(SRT_ERR_INVALID, "every panel must be packed before pack_tile");
    hipStream_t s = (hipStream_t)stream;
    Packed &pk = t->pk;
    const int P = (int)t->num_panels;
    const int NK = (int)t->map->patterns.size();
    pk.M = (int)t->rows;
    pk.K = (int)t->cols;
    pk.M_r = t->map->M_r;
    pk.P = P;
    pk.NK = NK;
    std::vector<int64_t> cp(P + 1, 0), vp(P + 1, 0);
    for (int p = 0; p < P; ++p) {
        cp[p + 1] = cp[p] + t->panels[p].ncols;
        vp[p + 1] = vp[p] + t->panels[p].nvals;
    }
    pk.total_cols = cp[P];
    pk.total_vals = vp[P];
    pk.padding = t->padding_total;
    SRT_TRY(dmalloc(&pk.counts, (size_t)P * NK));
    SRT_TRY(dmalloc(&pk.col_ptr, (size_t)P + 1));
    SRT_TRY(dmalloc(&pk.val_ptr, (size_t)P + 1));
    SRT_TRY(alloc_packed_arrays(pk, s));
    SRT_HIP_TRY(hipMemcpyAsync(pk.col_ptr, cp.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, s));
    SRT_HIP_TRY(hipMemcpyAsync(pk.val_ptr, vp.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, s));
    for (int p = 0; p < P; ++p) {
        const PanelBuf &pb = t->panels[p];
        if (NK)
            SRT_HIP_TRY(hipMemcpyAsync(pk.counts + (int64_t)p * NK, pb.counts, sizeof(int32_t) * NK,
                                       hipMemcpyDeviceToDevice, s));
        if (pb.ncols)
            SRT_HIP_TRY(hipMemcpyAsync(pk.cols + cp[p], pb.cols, sizeof(int32_t) * pb.ncols,
                                       hipMemcpyDeviceToDevice, s));
        if (pb.nvals)
            SRT_HIP_TRY(hipMemcpyAsync(pk.vals + vp[p], pb.vals, sizeof(float) * pb.nvals,
                                       hipMemcpyDeviceToDevice, s));
    }
    // The reference's single buffer: required bytes
    // sum(4 nnz + 4 ncols + 4 num_nkern + 3 * 64) (template:222-229), each
    // segment copied in and the pointer moved to the next 64-byte boundary
    // (template:231-258).
    auto align64 = [](size_t x) { return (x + 63) / 64 * 64; };
    size_t req = 0, off = 0;
    t->ref_offsets.assign(3 * (size_t)P, 0);
    for (int p = 0; p < P; ++p) {
        const PanelBuf &pb = t->panels[p];
        req += 4 * (size_t)pb.nvals + 4 * (size_t)pb.ncols + 4 * (size_t)NK + 3 * 64;
        t->ref_offsets[3 * p] = (int64_t)off;
        off = align64(off + 4 * (size_t)pb.nvals);
        t->ref_offsets[3 * p + 1] = (int64_t)off;
        off = align64(off + 4 * (size_t)pb.ncols);
        t->ref_offsets[3 * p + 2] = (int64_t)off;
        off = align64(off + 4 * (size_t)NK);
    }
    t->ref_required = req;
    t->ref_used = off;
    SRT_HIP_TRY(hipMalloc(&t->ref_buffer, std::max<size_t>(req, 64)));
    SRT_HIP_TRY(hipMemsetAsync(t->ref_buffer, 0, std::max<size_t>(req, 64), s));
    char *rb = static_cast<char *>(t->ref_buffer);
    for (int p = 0; p < P; ++p) {
        const PanelBuf &pb = t->panels[p];
        if (pb.nvals)
            SRT_HIP_TRY(hipMemcpyAsync(rb + t->ref_offsets[3 * p], pb.vals, sizeof(float) * pb.nvals,
                                       hipMemcpyDeviceToDevice, s));
        if (pb.ncols)
            SRT_HIP_TRY(hipMemcpyAsync(rb + t->ref_offsets[3 * p + 1], pb.cols, sizeof(int32_t) * pb.ncols,
                                       hipMemcpyDeviceToDevice, s));
        if (NK)
            SRT_HIP_TRY(hipMemcpyAsync(rb + t->ref_offsets[3 * p + 2], pb.counts, sizeof(int32_t) * NK,
                                       hipMemcpyDeviceToDevice, s));
    }
    SRT_HIP_TRY(hipStreamSynchronize(s));
    DeviceLut *lut = nullptr;
    SRT_TRY(t->map->device_lut(&lut));
    SRT_TRY(build_groups(pk, lut->id_pattern, s));
    for (PanelBuf &pb : t->panels) {
        hipFree(pb.counts);
        hipFree(pb.cols);
        hipFree(pb.vals);
        pb.counts = nullptr;
        pb.cols = nullptr;
        pb.vals = nullptr;
    }
    t->tile_packed = true;
    return SRT_OK;
}

int srt_tile_buffer(const srt_tile *t, const void **buffer, size_t *required_bytes, size_t *used_bytes,
                    int64_t *panel_offsets) {
    if (!t) return fail(SRT_ERR_INVALID, "null tile");
    if (!t->tile_packed) return fail(SRT_ERR_INVALID, "pack_tile has not been called");
    if (buffer) *buffer = t->ref_buffer;
    if (required_bytes) *required_bytes = t->ref_required;
    if (used_bytes) *used_bytes = t->ref_used;
    if (panel_offsets) std::copy(t->ref_offsets.begin(), t->ref_offsets.end(), panel_offsets);
    return SRT_OK;
}

int srt_tile_copy_buffer(const srt_tile *t, void *host_out, size_t bytes) {
    if (!t || (!host_out && bytes)) return fail(SRT_ERR_INVALID, "null argument");
    if (!t->tile_packed) return fail(SRT_ERR_INVALID, "pack_tile has not been called");
    if (bytes > t->ref_required) return fail(SRT_ERR_INVALID, "more bytes than the buffer holds");
    if (bytes) SRT_HIP_TRY(hipMemcpy(host_out, t->ref_buffer, bytes, hipMemcpyDeviceToHost));
    return SRT_OK;
}

int srt_tile_padding(const srt_tile *t, int64_t *padding) {
    if (!t || !padding) return fail(SRT_ERR_INVALID, "null argument");
    *padding = t->padding_total;
    return SRT_OK;
}

int srt_tile_execute(const srt_tile *t, int N_c, const float *B, int64_t ldb, float *C, int64_t ldc,
                     int load_c, const srt_tile_config *cfg, srt_stream_t stream) {
    if (!t) return fail(SRT_ERR_INVALID, "null tile");
    if (!t->tile_packed) return fail(SRT_ERR_INVALID, "pack_tile has not been called");
    srt_tile_config c;
    if (cfg) c = *cfg;
    else srt_tile_config_default(&c);
    realise_config(c, t->pk.M_r, t->pk.K, t->pk.K_c ? t->pk.K_c : t->pk.K, t->pk.total_cols, std::max(N_c, 1));
    std::lock_guard<std::mutex> g(t->mu);
    const int tiles = config_tiles(c, std::max(N_c, 1));
    SRT_TRY(switch_schedule(t->pk, c, tiles, t->exec, t->sch, t->sch_cfg, t->retired, (hipStream_t)stream));
    return run_executor(t->pk, t->sch, t->exec, c, B, ldb, C, ldc, N_c, load_c ? 1 : 0,
                        (hipStream_t)stream);
}

int srt_tile_split_counts(const srt_tile *t, int32_t *kp) {
    if (!t || !kp) return fail(SRT_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(t->mu);
    if (t->sch.kp.empty()) std::fill(kp, kp + t->num_panels, 1);
    else std::copy(t->sch.kp.begin(), t->sch.kp.end(), kp);
    return SRT_OK;
}

int srt_tile_copy_packed(const srt_tile *t, int32_t *counts, int64_t *col_ptr, int64_t *val_ptr,
                         int32_t *cols, float *vals, int64_t *packed_cols, int64_t *packed_values) {
    if (!t) return fail(SRT_ERR_INVALID, "null tile");
    if (!t->tile_packed) return fail(SRT_ERR_INVALID, "pack_tile has not been called");
    if (packed_cols) *packed_cols = t->pk.total_cols;
    if (packed_values) *packed_values = t->pk.total_vals;
    return copy_packed(t->pk, counts, col_ptr, val_ptr, cols, vals);
}

void srt_tile_destroy(srt_tile *t) {
    if (!t) return;
    for (PanelBuf &pb : t->panels) {
        hipFree(pb.counts);
        hipFree(pb.cols);
        hipFree(pb.vals);
    }
    t->pk.release();
    t->sch.release();
    for (RetiredSchedule &r : t->retired) r.sch.release();
    hipFree(t->ref_buffer);
    delete t;
}

// -------------------------------------------------------------------- CSR
int srt_csr_spmm(int m, int k, int n, int64_t nnz, const void *rowptr, const void *colind,
                 int index_bytes, const float *vals, const float *B, int64_t ldb, float *C,
                 int64_t ldc, int n_tile, srt_stream_t stream) {
    if (m < 0 || n < 0 || k < 0 || nnz < 0) return fail(SRT_ERR_INVALID, "negative dimension");
    if (index_bytes == 4 && nnz > 0x7fffffffLL) return fail(SRT_ERR_INVALID, "nnz exceeds int32 indices");
    if (index_bytes != 4 && index_bytes != 8) return fail(SRT_ERR_INVALID, "index_bytes must be 4 or 8");
    if (m == 0 || n == 0) return SRT_OK;
    if (!rowptr || !B || !C || (nnz > 0 && (!colind || !vals))) return fail(SRT_ERR_INVALID, "null array");
    if (ldb < n || ldc < n) return fail(SRT_ERR_INVALID, "ldb and ldc must be >= n");
    if (n_tile != 16 && n_tile != 32 && n_tile != 64 && n_tile != 128 && n_tile != 256)
        return fail(SRT_ERR_INVALID, "n_tile must be one of 16, 32, 64, 128, 256");
    int V = n_tile <= 64 ? 1 : n_tile / 64;
    while (V > 1 && (n % V || ldb % V || ldc % V || !aligned(B, 4 * V) || !aligned(C, 4 * V))) V >>= 1;
    SRT_HIP_TRY(launch_csr(m, k, n, nnz, rowptr, colind, index_bytes, vals, B, ldb, C, ldc, V,
                           (hipStream_t)stream));
    return SRT_OK;
}

}  // extern "C"
