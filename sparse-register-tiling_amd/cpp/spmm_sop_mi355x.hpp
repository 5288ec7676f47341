// C++ caller of the C ABI with the lifecycle of the reference's testbed
// functor (cpp_testbed/src/spmm/SpMMFunctor.h:19-206, SpMM_SOP.h:57-193):
// construct on a task (host CSR A, host B and C, B shape, thread count),
// set_config(map<string,int>) with the testbed's keys, setup() copies the
// operands to the device and packs A there, operator()() runs one SpMM,
// copy_output() brings C back, get_config_rep() describes the realised
// configuration.  A changed tiling key re-packs, as SpMM_SOP.h:95-135 does;
// the GPU keys of srt_tile_config are accepted the same way.
//
// Errors throw std::runtime_error carrying srt_last_error() (the testbed
// exits instead; a caller that wants that catches and exits).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "srt.h"

namespace srt_testbed {

// Host CSR as the testbed holds it (cpp_testbed/src/utils/Matrix.h: r, c, nz,
// Lp, Li, Lx).
struct CsrHost {
    int r = 0, c = 0, nz = 0;
    std::vector<int> Lp, Li;
    std::vector<float> Lx;
};

// The fields of SpMMTask<float> the functor reads (SpMMTask.h:10-29).
struct Task {
    const CsrHost *A = nullptr;
    const float *B = nullptr;  // bRows x bCols, row-major
    float *C = nullptr;        // A.r x bCols, row-major
    int bRows = 0, bCols = 0;
    int nThreads = 1;
};

inline void check(int rc, const char *what) {
    if (rc != SRT_OK) throw std::runtime_error(std::string(what) + ": " + srt_last_error());
}
inline void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

class SpMMSopMI355X {
  public:
    using Config = std::map<std::string, int>;

    SpMMSopMI355X(const Task &task, const char *mapping_id, const std::vector<std::string> &mapping_dirs)
        : task_(task) {
        std::vector<const char *> dirs;
        for (const auto &d : mapping_dirs) dirs.push_back(d.c_str());
        check(srt_mapping_read(mapping_id, dirs.data(), (int)dirs.size(), &map_), "srt_mapping_read");
        srt_tile_config_default(&cfg_);
    }
    SpMMSopMI355X(const SpMMSopMI355X &) = delete;
    SpMMSopMI355X &operator=(const SpMMSopMI355X &) = delete;
    ~SpMMSopMI355X() {
        srt_plan_destroy(plan_);
        srt_mapping_destroy(map_);
        for (void *p : {(void *)dLp_, (void *)dLi_, (void *)dLx_, (void *)dB_, (void *)dC_}) (void)hipFree(p);
    }

    // Testbed keys (SpMM_SOP.h:95-113): m_tile, k_tile, n_tile,
    // tiling_strategy, beta_10x, sparse_a, max_tlb_entries, tlb_page_size;
    // GPU keys: vec_width, waves_per_block, unroll, nsplit, chunk_cols,
    // layout, b_staging, xcd_remap.  Returns false for an unknown key.
    bool set_config(const Config &config) {
        srt_tile_config c = cfg_;
        for (const auto &kv : config) {
            const std::string &k = kv.first;
            const int v = kv.second;
            if (k == "m_tile") c.M_c = v;
            else if (k == "k_tile") c.K_c = v;
            else if (k == "n_tile") c.N_c = v;
            else if (k == "tiling_strategy") c.tiling_strategy = v;
            else if (k == "beta_10x") c.beta = float(v) / 10.f;
            else if (k == "sparse_a") c.sparse_a = v;
            else if (k == "max_tlb_entries") c.max_tlb_entries = v;
            else if (k == "tlb_page_size") c.tlb_page_size = v;
            else if (k == "vec_width") c.vec_width = v;
            else if (k == "waves_per_block") c.waves_per_block = v;
            else if (k == "unroll") c.unroll = v;
            else if (k == "nsplit") c.nsplit = v;
            else if (k == "chunk_cols") c.chunk_cols = v;
            else if (k == "layout") c.layout = v;
            else if (k == "b_staging") c.b_staging = v;
            else if (k == "xcd_remap") c.xcd_remap = v;
            else return false;
        }
        cfg_ = c;
        if (plan_) repack();
        return true;
    }

    // Copies A, B to the device and packs A (the MatMulSpecialized constructor).
    void setup() {
        const CsrHost &A = *task_.A;
        if (A.Lp.size() != (size_t)A.r + 1 || A.Li.size() != (size_t)A.nz || A.Lx.size() != (size_t)A.nz)
            throw std::runtime_error("setup: CSR arrays do not match r / nz");
        if (task_.bRows != A.c) throw std::runtime_error("setup: bRows must equal A.c");
        check_hip(hipMalloc(&dLp_, sizeof(int) * (A.r + 1)), "hipMalloc");
        check_hip(hipMalloc(&dLi_, sizeof(int) * (size_t)std::max(A.nz, 1)), "hipMalloc");
        check_hip(hipMalloc(&dLx_, sizeof(float) * (size_t)std::max(A.nz, 1)), "hipMalloc");
        check_hip(hipMalloc(&dB_, sizeof(float) * (size_t)task_.bRows * task_.bCols + 1), "hipMalloc");
        check_hip(hipMalloc(&dC_, sizeof(float) * (size_t)A.r * task_.bCols + 1), "hipMalloc");
        check_hip(hipMemcpy(dLp_, A.Lp.data(), sizeof(int) * (A.r + 1), hipMemcpyHostToDevice), "hipMemcpy");
        if (A.nz) {
            check_hip(hipMemcpy(dLi_, A.Li.data(), sizeof(int) * A.nz, hipMemcpyHostToDevice), "hipMemcpy");
            check_hip(hipMemcpy(dLx_, A.Lx.data(), sizeof(float) * A.nz, hipMemcpyHostToDevice), "hipMemcpy");
        }
        check_hip(hipMemcpy(dB_, task_.B, sizeof(float) * (size_t)task_.bRows * task_.bCols, hipMemcpyHostToDevice),
                  "hipMemcpy");
        repack();
    }

    // One SpMM on the device (operator()(C, B), SpMM_SOP.h:178).
    void operator()() {
        if (!plan_) throw std::runtime_error("operator(): setup() has not run");
        const int n = task_.bCols;
        check(srt_plan_execute(plan_, dC_, n, dB_, n, n, nullptr), "srt_plan_execute");
    }

    void copy_output() {
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check_hip(hipMemcpy(task_.C, dC_, sizeof(float) * (size_t)task_.A->r * task_.bCols, hipMemcpyDeviceToHost),
                  "hipMemcpy");
    }

    // The realised configuration (get_config(), SpMM_SOP.h:126-132).
    const srt_tile_config &config() const { return cfg_real_; }

    std::string get_config_rep() const {
        std::ostringstream s;
        s << "m_tile:" << cfg_real_.M_c << ",k_tile:" << cfg_real_.K_c << ",n_tile:" << cfg_real_.N_c
          << ",vec_width:" << cfg_real_.vec_width << ",waves_per_block:" << cfg_real_.waves_per_block
          << ",unroll:" << cfg_real_.unroll << ",nsplit:" << cfg_real_.nsplit << ",layout:" << cfg_real_.layout;
        return s.str();
    }

  private:
    void repack() {
        srt_plan_destroy(plan_);
        plan_ = nullptr;
        const CsrHost &A = *task_.A;
        check(srt_plan_create(A.r, A.c, task_.bCols, dLx_, dLp_, dLi_, sizeof(int), &cfg_, task_.nThreads, nullptr,
                              map_, nullptr, &plan_),
              "srt_plan_create");
        check(srt_plan_allocate_executor(plan_, task_.bCols), "srt_plan_allocate_executor");
        check(srt_plan_get_config(plan_, &cfg_real_), "srt_plan_get_config");
    }

    Task task_;
    srt_mapping *map_ = nullptr;
    srt_plan *plan_ = nullptr;
    srt_tile_config cfg_{}, cfg_real_{};
    int *dLp_ = nullptr, *dLi_ = nullptr;
    float *dLx_ = nullptr, *dB_ = nullptr, *dC_ = nullptr;
};

}  // namespace srt_testbed
