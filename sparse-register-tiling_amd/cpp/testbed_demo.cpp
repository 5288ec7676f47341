// A testbed-style driver for SpMMSopMI355X: random CSR A (uniform-random
// mask, N(0,1) values), B uniform[-1, 1], several configurations set through
// the testbed's keys, each result checked against a double-precision host
// product with the reference's tolerance rule (|x - y| <= 64 eps max(|x|,|y|)
// or |x - y| <= 64 eps, cpp_testbed/demo/SPMM_demo.cpp:99-137).
//
// usage: testbed_demo <mapping dir> [mapping id] [m k n density seed]
// prints one line per configuration and PASS (exit 0) or FAIL (exit 1).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <random>

#include "spmm_sop_mi355x.hpp"

using srt_testbed::CsrHost;
using srt_testbed::SpMMSopMI355X;
using srt_testbed::Task;

static CsrHost random_csr(int m, int k, double density, unsigned seed) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::normal_distribution<float> nv(0.f, 1.f);
    CsrHost A;
    A.r = m;
    A.c = k;
    A.Lp.push_back(0);
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < k; ++j)
            if (u(rng) < density) {
                A.Li.push_back(j);
                A.Lx.push_back(nv(rng));
            }
        A.Lp.push_back((int)A.Li.size());
    }
    A.nz = (int)A.Li.size();
    return A;
}

static bool within_tol(float x, float y) {
    const double eps = std::numeric_limits<float>::epsilon();
    const double d = std::fabs((double)x - (double)y);
    return d <= 64 * eps * std::fmax(std::fabs((double)x), std::fabs((double)y)) || d <= 64 * eps;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <mapping dir> [mapping id] [m k n density seed]\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const char *mapping_id = argc > 2 ? argv[2] : "da01e";
    const int m = argc > 3 ? std::atoi(argv[3]) : 2048;
    const int k = argc > 4 ? std::atoi(argv[4]) : 512;
    const int n = argc > 5 ? std::atoi(argv[5]) : 256;
    const double density = argc > 6 ? std::atof(argv[6]) : 0.1;
    const unsigned seed = argc > 7 ? (unsigned)std::atoi(argv[7]) : 1u;

    CsrHost A = random_csr(m, k, density, seed);
    std::vector<float> B((size_t)k * n), C((size_t)m * n, std::nanf("")), ref((size_t)m * n, 0.f);
    std::mt19937_64 rng(seed + 7);
    std::uniform_real_distribution<float> ub(-1.f, 1.f);
    for (auto &b : B) b = ub(rng);
    for (int i = 0; i < m; ++i) {
        std::vector<double> acc(n, 0.0);
        for (int p = A.Lp[i]; p < A.Lp[i + 1]; ++p)
            for (int j = 0; j < n; ++j) acc[j] += (double)A.Lx[p] * B[(size_t)A.Li[p] * n + j];
        for (int j = 0; j < n; ++j) ref[(size_t)i * n + j] = (float)acc[j];
    }

    Task task;
    task.A = &A;
    task.B = B.data();
    task.C = C.data();
    task.bRows = k;
    task.bCols = n;
    task.nThreads = 1;

    const std::vector<SpMMSopMI355X::Config> configs = {
        {},
        {{"k_tile", 256}},
        {{"k_tile", 128}, {"m_tile", 64}, {"n_tile", 64}, {"tiling_strategy", 1}, {"beta_10x", 10}},
        {{"vec_width", 2}, {"waves_per_block", 16}, {"nsplit", 16}, {"layout", 2}},
        {{"vec_width", 4}, {"unroll", 2}, {"waves_per_block", 8}, {"nsplit", 1}, {"layout", 1}},
    };
    bool ok = true;
    try {
        SpMMSopMI355X f(task, mapping_id, {dir});
        f.setup();
        for (const auto &cfg : configs) {
            if (!f.set_config(cfg)) throw std::runtime_error("unknown configuration key");
            std::fill(C.begin(), C.end(), std::nanf(""));
            f();
            f.copy_output();
            size_t bad = 0;
            for (size_t i = 0; i < C.size(); ++i)
                if (!within_tol(C[i], ref[i])) ++bad;
            std::printf("config {%s}: %zu of %zu outside tolerance\n", f.get_config_rep().c_str(), bad, C.size());
            ok = ok && bad == 0;
        }
        if (f.set_config({{"no_such_key", 1}})) {
            std::printf("an unknown key was accepted\n");
            ok = false;
        }
    } catch (const std::exception &e) {
        std::printf("error: %s\n", e.what());
        ok = false;
    }
    std::printf("%s\n", ok ? "PASS" : "FAIL");
    return ok ? 0 : 1;
}
