#!/usr/bin/env python3
"""Benchmark: plain FP32 sparse-register-tiling SpMM on one MI355X.

Workload: A = 2048 x 512 magnitude-pruned to 90% sparsity (Gaussian weights,
rows scaled by lognormal(0, 0.5)), B = 512 x 512 U[-1, 1], C = A . B in FP32.
One step = one SpMM launch.  Launches rotate over R replicas of (packed A,
B, C) whose total footprint exceeds twice the 256 MB last-level cache, so
every launch starts cold in cache; inputs are resident in HBM before timing.

Prints one JSON line with the driver's fields plus:
  roofline      HBM roofline of the executor kernel: algorithmic bytes per
                launch (nnz*4 + K*N*4 + M*N*4) / mean kernel duration
                (HIP event pairs around each launch on the launch stream)
  cpu_baseline  the C restatement of the reference's nano-kernel executor
                (oracle/srt_oracle.c, OpenMP over panels) on the host cores,
                same matrix and formula; MKL CSR (torch CPU sparse mm) beside it
"""
from __future__ import annotations

import argparse
import importlib
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "SpMM GFLOP/s (2·nnz·N/s) + %HBM roofline; served p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
MALL_BYTES = 256 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--mapping", default="auto", help="M_r:name (e.g. 8:identity) or auto")
    ap.add_argument("--vec-width", type=int, default=0)
    ap.add_argument("--unroll", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--replicas", type=int, default=0, help="0 = enough to exceed 2x the 256 MB cache")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tune-reps", type=int, default=30)
    return ap.parse_args()


def cpu_baselines(A, B, M_r, name, seconds):
    """Time the oracle's C executor (all host threads and 1 thread) and MKL CSR."""
    import numpy as np
    import torch
    from oracle import cbind, sop_oracle as so
    nthreads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    table = so.read_mapping(os.path.join(ROOT, "sparse-register-tiling_amd", "mappings",
                                         "mapping_%s.txt" % {(8, "identity"): "161fb", (8, "orig"): "400fa",
                                                             (8, "alt"): "747f9", (4, "identity"): "61fee",
                                                             (4, "orig"): "da01e"}[(M_r, name)]))[1]
    pk = cbind.pack_dense(A, M_r, table)
    nnz = int(np.count_nonzero(A))
    flop = 2.0 * nnz * B.shape[1]

    def measure(fn, budget):
        fn()
        t0 = time.perf_counter()
        fn()
        one = max(time.perf_counter() - t0, 1e-6)
        per_rep = max(1, math.ceil(0.015 / one))          # each repetition >= 15 ms
        reps = max(7, int(budget / (one * per_rep)))
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            for _ in range(per_rep):
                fn()
            times.append((time.perf_counter() - t0) / per_rep)
        return statistics.median(times), reps * per_rep

    t_all, runs_all = measure(lambda: cbind.exec_packed(pk, B, nthreads=nthreads), seconds * 0.6)
    t_one, runs_one = measure(lambda: cbind.exec_packed(pk, B, nthreads=1), seconds * 0.2)
    csr = torch.from_numpy(A).to_sparse_csr()
    Bt = torch.from_numpy(B)
    torch.set_num_threads(nthreads)
    t_mkl, runs_mkl = measure(lambda: torch.sparse.mm(csr, Bt), seconds * 0.2)
    return {
        "value": flop / t_all / 1e9, "unit": "GFLOP/s", "cores": nthreads, "kind": "port",
        "sample": "%d runs of the full 2048x512 (90%% sparse) x 512x512 SpMM, executor restated in C "
                  "(M_r=%d %s mapping), median of repetitions >= 15 ms" % (runs_all, M_r, name),
        "single_thread_gflops": flop / t_one / 1e9,
        "mkl_csr_gflops": flop / t_mkl / 1e9,
        "mkl_csr": "torch CPU sparse_csr @ dense (MKL=%s), %d threads, %d runs"
                   % (torch.backends.mkl.is_available(), nthreads, runs_mkl),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: multi-GPU runs are not part of this build "
              "(the SpMM shards by row panels without communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    srt = importlib.import_module("sparse-register-tiling_amd")
    from importlib import import_module
    autotune = import_module("sparse-register-tiling_amd.autotune")
    wl = srt.workloads
    case = wl.NORTH_STAR
    A, B = wl.case_inputs(case, seed=0)
    nnz = int(np.count_nonzero(A))
    M, K, N = case.M, case.K, case.N
    csr = torch.from_numpy(A).to_sparse_csr()
    Bd = torch.from_numpy(B).cuda()
    log("workload %s nnz=%d" % (case.name, nnz))

    # ---- configuration: autotune unless fixed on the command line
    if args.mapping == "auto":
        recs = autotune.tune(csr, N, reps=args.tune_reps, B=Bd)
        (M_r, name), cfg = autotune.best_config(recs)
        log("autotune best:", recs[0])
        tune_top = recs[:5]
    else:
        mr, name = args.mapping.split(":")
        M_r = int(mr)
        cfg = srt.TileConfig(vec_width=args.vec_width, unroll=args.unroll, waves_per_block=args.waves)
        tune_top = None
    mapping = srt.resolve_mapping(srt.named_mapping(M_r, name))

    # ---- correctness gate against torch fp32 dense (cheap, every run)
    p0 = srt.MatMulSpecialized.from_csr(csr, N, mapping_id=mapping, tile_config=cfg)
    C0 = torch.empty((M, N), device="cuda")
    p0(C0, Bd)
    ref = torch.from_numpy(A).cuda() @ Bd
    err = (C0 - ref).abs().max().item()
    if not err <= 1e-3:
        raise SystemExit("result check failed: max |C - A@B| = %g" % err)
    realised = p0.get_config()
    footprint = p0.require_storage + 4 * (K * N + M * N)
    R = args.replicas or max(2, math.ceil(2 * MALL_BYTES / footprint))
    log("config M_r=%d %s %s; %d replicas of %.2f MB" % (M_r, name, realised, R, footprint / 1e6))

    # ---- replicas (distinct packed A, B, C per replica)
    plans = [p0] + [srt.MatMulSpecialized.from_csr(csr, N, mapping_id=mapping, tile_config=cfg)
                    for _ in range(R - 1)]
    Bs = [Bd] + [Bd.clone() for _ in range(R - 1)]
    Cs = [C0] + [torch.empty_like(C0) for _ in range(R - 1)]
    torch.cuda.synchronize()

    def step(i):
        j = i % R
        plans[j](Cs[j], Bs[j])

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    # whole-job timing: K back-to-back launches
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for i in range(args.steps):
        step(i)
    e.record()
    torch.cuda.synchronize()
    ms_total = s.elapsed_time(e)
    ms_per_step = ms_total / args.steps

    # per-launch kernel duration: an event pair around every launch
    n_kt = min(args.steps, 1000)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n_kt)]
    torch.cuda.synchronize()
    for i in range(n_kt):
        ev[i][0].record()
        step(i)
        ev[i][1].record()
    torch.cuda.synchronize()
    kt = [a.elapsed_time(b) * 1e-3 for a, b in ev]
    t_kernel = statistics.mean(kt)
    # warm (single replica) kernel time, labelled separately
    for i in range(20):
        plans[0](Cs[0], Bs[0])
    torch.cuda.synchronize()
    ev2 = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(200)]
    for a, b in ev2:
        a.record()
        plans[0](Cs[0], Bs[0])
        b.record()
    torch.cuda.synchronize()
    t_warm = statistics.mean(x.elapsed_time(y) * 1e-3 for x, y in ev2)

    flop = wl.flops(nnz, N)
    bytes_alg = wl.bytes_alg(nnz, M, K, N)
    achieved = bytes_alg / t_kernel / 1e9
    value = flop / (ms_per_step * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain FP32 SpMM, A 2048x512 magnitude-pruned 90% (lognormal row "
                               "scales, seed 0) x B 512x512 U[-1,1]",
                   "M": M, "K": K, "N": N, "sparsity": case.sparsity, "nnz": nnz,
                   "M_r": M_r, "mapping": name, "mapping_id": mapping.mapping.id if mapping.mapping else None,
                   "vec_width": realised.vec_width, "unroll": realised.unroll,
                   "waves_per_block": realised.waves_per_block, "xcd_remap": realised.xcd_remap,
                   "replicas": R, "cache_cold": True},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_alg_per_launch": bytes_alg, "kernel_us": round(t_kernel * 1e6, 3),
                     "kernel_us_warm": round(t_warm * 1e6, 3),
                     "fp32_vector_frac": round(flop / t_kernel / 157.3e12, 4)},
        "kernel_gflops": round(flop / t_kernel / 1e9, 1),
        "max_abs_err_vs_fp32_dense": err,
        "packed": p0.stats(),
    }
    if tune_top:
        out["autotune_top"] = tune_top
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baselines(A, B, M_r, name, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
