"""Autotuner: picks the pattern set (panel height M_r and mapping) and the GPU
realisation of the tile config (lane width, unroll, waves per workgroup,
panel split) for one sparse matrix and B width by timing the candidates.

Replaces the reference's empirical heuristic (tools/empirical_heuristic_*.json,
built by tools/paper/heuristic.py:26-50 and consumed by
tools/experiment_runner.py:86-113: sparsity bucket x bcols -> up to 6 nano
configs) and the testbed's grid search (cpp_testbed/src/spmm/SpMMFunctor.h:
56-185: set_config, setup, warm-up, median of timed runs, keep the best).
Each candidate is timed as a HIP graph of back-to-back launches replayed on
one stream, bracketed by HIP events, so host launch cost does not enter.
Results are cached per (M, K, sparsity bucket, N) in a JSON file whose layout
follows the heuristic file: bucket -> bcols -> ranked configs.
"""
from __future__ import annotations

import json
import os
import statistics
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from .matmul import MatMulSpecialized, TileConfig
from .mapping import named_mapping

DEFAULT_MAPPINGS = [(8, "orig"), (8, "identity"), (4, "identity"), (4, "orig")]
# GPU realisations tried per mapping: B path x lane width x unroll x (waves, split).
DEFAULT_SHAPES = ([dict(b_staging=1, vec_width=1, unroll=u, waves_per_block=16, nsplit=s)
                   for u in (4, 8) for s in (2, 4, 16)]
                  + [dict(b_staging=2, vec_width=v, unroll=u, waves_per_block=w, nsplit=s)
                     for v in (2, 4) for u in (4, 8) for (w, s) in ((8, 8), (16, 16), (4, 4))])
# The reduced space the shipped table is tuned over (every pruned-FFN workload):
# lane width x unroll x workgroup shape x K tile, global-memory B path.
TABLE_SHAPES = [dict(b_staging=2, vec_width=v, unroll=u, waves_per_block=w, nsplit=s, K_c=kc, layout=lay)
                for v in (1, 2, 4) for u in (2, 4) for (w, s) in ((16, 16), (16, 8), (8, 8), (4, 4))
                for kc in (0, 256) for lay in (1, 2)] + \
               [dict(b_staging=1, vec_width=1, unroll=u, waves_per_block=w, nsplit=s, K_c=0, layout=lay)
                for u in (2, 4) for (w, s) in ((16, 16), (16, 8), (8, 8)) for lay in (1, 2)]
TABLE_MAPPINGS = [(4, "orig"), (8, "orig"), (4, "identity"), (2, "identity")]
# Matrix-core candidate: the full 4-row packing on v_mfma_f32_4x4x1_16b
# (its B slice, K rows x 64 columns, must fit the LDS: K <= 640).
MATRIX_CORE_CANDIDATE = ((4, "full"), dict(matrix_cores=1))
MATRIX_CORE_MAX_K = 640
BUCKETS = [(0.0, 0.7), (0.7, 0.8), (0.8, 0.9), (0.9, 1.0)]


def sparsity_bucket(sparsity: float) -> str:
    """The heuristic file's bucket key, e.g. '(0.8,0.9]'.  The sparsity is
    rounded to 4 decimals first: a matrix pruned to 70 % keeps
    floor(0.3 * M * K) entries and measures 0.7000008, which belongs with the
    70 % workloads the table was tuned on."""
    sparsity = round(float(sparsity), 4)
    for lo, hi in BUCKETS:
        if lo < sparsity <= hi or (lo == 0.0 and sparsity <= hi):
            return "(%g,%g]" % (lo, hi)
    return "(0.9,1]"


def time_graph(launch, reps: int = 50, blocks: int = 3) -> float:
    """Median over ``blocks`` replays of a graph holding ``reps`` calls of
    ``launch()``; returns device seconds per call."""
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for _ in range(3):
            launch()
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        for _ in range(reps):
            launch()
    g.replay()
    torch.cuda.synchronize()
    times = []
    for _ in range(blocks):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        cur = torch.cuda.current_stream()
        a.record(cur)
        g.replay()
        b.record(cur)
        b.synchronize()
        times.append(a.elapsed_time(b) / 1e3 / reps)
    del g
    return statistics.median(times)


def time_plan(plan: MatMulSpecialized, B: torch.Tensor, C: torch.Tensor, reps: int = 50,
              blocks: int = 3) -> float:
    return time_graph(lambda: plan(C, B), reps, blocks)


def tune(csr: torch.Tensor, N: int, mappings: Sequence[Tuple[int, str]] = DEFAULT_MAPPINGS,
         shapes: Sequence[Dict] = DEFAULT_SHAPES, reps: int = 50,
         B: Optional[torch.Tensor] = None, refine: int = 4) -> List[Dict]:
    """Time every (mapping, shape) candidate; returns records sorted fastest
    first.  A shape is a dict of TileConfig fields.  The `refine` fastest
    are timed again with three times the replays (their order decides the
    choice, so timing noise of the short first pass does not); they lead the
    list ("refined": True), sorted among themselves, then the rest, sorted."""
    M, K = csr.shape
    if B is None:
        B = torch.rand((K, N), device="cuda") * 2 - 1
    C = torch.empty((M, N), device="cuda")
    out = []
    bases = {}
    for M_r, name in mappings:
        base = MatMulSpecialized.from_csr(csr, N, mapping_id=named_mapping(M_r, name))
        bases[(M_r, name)] = base
        for shape in shapes:
            if N % max(1, shape.get("vec_width", 1)):
                continue
            cfg = TileConfig(xcd_remap=1, **shape)
            plan = MatMulSpecialized.from_csr(csr, N, mapping_id=base.mapping, tile_config=cfg)
            real = plan.get_config()
            t = time_plan(plan, B, C, reps)
            rec = {"M_r": M_r, "mapping": name, "seconds": t, "padding": plan.stats()["padding"]}
            rec.update(shape)
            rec["realised"] = {k: getattr(real, k) for k in ("b_staging", "vec_width", "waves_per_block",
                                                             "nsplit", "chunk_cols", "K_c", "layout")}
            out.append(rec)
            plan.close()
    out.sort(key=lambda r: r["seconds"])
    for i, rec in enumerate(out):
        rec["refined"] = i < refine
        if i >= refine:
            continue
        shape = {k: v for k, v in rec.items() if k in TileConfig.__dataclass_fields__}
        plan = MatMulSpecialized.from_csr(csr, N, mapping_id=bases[(rec["M_r"], rec["mapping"])].mapping,
                                          tile_config=TileConfig(xcd_remap=1, **shape))
        rec["seconds"] = time_plan(plan, B, C, 3 * reps)
        plan.close()
    for base in bases.values():
        base.close()
    out[:refine] = sorted(out[:refine], key=lambda r: r["seconds"])
    return out


# ---------------------------------------------------------------- cold timing
MALL_BYTES = 256 << 20


def time_plan_cold(csr: torch.Tensor, N: int, mapping, cfg: TileConfig, reps: int = 60,
                   B: Optional[torch.Tensor] = None, blocks: int = 3, rounds: int = 10) -> float:
    """Seconds per launch with every launch cold in cache: launches rotate
    over R replicas of (packed A, B, C) whose footprint exceeds twice the
    256 MB last-level cache, captured in one graph of ``rounds`` rounds of
    the replicas (the headline protocol); median over ``blocks`` replays."""
    M, K = csr.shape
    if B is None:
        B = torch.rand((K, N), device="cuda") * 2 - 1
    first = MatMulSpecialized.from_csr(csr, N, mapping_id=mapping, tile_config=cfg)
    footprint = first.require_storage + first.schedule_bytes() + 4 * (K * N + M * N)
    R = max(2, min(reps, -(-2 * MALL_BYTES // footprint)))
    plans = [first] + [MatMulSpecialized.from_csr(csr, N, mapping_id=mapping, tile_config=cfg)
                       for _ in range(R - 1)]
    Bs = [B] + [B.clone() for _ in range(R - 1)]
    Cs = [torch.empty((M, N), device="cuda") for _ in range(R)]
    i = [0]

    def launch():
        j = i[0] % R
        i[0] += 1
        plans[j](Cs[j], Bs[j])

    t = time_graph(launch, reps=max(reps, rounds * R), blocks=blocks)
    for p in plans:
        p.close()
    return t


# ----------------------------------------------------------- counter evidence
def _probe_command(npz: str, N: int, rec: Dict, launches: int) -> List[str]:
    code = ("import importlib; importlib.import_module('sparse-register-tiling_amd.autotune')"
            ".probe_main(['--npz', %r, '--n', '%d', '--config', %r, '--launches', '%d'])"
            % (npz, N, json.dumps({k: v for k, v in rec.items() if k != "realised"}), launches))
    import sys
    return [sys.executable, "-c", code]


def probe_main(argv=None) -> None:
    """Counter probe (run under rocprofv3 --pmc by counter_evidence): one
    launch per replica (first touch of their pages), then ``--launches``
    launches rotated over the cold replicas (the ones measured), then one
    elementwise copy of COPY_BYTES whose FETCH_SIZE / WRITE_SIZE readings
    calibrate the counters."""
    import argparse
    import numpy as np
    ap = argparse.ArgumentParser()
    ap.add_argument("--npz", required=True)
    ap.add_argument("--n", type=int, required=True)
    ap.add_argument("--config", required=True)
    ap.add_argument("--launches", type=int, default=40)
    a = ap.parse_args(argv)
    z = np.load(a.npz)
    M, K = int(z["shape"][0]), int(z["shape"][1])
    csr = torch.sparse_csr_tensor(torch.from_numpy(z["rowptr"]), torch.from_numpy(z["col"]),
                                  torch.from_numpy(z["val"]), size=(M, K))
    rec = json.loads(a.config)
    shape = {k: v for k, v in rec.items() if k in TileConfig.__dataclass_fields__}
    mapping = named_mapping(rec["M_r"], rec["mapping"])
    cfg = TileConfig(xcd_remap=1, **shape)
    B = torch.rand((K, a.n), device="cuda") * 2 - 1
    first = MatMulSpecialized.from_csr(csr, a.n, mapping_id=mapping, tile_config=cfg)
    footprint = first.require_storage + first.schedule_bytes() + 4 * (K * a.n + M * a.n)
    R = max(2, -(-2 * MALL_BYTES // footprint))
    plans = [first] + [MatMulSpecialized.from_csr(csr, a.n, mapping_id=mapping, tile_config=cfg)
                       for _ in range(R - 1)]
    Bs = [B] + [B.clone() for _ in range(R - 1)]
    Cs = [torch.empty((M, a.n), device="cuda") for _ in range(R)]
    torch.cuda.synchronize()
    for i in range(R + a.launches):
        plans[i % R](Cs[i % R], Bs[i % R])
    src = torch.ones(COPY_BYTES // 4, device="cuda")
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()


COPY_BYTES = 256 << 20
EXEC_KERNEL = "exec_kernel"
# the directory holding this package (the probe child imports it from there)
_IMPORT_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _counter_values(outdir: str) -> Dict[str, Dict[str, List[float]]]:
    """Every counter reading, per kernel name, in dispatch order, from
    rocprofv3 CSV."""
    import csv
    import glob
    acc: Dict[str, Dict[str, List[Tuple[int, float]]]] = {}
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                acc.setdefault(r["Kernel_Name"], {}).setdefault(r["Counter_Name"], []).append(
                    (int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    return {k: {c: [x for _, x in sorted(v)] for c, v in d.items()} for k, d in acc.items()}


def _counter_means(outdir: str, last: Optional[int] = None) -> Dict[str, Dict[str, float]]:
    """Mean per dispatch of every counter, per kernel name, from rocprofv3
    CSV; with ``last``, over each kernel's last ``last`` dispatches only."""
    return {k: {c: statistics.mean(v[-last:] if last else v) for c, v in d.items()}
            for k, d in _counter_values(outdir).items()}


def counter_evidence(csr: torch.Tensor, N: int, records: List[Dict], top: int = 4, launches: int = 40,
                     rocprof: str = "rocprofv3", timeout: int = 240, log=None,
                     keep_dir: Optional[str] = None) -> List[Dict]:
    """rocprofv3 counter evidence for the ``top`` leading records: each runs
    in child processes under ``rocprofv3 --pmc FETCH_SIZE``, then
    ``--pmc WRITE_SIZE`` (cold rotated launches plus a calibration copy),
    then ``--pmc MfmaUtil``, and gets ``hbm_bytes`` (HBM bytes per launch,
    each counter scaled by what it reads for the copy of known size),
    ``hbm_gbps`` (those bytes over its cold time) and ``mfma_util_pct``
    (matrix-core busy share of the executor's SIMD cycles).  Records the
    counters could not be read for are left without them."""
    import subprocess
    import tempfile
    import numpy as np
    csr = csr.to_sparse_csr() if csr.layout != torch.sparse_csr else csr
    with tempfile.TemporaryDirectory(prefix="srt_counters_") as tmp_own:
        tmp = keep_dir or tmp_own  # keep_dir: the rocprofv3 output stays there
        os.makedirs(tmp, exist_ok=True)
        npz = os.path.join(tmp, "a.npz")
        np.savez(npz, shape=np.array(csr.shape), rowptr=csr.crow_indices().cpu().numpy(),
                 col=csr.col_indices().cpu().numpy(), val=csr.values().cpu().float().numpy())
        for i, rec in enumerate(records[:top]):
            scaled = {}
            # One pass per counter: FETCH_SIZE takes 3 of the 4 TCC counter
            # slots and WRITE_SIZE 2, so they cannot share a pass.
            for counter in ("FETCH_SIZE", "WRITE_SIZE"):
                out = os.path.join(tmp, "p%d_%s" % (i, counter))
                cmd = [rocprof, "--pmc", counter, "-d", out, "-o", "run", "--output-format", "csv",
                       "--"] + _probe_command(npz, N, rec, launches)
                try:
                    subprocess.run(cmd, check=True, capture_output=True, timeout=timeout, cwd=_IMPORT_ROOT)
                except (OSError, subprocess.SubprocessError) as e:
                    if log:
                        log("counter pass %s of record %d failed: %s" % (counter, i, type(e).__name__))
                    break
                ex = [d[counter] for k, d in _counter_means(out, last=launches).items()
                      if EXEC_KERNEL in k and counter in d]
                # the calibration copy: the largest single elementwise dispatch
                # (kernels of the same name also ran on small tensors)
                cp = [x for k, d in _counter_values(out).items() if "elementwise" in k
                      for x in d.get(counter, [])]
                if not ex or not cp or max(cp) <= 0:
                    break
                scaled[counter] = ex[0] * COPY_BYTES / max(cp)
            if len(scaled) != 2:
                continue
            rec["hbm_bytes"] = int(round(scaled["FETCH_SIZE"] + scaled["WRITE_SIZE"]))
            t = rec.get("seconds_cold") or rec["seconds"]
            rec["hbm_gbps"] = round(rec["hbm_bytes"] / t / 1e9, 1)
            # Matrix-core utilisation (rocprofv3's derived MfmaUtil: MFMA-busy
            # SIMD cycles over active cycles x SIMDs, percent), its own pass.
            out = os.path.join(tmp, "p%d_MfmaUtil" % i)
            cmd = [rocprof, "--pmc", "MfmaUtil", "-d", out, "-o", "run", "--output-format", "csv",
                   "--"] + _probe_command(npz, N, rec, launches)
            try:
                subprocess.run(cmd, check=True, capture_output=True, timeout=timeout, cwd=_IMPORT_ROOT)
                util = [d["MfmaUtil"] for k, d in _counter_means(out, last=launches).items()
                        if EXEC_KERNEL in k and "MfmaUtil" in d]
                if util:
                    rec["mfma_util_pct"] = round(util[0], 2)
            except (OSError, subprocess.SubprocessError) as e:
                if log:
                    log("counter pass MfmaUtil of record %d failed: %s" % (i, type(e).__name__))
            if log:
                log("record %d: %d HBM bytes per launch, MFMA util %s %%" % (i, rec["hbm_bytes"],
                                                                          rec.get("mfma_util_pct")))
    return records


def rank_with_evidence(records: List[Dict], slack: float = 0.03) -> List[Dict]:
    """Order records by cold time; among those within ``slack`` of the
    fastest that carry counter evidence, the one with the lowest
    cold time x sqrt(HBM bytes / fewest bytes among them) leads: less traffic
    wins a near-tie only when it saves at least about twice the relative time
    it costs (the same speed with less traffic leaves more of the memory
    system to concurrent work)."""
    import math
    key = lambda r: r.get("seconds_cold", r["seconds"])
    recs = sorted(records, key=key)
    if not recs:
        return recs
    best = key(recs[0])
    near = [r for r in recs if key(r) <= best * (1 + slack) and r.get("hbm_bytes")]
    if near:
        fewest = min(r["hbm_bytes"] for r in near)
        lead = min(near, key=lambda r: (key(r) * math.sqrt(r["hbm_bytes"] / fewest), key(r)))
        recs.remove(lead)
        recs.insert(0, lead)
    return recs


def tune_with_evidence(csr: torch.Tensor, N: int, records: List[Dict], top: int = 4,
                       B: Optional[torch.Tensor] = None, cold_reps: int = 60, counters: bool = True,
                       log=None, keep_dir: Optional[str] = None) -> List[Dict]:
    """Second stage of tuning: the ``top`` records of a warm search re-timed
    on the cold-rotated protocol (``seconds_cold``) and, with ``counters``,
    given rocprofv3 HBM-traffic evidence; returned ranked by
    ``rank_with_evidence``, each marked ``evidence``."""
    head = [dict(r) for r in records[:top]]
    for rec in head:
        shape = {k: v for k, v in rec.items() if k in TileConfig.__dataclass_fields__}
        rec["seconds_cold"] = time_plan_cold(csr, N, named_mapping(rec["M_r"], rec["mapping"]),
                                             TileConfig(xcd_remap=1, **shape), reps=cold_reps, B=B)
        if log:
            log("cold %.2f us (warm %.2f us) %s" % (rec["seconds_cold"] * 1e6, rec["seconds"] * 1e6, shape))
    if counters:
        counter_evidence(csr, N, head, top=top, log=log, keep_dir=keep_dir)
    head = rank_with_evidence(head)
    for rec in head:
        rec["evidence"] = "cold" + ("+counters" if "hbm_bytes" in rec else "")
    return head + [dict(r) for r in records[top:]]


# Measured choices shipped with the package (same layout as the cache).
SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "heuristics", "mi355x.json")


class TuningCache:
    """JSON cache: {"MxK": {bucket: {N: [records...]}}}.  ``path=SHIPPED``
    reads the measured table that ships with the package."""

    def __init__(self, path: Optional[str] = None):
        self.path = path or os.path.join(os.path.expanduser("~"), ".cache", "srt_autotune.json")
        try:
            with open(self.path) as f:
                self.data = json.load(f)
        except (OSError, ValueError):
            self.data = {}

    def get(self, M: int, K: int, sparsity: float, N: int) -> Optional[List[Dict]]:
        return self.data.get("%dx%d" % (M, K), {}).get(sparsity_bucket(sparsity), {}).get(str(N))

    def put(self, M: int, K: int, sparsity: float, N: int, records: List[Dict], keep: int = 6) -> None:
        self.data.setdefault("%dx%d" % (M, K), {}).setdefault(sparsity_bucket(sparsity), {})[str(N)] = records[:keep]
        os.makedirs(os.path.dirname(self.path), exist_ok=True)
        with open(self.path + ".tmp", "w") as f:
            json.dump(self.data, f, indent=1)
        os.replace(self.path + ".tmp", self.path)


def best_config(records: List[Dict]) -> Tuple[Tuple[int, str], TileConfig]:
    r = records[0]
    kw = {k: v for k, v in r.items() if k in TileConfig.__dataclass_fields__}
    return (r["M_r"], r["mapping"]), TileConfig(xcd_remap=1, **kw)


def tune_workloads(cases, path: str, reps: int = 20, seed: int = 0, log=None) -> "TuningCache":
    """Tune every workload (workloads.SpMMCase) over TABLE_MAPPINGS x
    TABLE_SHAPES and store the ranked records in a TuningCache at `path`."""
    from . import workloads as wl
    cache = TuningCache(path)
    for case in cases:
        A, B = wl.case_inputs(case, seed=seed)
        csr = torch.from_numpy(A).to_sparse_csr()
        recs = tune(csr, case.N, mappings=TABLE_MAPPINGS, shapes=TABLE_SHAPES, reps=reps,
                    B=torch.from_numpy(B).cuda())
        cache.put(case.M, case.K, case.sparsity, case.N, recs)
        if log:
            log(case, recs[0])
    return cache


def add_candidate(cache: "TuningCache", case, mapping: Tuple[int, str], shape: Dict, reps: int = 60,
                  seed: int = 0) -> Dict:
    """Time one more candidate for a workload already in the table and file
    it among the refined records (timed with the refined records' 3x
    replays) at its place by time; evidence marks are dropped (the order
    they certified changed: run the evidence stage again)."""
    from . import workloads as wl
    A, B = wl.case_inputs(case, seed=seed)
    csr = torch.from_numpy(A).to_sparse_csr()
    Bd = torch.from_numpy(B).cuda()
    C = torch.empty((case.M, case.N), device="cuda")
    M_r, name = mapping
    cfg = TileConfig(xcd_remap=1, **shape)
    plan = MatMulSpecialized.from_csr(csr, case.N, mapping_id=named_mapping(M_r, name), tile_config=cfg)
    real = plan.get_config()
    rec = {"M_r": M_r, "mapping": name, "seconds": time_plan(plan, Bd, C, 3 * reps),
           "padding": plan.stats()["padding"]}
    rec.update(shape)
    rec["realised"] = {k: getattr(real, k) for k in ("b_staging", "vec_width", "waves_per_block", "nsplit",
                                                     "chunk_cols", "K_c", "layout", "matrix_cores")}
    rec["refined"] = True
    plan.close()
    old = cache.get(case.M, case.K, case.sparsity, case.N) or []
    keys = ("b_staging", "vec_width", "unroll", "waves_per_block", "nsplit", "K_c", "layout", "matrix_cores")
    same = lambda r: r["M_r"] == M_r and r["mapping"] == name and all(r.get(k) == rec.get(k) for k in keys)
    old = [{k: v for k, v in r.items() if k not in ("evidence", "seconds_cold", "hbm_bytes", "hbm_gbps",
                                                        "mfma_busy", "mfma_util_pct")} for r in old if not same(r)]
    refined = sorted([r for r in old if r.get("refined")] + [rec], key=lambda r: r["seconds"])
    rest = [r for r in old if not r.get("refined")]
    cache.put(case.M, case.K, case.sparsity, case.N, refined + rest, keep=len(refined) + len(rest))
    return rec


def tuned_plan(csr: torch.Tensor, N: int, path: str = SHIPPED) -> MatMulSpecialized:
    """A plan with the best measured configuration of the tuning table for
    this shape, sparsity bucket and N (the reference's heuristic-JSON lookup,
    tools/paper/heuristic.py:26-50, experiment_runner.py:86-113); the default
    configuration when the table has no entry."""
    M, K = csr.shape
    sparsity = 1.0 - csr.values().numel() / float(max(1, M * K))
    recs = TuningCache(path).get(M, K, sparsity, N)
    if not recs:
        return MatMulSpecialized.from_csr(csr, N, mapping_id=named_mapping(4, "orig"))
    (M_r, name), cfg = best_config(recs)
    return MatMulSpecialized.from_csr(csr, N, mapping_id=named_mapping(M_r, name), tile_config=cfg)


HYBRID_THRESHOLDS = (0.2, 0.25, 0.3, 0.35, 0.4)


def tuned_hybrid(csr: torch.Tensor, N: int, thresholds: Sequence[float] = HYBRID_THRESHOLDS, reps: int = 30,
                 path: str = SHIPPED) -> Tuple[MatMulSpecialized, Dict[str, object]]:
    """The fastest of the tuned plain plan and its hybrids (16-row panels at
    least `threshold` dense on FP32 matrix cores, the rest register-tiled
    with the plain plan's mapping and configuration), graph-timed on the
    device at width N.  The reference fixes its dense-tile threshold at 0.30
    (sbench/padding_reordering.py:36); here each threshold that yields a
    different set of dense panels is timed and the measurement decides.
    Returns the plan and a report {"plain_s", "choice": threshold or None,
    "candidates": [(threshold, dense panels, seconds)]}; the other plans are
    closed."""
    M, K = csr.shape
    dev = torch.device("cuda", torch.cuda.current_device())
    g = torch.Generator(device="cpu").manual_seed(1)
    B = (torch.rand((K, N), generator=g) * 2 - 1).to(dev)
    C = torch.empty((M, N), device=dev)
    best = tuned_plan(csr, N, path)
    best_s = time_graph(lambda: best(C, B), reps=reps)
    report: Dict[str, object] = {"plain_s": best_s, "choice": None, "candidates": []}
    cfg, mapping = best.get_config(), best.mapping
    seen = set()
    for tau in sorted(thresholds):
        hp = MatMulSpecialized.from_csr(csr, N, mapping_id=mapping, tile_config=cfg, dense_threshold=tau)
        rows = tuple(hp.dense_panels()["rows"])
        if not rows or rows in seen:  # no dense panel (the plain plan) or a set already timed
            hp.close()
            continue
        seen.add(rows)
        t = time_graph(lambda: hp(C, B), reps=reps)
        report["candidates"].append((tau, len(rows), t))
        if t < best_s:
            best.close()
            best, best_s, report["choice"] = hp, t, tau
        else:
            hp.close()
    return best, report


def main(argv=None) -> None:
    """Re-measure the tuning table over the 48 pruned-FFN workloads on the
    current GPU (a few minutes on an MI355X):
    python -c "import importlib; importlib.import_module('sparse-register-tiling_amd.autotune').main()" --out t.json"""
    import argparse
    from . import workloads as wl
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--out", required=True, help="JSON table to write (the shipped one is %s)" % SHIPPED)
    ap.add_argument("--reps", type=int, default=20, help="launches per timing (the shortlist gets 3x)")
    ap.add_argument("--evidence", default="",
                    help="comma-separated workloads, MxK (every sparsity and N) or MxK:sparsity:N "
                         "(e.g. 2048x512:0.9:512), whose records get the second stage: cold-rotated timing "
                         "and rocprofv3 counter evidence (from the table --out holds)")
    ap.add_argument("--no-counters", action="store_true", help="evidence stage without rocprofv3")
    ap.add_argument("--top", type=int, default=4)
    ap.add_argument("--counter-dir", help="keep the rocprofv3 counter output of the evidence stage here")
    ap.add_argument("--add-matrix-cores", action="store_true",
                    help="time the matrix-core candidate for every workload with K <= 640 and file it "
                         "among the refined records of the table --out holds")
    a = ap.parse_args(argv)
    if a.add_matrix_cores:
        cache = TuningCache(a.out)
        for case in wl.gpu_cases():
            if case.K > MATRIX_CORE_MAX_K or not cache.get(case.M, case.K, case.sparsity, case.N):
                continue
            rec = add_candidate(cache, case, *MATRIX_CORE_CANDIDATE)
            best = cache.get(case.M, case.K, case.sparsity, case.N)[0]
            print(case, "matrix cores %.2f us, best %.2f us (%s)" % (rec["seconds"] * 1e6, best["seconds"] * 1e6,
                                                                    best["mapping"]), flush=True)
        return
    if a.evidence:
        wanted = set(a.evidence.split(","))
        cache = TuningCache(a.out)
        for case in wl.gpu_cases():
            shape = "%dx%d" % (case.M, case.K)
            if shape not in wanted and "%s:%g:%d" % (shape, case.sparsity, case.N) not in wanted:
                continue
            recs = cache.get(case.M, case.K, case.sparsity, case.N)
            if not recs:
                continue
            A, B = wl.case_inputs(case, seed=0)
            csr = torch.from_numpy(A).to_sparse_csr()
            recs = tune_with_evidence(csr, case.N, recs, top=a.top, B=torch.from_numpy(B).cuda(),
                                      counters=not a.no_counters, log=lambda m: print(case, m, flush=True),
                                      keep_dir=a.counter_dir)
            cache.put(case.M, case.K, case.sparsity, case.N, recs)
            r = recs[0]
            print(case, "cold %.2f us" % (r["seconds_cold"] * 1e6), "hbm %s B" % r.get("hbm_bytes"), flush=True)
        return
    tune_workloads(wl.gpu_cases(), a.out, reps=a.reps,
                   log=lambda c, r: print(c, "%.2f us" % (r["seconds"] * 1e6), flush=True))
