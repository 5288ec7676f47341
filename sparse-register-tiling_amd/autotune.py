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


def main(argv=None) -> None:
    """Re-measure the tuning table over the 48 pruned-FFN workloads on the
    current GPU (a few minutes on an MI355X):
    python -c "import importlib; importlib.import_module('sparse-register-tiling_amd.autotune').main()" --out t.json"""
    import argparse
    from . import workloads as wl
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--out", required=True, help="JSON table to write (the shipped one is %s)" % SHIPPED)
    ap.add_argument("--reps", type=int, default=20, help="launches per timing (the shortlist gets 3x)")
    a = ap.parse_args(argv)
    tune_workloads(wl.gpu_cases(), a.out, reps=a.reps,
                   log=lambda c, r: print(c, "%.2f us" % (r["seconds"] * 1e6), flush=True))
