"""Autotuner: picks panel height / pattern set, lane width, unroll and
workgroup shape for one sparse matrix and B width by timing the candidates.

Replaces the reference's empirical heuristic (tools/empirical_heuristic_*.json,
built by tools/paper/heuristic.py:26-50 and consumed by
tools/experiment_runner.py:86-113: sparsity bucket x bcols -> up to 6 nano
configs) and the testbed's grid search (cpp_testbed/src/spmm/SpMMFunctor.h:
56-185: set_config, setup, warm-up, median of timed runs, keep the best).
Timings are device times from HIP events on the launch stream.  Results are
cached per (M, K, sparsity bucket, N) in a JSON file whose layout follows the
heuristic file: bucket -> bcols -> ranked configs.
"""
from __future__ import annotations

import json
import os
import statistics
from dataclasses import asdict
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .matmul import MatMulSpecialized, TileConfig
from .mapping import named_mapping

DEFAULT_MAPPINGS = [(8, "orig"), (8, "identity"), (4, "identity"), (4, "orig")]
DEFAULT_SHAPES = [(v, u, w) for v in (1, 2, 4) for u in (4, 8) for w in (1, 2, 4)]
BUCKETS = [(0.0, 0.7), (0.7, 0.8), (0.8, 0.9), (0.9, 1.0)]


def sparsity_bucket(sparsity: float) -> str:
    """The heuristic file's bucket key, e.g. '(0.8,0.9]'."""
    for lo, hi in BUCKETS:
        if lo < sparsity <= hi or (lo == 0.0 and sparsity <= hi):
            return "(%g,%g]" % (lo, hi)
    return "(0.9,1]"


def time_plan(plan: MatMulSpecialized, B: torch.Tensor, C: torch.Tensor, reps: int = 50,
              blocks: int = 3) -> float:
    """Median over ``blocks`` of the mean device time of ``reps`` launches, in seconds."""
    for _ in range(5):
        plan(C, B)
    times = []
    for _ in range(blocks):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            plan(C, B)
        e.record()
        e.synchronize()
        times.append(s.elapsed_time(e) / 1e3 / reps)
    return statistics.median(times)


def tune(csr: torch.Tensor, N: int, mappings: Sequence[Tuple[int, str]] = DEFAULT_MAPPINGS,
         shapes: Sequence[Tuple[int, int, int]] = DEFAULT_SHAPES, reps: int = 50,
         B: Optional[torch.Tensor] = None) -> List[Dict]:
    """Time every (mapping, V, U, waves) candidate; returns records sorted fastest first."""
    M, K = csr.shape
    if B is None:
        B = torch.rand((K, N), device="cuda") * 2 - 1
    C = torch.empty((M, N), device="cuda")
    out = []
    for M_r, name in mappings:
        base = MatMulSpecialized.from_csr(csr, N, mapping_id=named_mapping(M_r, name))
        for V, U, W in shapes:
            if N % V:
                continue
            cfg = TileConfig(vec_width=V, unroll=U, waves_per_block=W, xcd_remap=1)
            plan = MatMulSpecialized.from_csr(csr, N, mapping_id=base.mapping, tile_config=cfg)
            t = time_plan(plan, B, C, reps)
            out.append({"M_r": M_r, "mapping": name, "vec_width": V, "unroll": U, "waves_per_block": W,
                        "seconds": t, "padding": plan.stats()["padding"]})
            plan.close()
        base.close()
    out.sort(key=lambda r: r["seconds"])
    return out


class TuningCache:
    """JSON cache: {"MxK": {bucket: {N: [records...]}}}."""

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
    return (r["M_r"], r["mapping"]), TileConfig(vec_width=r["vec_width"], unroll=r["unroll"],
                                                waves_per_block=r["waves_per_block"], xcd_remap=1)
