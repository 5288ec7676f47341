"""CNN batched form: the 1x1-conv layers of a ResNet-50-style network as
SpMMs, all layers in one grouped launch, captured in a graph.

A 1x1 convolution with M output and K input channels over an activation of
H*W pixels and a batch of `batch` images is C (M x N) = W (M x K) . X (K x N)
with N = H*W*batch (channels-first activations).  The layer list follows the
1x1 convolutions of ResNet-50's bottleneck groups (the reference's DLMC file
list, tools/filelists/rn50_magnitude_70.txt: bottleneck_1 / bottleneck_3 /
projection of each block group) at the four resolutions of its stages
(H*W = 3136, 784, 196, 49; tools/filelists/mobilenet_70.txt:1-11 pairs each
matrix with its H*W the same way).

Two forms:
  * FP32 (`GroupedPlans`): the register-tiled executors, one grouped launch
    over every layer (srt_group_*); each layer's result is bit-identical to
    its own MatMulSpecialized launch.
  * BF16 / FP8 (`GroupedPanels`): the 16-row MFMA panels, one grouped launch
    (srt_mfma_grouped_execute); activations are kept channels-last (N x K,
    the transposed operand the MFMA form reads), so no conversion runs in
    the step.

The launch shape (mapping = pattern set and panel height, lane width,
workgroup size, chunk split; MFMA tile shape) comes from `tune_fp32` /
`tune_panels`, which time whole grouped launches in graphs.
"""
from __future__ import annotations

import ctypes as ct
import warnings
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .autotune import time_graph
from .mapping import named_mapping
from .matmul import MatMulSpecialized, TileConfig
from .native import ptr, resolve_mapping, stream_handle
from .precision import DTYPES, SHAPES, PanelPlan
from .serving import PROBLEM_DTYPE

# (name, H*W, M = output channels, K = input channels)
RN50_1X1: List[Tuple[str, int, int, int]] = [
    ("group1_proj_b1", 3136, 64, 64), ("group1_proj_b3", 3136, 256, 64),
    ("group1_projection", 3136, 256, 64), ("group1_2_b1", 3136, 64, 256),
    ("group2_proj_b1", 784, 128, 256), ("group2_proj_b3", 784, 512, 128),
    ("group2_projection", 784, 512, 256), ("group2_2_b1", 784, 128, 512),
    ("group3_proj_b1", 196, 256, 512), ("group3_proj_b3", 196, 1024, 256),
    ("group3_projection", 196, 1024, 512), ("group3_2_b1", 196, 256, 1024),
    ("group4_proj_b1", 49, 512, 1024), ("group4_proj_b3", 49, 2048, 512),
    ("group4_projection", 49, 2048, 1024), ("group4_2_b1", 49, 512, 2048),
]


@dataclass
class CnnLayer:
    name: str
    hw: int
    M: int
    K: int
    sparsity: float
    A: torch.Tensor  # CPU float32 CSR

    @property
    def nnz(self) -> int:
        return int(self.A.values().numel())


def cnn_layers(seed: int = 0, sparsity: Tuple[float, float] = (0.80, 0.98),
               layers: Sequence[Tuple[str, int, int, int]] = RN50_1X1) -> List[CnnLayer]:
    """Weights of every layer: sparsity s ~ U[lo, hi] per layer, each entry
    kept independently with probability 1 - s (uniform-random mask), kept
    values N(0, 1).  Deterministic in `seed`."""
    rng = np.random.default_rng(seed)
    out = []
    for name, hw, M, K in layers:
        s = float(rng.uniform(*sparsity))
        keep = rng.random((M, K)) >= s
        w = (rng.standard_normal((M, K)) * keep).astype(np.float32)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", UserWarning)  # torch's "sparse CSR is beta" notice
            csr = torch.from_numpy(w).to_sparse_csr()
        out.append(CnnLayer(name, hw, M, K, s, csr))
    return out


def layer_n(layer: CnnLayer, batch: int) -> int:
    return layer.hw * batch


def activations(layers: Sequence[CnnLayer], batch: int, seed: int = 1, device="cuda") -> List[torch.Tensor]:
    """Channels-first fp32 activations X (K x N), uniform[-1, 1]."""
    g = torch.Generator(device=device).manual_seed(seed)
    return [torch.rand((l.K, layer_n(l, batch)), generator=g, device=device) * 2 - 1 for l in layers]


# ---------------------------------------------------------------- FP32 group
class GroupedPlans:
    """Several MatMulSpecialized plans run as one grouped register-tiled
    launch: C_i = A_i . B_i for every i.  The plans must share a mapping
    (executor) and waves_per_block."""

    def __init__(self, plans: Sequence[MatMulSpecialized], B: Sequence[torch.Tensor],
                 C: Sequence[torch.Tensor], n: Optional[Sequence[int]] = None,
                 stream: Optional[torch.cuda.Stream] = None):
        if not plans:
            raise ValueError("no plans")
        self.plans = list(plans)
        hp = (ct.c_void_p * len(plans))(*[p._h.value for p in plans])
        args = self._operands(B, C, n)
        h = ct.c_void_p()
        _lib.call("srt_group_create", hp, len(plans), *args, stream_handle(stream), ct.byref(h))
        self._h = h
        self._keep = (hp, args, list(B), list(C))

    def _operands(self, B, C, n):
        cnt = len(self.plans)
        if len(B) != cnt or len(C) != cnt:
            raise ValueError("one B and one C per plan")
        n = [b.shape[1] for b in B] if n is None else [int(x) for x in n]
        for p, b, c, ni in zip(self.plans, B, C, n):
            for t, name in ((b, "B"), (c, "C")):
                if not (t.is_cuda and t.dtype == torch.float32 and t.dim() == 2 and t.stride(1) == 1):
                    raise ValueError("%s must be a row-major 2-D float32 CUDA tensor" % name)
            if b.shape[0] < p.k or c.shape[0] < p.m or b.shape[1] < ni or c.shape[1] < ni:
                raise ValueError("operand shapes do not match the plan")
        Bp = (ct.c_void_p * cnt)(*[b.data_ptr() for b in B])
        Cp = (ct.c_void_p * cnt)(*[c.data_ptr() for c in C])
        ldb = (ct.c_int64 * cnt)(*[b.stride(0) for b in B])
        ldc = (ct.c_int64 * cnt)(*[c.stride(0) for c in C])
        nn = (ct.c_int * cnt)(*n)
        return Bp, ldb, Cp, ldc, nn

    def rebind(self, B, C, n=None, stream=None) -> None:
        args = self._operands(B, C, n)
        _lib.call("srt_group_rebind", self._h, *args, stream_handle(stream))
        self._keep = (self._keep[0], args, list(B), list(C))

    def __call__(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        _lib.call("srt_group_execute", self._h, stream_handle(stream))

    def info(self) -> Dict[str, int]:
        v, w, l = ct.c_int(), ct.c_int(), ct.c_int()
        _lib.call("srt_group_info", self._h, ct.byref(v), ct.byref(w), ct.byref(l))
        return {"vec_width": v.value, "workgroups": w.value, "lds_bytes": l.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().srt_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- MFMA group
class GroupedPanels:
    """Several PanelPlans (BF16 or FP8) as one grouped MFMA launch.  Bt_i is
    layer i's activation channels-last (N x ldk, the plan's element type),
    C_i its fp32 output (M x N)."""

    def __init__(self, plans: Sequence[PanelPlan], Bt: Sequence[torch.Tensor], C: Sequence[torch.Tensor],
                 col_scales: Optional[Sequence[Optional[torch.Tensor]]] = None, shape: int = 0,
                 grid: Optional[int] = None):
        if not plans:
            raise ValueError("no plans")
        dt = plans[0].dtype
        if any(p.dtype != dt for p in plans):
            raise ValueError("plans must share one dtype")
        if shape not in SHAPES:
            raise ValueError("unknown launch shape %r" % shape)
        self.plans, self.dtype, self.shape = list(plans), dt, shape
        dev = plans[0].device
        cnt = len(plans)
        col_scales = [None] * cnt if col_scales is None else list(col_scales)
        self.table = np.zeros(cnt, dtype=PROBLEM_DTYPE)
        tiles = 0
        for i, (p, b, c, s) in enumerate(zip(plans, Bt, C, col_scales)):
            n = b.shape[0]
            if c.shape[0] < p.m or c.shape[1] < n or c.dtype != torch.float32 or c.stride(1) != 1:
                raise ValueError("C of problem %d does not match its plan" % i)
            _lib.call("srt_mfma_problem_fill", p._h, ptr(b), b.stride(0), ptr(s), ptr(c), c.stride(0), n, tiles,
                      self.table[i:i + 1].ctypes.data_as(ct.c_void_p))
            t = _lib.lib().srt_mfma_problem_tiles(self.table[i:i + 1].ctypes.data_as(ct.c_void_p), shape)
            if t < 0:
                raise RuntimeError(_lib.last_error())
            tiles += t
        self.tiles = tiles
        self.grid = max(1, tiles if grid is None else min(grid, tiles))
        self.d_table = torch.from_numpy(self.table.view(np.uint8).copy()).to(dev)
        self.d_hdr = torch.tensor([cnt, tiles], dtype=torch.int32, device=dev)
        self._keep = (list(Bt), list(C), col_scales)

    def __call__(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        _lib.call("srt_mfma_grouped_execute", DTYPES[self.dtype], self.shape, ptr(self.d_table), ptr(self.d_hdr),
                  self.grid, stream_handle(stream))


# ---------------------------------------------------------------- the network
# Candidates of the FP32 autotuner: mapping (panel height, pattern set) x
# (lane width, waves per workgroup, waves per panel, K tile).
FP32_MAPPINGS = [(4, "orig"), (8, "orig"), (4, "identity"), (8, "identity")]
FP32_SHAPES = [dict(vec_width=v, waves_per_block=w, nsplit=s, K_c=kc, layout=lay)
               for v in (4, 2) for (w, s) in ((8, 1), (8, 4), (8, 8), (16, 16), (4, 4)) for kc in (0, 256)
               for lay in (1, 2)]


# Tile-config fields one layer of a group may set for itself (the grouped
# launch shares the executor, the lane width and the workgroup size).
LAYER_FIELDS = ("nsplit", "layout", "K_c", "chunk_cols")


class CnnFp32:
    """The FP32 network form: one plan per layer (shared mapping and
    waves_per_block) and one grouped launch over all of them.  `layer_shapes`
    (one dict per layer, keys from LAYER_FIELDS) overrides `shape` per layer."""

    def __init__(self, layers: Sequence[CnnLayer], batch: int, X: Sequence[torch.Tensor],
                 mapping: Tuple[int, str] = (4, "orig"), shape: Optional[Dict] = None,
                 out: Optional[Sequence[torch.Tensor]] = None, layer_shapes: Optional[Sequence[Dict]] = None):
        self.layers, self.batch = list(layers), batch
        shape = dict(FP32_SHAPES[0] if shape is None else shape)
        self.mapping, self.shape = mapping, shape
        if layer_shapes is not None:
            if len(layer_shapes) != len(self.layers):
                raise ValueError("one layer shape per layer")
            for ls in layer_shapes:
                bad = set(ls) - set(LAYER_FIELDS)
                if bad:
                    raise ValueError("a layer may set only %s, not %s" % (LAYER_FIELDS, sorted(bad)))
        self.layer_shapes = [dict(ls) for ls in layer_shapes] if layer_shapes is not None else None
        nm = resolve_mapping(named_mapping(*mapping))

        def cfg(i):
            sh = dict(shape)
            if self.layer_shapes is not None:
                sh.update(self.layer_shapes[i])
            return TileConfig(xcd_remap=1, b_staging=2, unroll=4, **sh)
        self.plans = [MatMulSpecialized.from_csr(l.A, layer_n(l, batch), mapping_id=nm, tile_config=cfg(i))
                      for i, l in enumerate(self.layers)]
        self.X = list(X)
        self.Y = list(out) if out is not None else [
            torch.empty((l.M, layer_n(l, batch)), device="cuda") for l in self.layers]
        self.group = GroupedPlans(self.plans, self.X, self.Y)

    def __call__(self, stream=None) -> List[torch.Tensor]:
        self.group(stream)
        return self.Y

    def per_layer(self, stream=None) -> None:
        """The same work as one launch per layer."""
        for p, x, y in zip(self.plans, self.X, self.Y):
            p(y, x, stream=stream)

    def close(self) -> None:
        self.group.close()
        for p in self.plans:
            p.close()


class CnnPanels:
    """The BF16 / FP8 network form: one PanelPlan per layer and one grouped
    MFMA launch.  X are channels-first fp32 activations; they are converted
    once at set-up to the channels-last operand the step reads."""

    def __init__(self, layers: Sequence[CnnLayer], batch: int, X: Sequence[torch.Tensor], dtype: str = "bf16",
                 shape: int = 0, grid: Optional[int] = None):
        self.layers, self.batch, self.dtype = list(layers), batch, dtype
        self.plans = [PanelPlan.from_csr(l.A, dtype) for l in self.layers]
        prepped = [p.prepare_b(x) for p, x in zip(self.plans, X)]
        self.Xt = [b for b, _ in prepped]
        self.scales = [s for _, s in prepped]
        self.Y = [torch.empty((l.M, layer_n(l, batch)), device="cuda") for l in self.layers]
        self.group = GroupedPanels(self.plans, self.Xt, self.Y, self.scales, shape=shape, grid=grid)

    def __call__(self, stream=None) -> List[torch.Tensor]:
        self.group(stream)
        return self.Y

    def close(self) -> None:
        for p in self.plans:
            p.close()


def tune_fp32(layers: Sequence[CnnLayer], batch: int, X: Sequence[torch.Tensor],
              mappings: Sequence[Tuple[int, str]] = FP32_MAPPINGS, shapes: Sequence[Dict] = FP32_SHAPES,
              reps: int = 20) -> List[Dict]:
    """Time the grouped FP32 launch of every (mapping, shape) candidate;
    records sorted fastest first."""
    out = []
    Y = [torch.empty((l.M, layer_n(l, batch)), device="cuda") for l in layers]
    for mp in mappings:
        for sh in shapes:
            net = CnnFp32(layers, batch, X, mapping=mp, shape=sh, out=Y)
            t = time_graph(net.group, reps)
            out.append({"M_r": mp[0], "mapping": mp[1], "seconds": t, **sh, **net.group.info()})
            net.close()
    out.sort(key=lambda r: r["seconds"])
    return out


def layer_candidates(waves_per_block: int) -> List[Dict]:
    """Per-layer settings refine_fp32_layers tries."""
    splits = sorted({1, 4, 8, waves_per_block} & set(range(1, waves_per_block + 1)))
    return [dict(nsplit=s, layout=lay, K_c=kc) for s in splits for lay in (1, 2) for kc in (0, 256)]


def refine_fp32_layers(layers: Sequence[CnnLayer], batch: int, X: Sequence[torch.Tensor],
                       mapping: Tuple[int, str], shape: Dict, reps: int = 10,
                       candidates: Optional[Sequence[Dict]] = None) -> Tuple[List[Dict], float]:
    """Per-layer split / layout / K tile for a tuned group shape: each layer
    takes the candidate its own launch runs fastest with (same mapping, lane
    width and workgroup size as the group); returns (layer_shapes, seconds of
    the grouped launch with them)."""
    cands = list(candidates) if candidates is not None else layer_candidates(shape["waves_per_block"])
    nm = resolve_mapping(named_mapping(*mapping))
    chosen = []
    for l, x in zip(layers, X):
        y = torch.empty((l.M, layer_n(l, batch)), device="cuda")
        best = None
        for c in cands:
            sh = dict(shape)
            sh.update(c)
            plan = MatMulSpecialized.from_csr(l.A, layer_n(l, batch), mapping_id=nm,
                                              tile_config=TileConfig(xcd_remap=1, b_staging=2, unroll=4, **sh))
            t = time_graph(lambda: plan(y, x), reps)
            plan.close()
            if best is None or t < best[0]:
                best = (t, c)
        chosen.append(dict(best[1]))
    net = CnnFp32(layers, batch, X, mapping=mapping, shape=shape, layer_shapes=chosen)
    t = time_graph(net.group, reps)
    net.close()
    return chosen, t


def tune_panels(layers: Sequence[CnnLayer], batch: int, X: Sequence[torch.Tensor], dtype: str = "bf16",
                shapes: Sequence[int] = tuple(SHAPES), reps: int = 20) -> List[Dict]:
    """Time the grouped MFMA launch for every tile shape; fastest first."""
    net = CnnPanels(layers, batch, X, dtype, shape=shapes[0])
    out = []
    for sh in shapes:
        g = GroupedPanels(net.plans, net.Xt, net.Y, net.scales, shape=sh)
        out.append({"shape": sh, "tile": SHAPES[sh], "seconds": time_graph(g, reps), "tiles": g.tiles})
    net.close()
    out.sort(key=lambda r: r["seconds"])
    return out


def algorithmic_bytes(layers: Sequence[CnnLayer], batch: int, b_elt: int = 4) -> int:
    """HBM bytes one pass over the network must move: the B rows some
    nonzero uses (b_elt bytes per element), C (fp32) and A's nonzeros
    (value + 4-byte index)."""
    total = 0
    for l in layers:
        n = layer_n(l, batch)
        used = int(torch.unique(l.A.col_indices()).numel())
        total += used * n * b_elt + l.M * n * 4 + l.nnz * (b_elt + 4)
    return total


def flops(layers: Sequence[CnnLayer], batch: int) -> int:
    """2 * nnz * N over the layers (the sparse work)."""
    return sum(2 * l.nnz * layer_n(l, batch) for l in layers)
