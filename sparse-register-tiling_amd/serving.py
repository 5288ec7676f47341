"""Serving form: resident pre-packed sparse weights, continuous batching.

No counterpart in the reference (its bindings run one SpMM per call,
SpMM_SOP.h:178, template:307-392).  A ``Server`` keeps packed weights
resident in HBM and serves requests C = W . B (B: K x N BF16, any N) as they
arrive.  Two weight forms (backends):

* ``RegTileWeights`` (the default served form): the register-tiled sparse
  panels of the plain path (``MatMulSpecialized``, 4-row "orig" pattern set)
  over BF16-valued weights (each value rounded to BF16, held as FP32 for the
  scalar unit), so resident bytes follow the nonzeros; the requests' BF16 B
  is widened to FP32 on the device and every SpMM of a batch runs in one
  persistent grouped executor launch.  FP32 accumulation, the same fused
  multiply-add chain as a lone launch: results equal the oracle's chain over
  the BF16 operands bit for bit;
* ``PanelPlan`` weights: 16-row BF16 / FP8 panels on matrix cores (every
  16 x 32 block holding a nonzero stored dense), B transposed on the device,
  one persistent grouped MFMA launch.

Common to both:

* admission by HBM budget: each admitted request holds its device operand
  and its FP32 C in one block of a reserved slab (power-of-two size
  classes); requests wait while the budget is full;
* continuous batching: whenever one of the launch slots is free, every
  admitted request not yet running (up to ``max_batch``) goes into one batch;
  slots run on their own streams, so batches admitted while others are on
  the GPU overlap on the device;
* one batch = one replay of a HIP graph captured per slot: upload of the
  problem tables (pinned host -> device), one grouped launch preparing every
  request's B operand, one persistent grouped SpMM launch.  Problem counts
  live in device memory, so the same graph replays any batch;
* latency = completion (HIP event on the slot's stream, relative to a
  reference event) - arrival; p50 / p99 and throughput are reported.

Every output tile is computed by one workgroup from the same data in the same
order whatever else is in the batch, so a request's result is bit-identical
alone and in any mix.
"""
from __future__ import annotations

import ctypes as ct
import math
import time
from dataclasses import dataclass
from collections import deque
from typing import Deque, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .precision import DTYPES, SHAPES, PanelPlan

PROBLEM_DTYPE = np.dtype([("blocks", "<u8"), ("Bt", "<u8"), ("C", "<u8"), ("rscale", "<u8"), ("cscale", "<u8"),
                          ("ldb", "<i8"), ("ldc", "<i8"), ("M", "<i4"), ("N", "<i4"), ("KB", "<i4"), ("P", "<i4"),
                          ("tile_begin", "<i4"), ("vec_ok", "<i4"), ("slot", "<u8")])
BPREP_DTYPE = np.dtype([("src", "<u8"), ("dst", "<u8"), ("cscale", "<u8"), ("ld_src", "<i8"), ("ldk", "<i8"),
                        ("K", "<i4"), ("N", "<i4"), ("tile_begin", "<i4"), ("src_bf16", "<i4"),
                        ("pad_to", "<i4"), ("reserved", "<i4")])
assert PROBLEM_DTYPE.itemsize == 88 and BPREP_DTYPE.itemsize == 64


class SlabPool:
    """The HBM budget: one device allocation handed out in power-of-two
    size classes (16 KiB .. the whole slab) with O(1) allocate and free:
    fresh space is carved from the top, freed blocks go to their class's
    list.  ``used`` counts class bytes of live blocks; ``alloc`` returns None
    when the request does not fit.  With nothing live the slab is reset, so
    free lists never strand the budget."""

    MIN_CLASS = 14

    def __init__(self, nbytes: int, device):
        self.nbytes = int(nbytes) >> self.MIN_CLASS << self.MIN_CLASS
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device)
        self.base = self.buf.data_ptr()
        self.free_lists: Dict[int, List[int]] = {}
        self.top = 0
        self.used = 0
        self.live = 0

    @classmethod
    def size_class(cls, nbytes: int) -> int:
        return max(cls.MIN_CLASS, (int(nbytes) - 1).bit_length())

    def alloc(self, nbytes: int) -> Optional[int]:
        c = self.size_class(nbytes)
        size = 1 << c
        if self.used + size > self.nbytes:
            return None
        fl = self.free_lists.get(c)
        if fl:
            off = fl.pop()
        elif self.top + size <= self.nbytes:
            off = self.top
            self.top += size
        else:
            return None
        self.used += size
        self.live += 1
        return off

    def free(self, off: int, nbytes: int) -> None:
        c = self.size_class(nbytes)
        self.used -= 1 << c
        self.live -= 1
        self.free_lists.setdefault(c, []).append(off)
        if self.live == 0:
            self.free_lists.clear()
            self.top = 0

    def view(self, off: int, shape, dtype) -> torch.Tensor:
        n = int(np.prod(shape)) * torch.tensor([], dtype=dtype).element_size()
        return self.buf[off:off + n].view(dtype).view(*shape)


@dataclass
class Request:
    rid: int
    weight: int
    B: torch.Tensor              # K x N BF16, row-major, on the device
    arrival: float = 0.0         # seconds, relative to the server's clock
    done: float = float("nan")
    C: Optional[torch.Tensor] = None
    _off: int = -1
    _bytes: int = 0
    _bt_off: int = 0   # slab offsets of the transposed operand and the result
    _c_off: int = 0
    _group: Optional["_Merged"] = None   # set while the request runs merged with others of its weight

    @property
    def N(self) -> int:
        return self.B.shape[1]

    @property
    def latency(self) -> float:
        return self.done - self.arrival


EXEC_PROBLEM_DTYPE = np.dtype([("prog", "<u8"), ("desc", "<u8"), ("vals", "<u8"), ("B", "<u8"), ("C", "<u8"),
                               ("ldc", "<i8"), ("b_bytes", "<u4"), ("ldb", "<u4"), ("M", "<i4"), ("N", "<i4"),
                               ("K", "<i4"), ("n_wblocks", "<i4"), ("reduce", "<i4"), ("load_c", "<i4"),
                               ("block_begin", "<i4"), ("pad", "<i4")])
assert EXEC_PROBLEM_DTYPE.itemsize == 88


def _align(x: int, a: int = 256) -> int:
    return -(-int(x) // a) * a


class MfmaBackend:
    """16-row BF16 / FP8 panels on matrix cores (``PanelPlan`` weights)."""

    kind = "mfma"
    prob_dtype = PROBLEM_DTYPE

    def __init__(self, weights: Sequence[PanelPlan], shape: int, grid: int):
        dt = {w.dtype for w in weights}
        if len(dt) != 1:
            raise ValueError("all weights must share one dtype")
        self.dtype = dt.pop()
        self.code = DTYPES[self.dtype]
        self.weights = list(weights)
        self.shape, self.grid = int(shape), int(grid)
        self.esz = 2 if self.dtype == "bf16" else 1
        self.wM = np.array([w.m for w in weights], dtype=np.int64)
        self.wK = np.array([w.k for w in weights], dtype=np.int64)
        self.wldk = -(-self.wK // 32) * 32
        nt, wp, pw = SHAPES[self.shape]
        self.wrows = -(-self.wM // (16 * wp * pw))  # row tiles per weight

    def weight_records(self, base: int) -> None:
        n = len(self.weights)
        self.rec = np.zeros(n, dtype=PROBLEM_DTYPE)
        for i, w in enumerate(self.weights):
            _lib.call("srt_mfma_problem_fill", w._h, ct.c_void_p(base), int(self.wldk[i]), None,
                      ct.c_void_p(base), 1, 1, 0, self.rec[i:i + 1].ctypes.data_as(ct.c_void_p))

    @staticmethod
    def result_ld(N: int) -> int:
        return int(N)

    def operand_bytes(self, wi: int, N: int) -> int:
        return _align(N * int(self.wldk[wi]) * self.esz)

    def resident_bytes(self) -> int:
        return sum(w.stats()["stored_blocks"] * 512 * self.esz for w in self.weights)

    def fill(self, slot: "_Slot", wi, N, bt, cp, bp) -> None:
        nt, wp, pw = SHAPES[self.shape]
        n = len(wi)
        ldk = self.wldk[wi]
        tiles = self.wrows[wi] * -(-N // nt)
        ptiles = (ldk // 32) * -(-N // 32)
        pr = slot.prob[:n]
        pr[:] = self.rec[wi]
        pr["Bt"] = bt
        pr["C"] = cp
        pr["ldb"] = ldk
        pr["ldc"] = N
        pr["N"] = N
        pr["tile_begin"] = np.cumsum(tiles) - tiles
        pr["vec_ok"] = 1
        pp = slot.prep[:n]
        pp["src"] = bp
        pp["dst"] = bt
        pp["cscale"] = 0
        pp["ld_src"] = N
        pp["ldk"] = ldk
        pp["K"] = self.wK[wi]
        pp["N"] = N
        pp["tile_begin"] = np.cumsum(ptiles) - ptiles
        pp["src_bf16"] = 1
        pp["pad_to"] = 0
        slot.hdr[:] = (n, int(ptiles.sum()), n, int(tiles.sum()), 0, 0, 0, 0)

    def enqueue(self, slot: "_Slot", s: ct.c_void_p) -> None:
        _lib.call("srt_mfma_grouped_b_prep", self.code, ct.c_void_p(slot.d_prep.data_ptr()),
                  ct.c_void_p(slot.d_hdr.data_ptr()), self.grid, s)
        _lib.call("srt_mfma_grouped_execute", self.code, self.shape, ct.c_void_p(slot.d_prob.data_ptr()),
                  ct.c_void_p(slot.d_hdr.data_ptr() + 8), self.grid, s)


class RegTileWeights:
    """Register-tiled sparse weights for serving: each W rounded to BF16
    values (held as FP32), packed with one mapping and one launch shape so
    every weight runs in the same persistent grouped launch.  The launch
    shape has four floats per lane (B and C strides rounded up to 4, so any
    N), panel-order workgroups (a schedule that serves every N) and unroll 4
    (the persistent variant)."""

    def __init__(self, weights, mapping=(4, "orig"), waves_per_block: int = 8, nsplit: int = 8,
                 chunk_cols: int = 128, n_max: int = 512):
        from .mapping import named_mapping
        from .matmul import MatMulSpecialized, TileConfig
        from .native import resolve_mapping
        nm = resolve_mapping(named_mapping(*mapping))
        cfg = TileConfig(vec_width=4, unroll=4, waves_per_block=waves_per_block, nsplit=nsplit,
                         chunk_cols=chunk_cols, layout=1, b_staging=2, xcd_remap=1)
        self.mapping = mapping
        self.plans = []
        self.dense_shapes = []
        self.nnz = 0
        for W in weights:
            W = torch.as_tensor(W, dtype=torch.float32).cpu()
            Wq = W.to(torch.bfloat16).to(torch.float32)    # BF16 values, exact in FP32
            csr = Wq.to_sparse_csr()
            self.nnz += int(csr.values().numel())
            plan = MatMulSpecialized.from_csr(csr, n_max, mapping_id=nm, tile_config=cfg)
            plan.release_columns()   # the batch records carry the B rows; the schedule never changes
            self.plans.append(plan)
            self.dense_shapes.append(tuple(W.shape))

    def __len__(self) -> int:
        return len(self.plans)

    def __getitem__(self, i):
        return self.plans[i]

    def resident_bytes(self) -> int:
        """Device bytes held per weight: packed values, batch records, wave
        descriptors and the small per-panel arrays (counts, offsets, groups)."""
        return sum(p.require_storage + p.schedule_bytes() for p in self.plans)

    def resident_breakdown(self) -> Dict[str, int]:
        st = [p.stats() for p in self.plans]
        return {"values": sum(4 * s["packed_values"] for s in st),
                "schedules": sum(p.schedule_bytes() for p in self.plans),
                "other": sum(p.require_storage - 4 * s["packed_values"] for p, s in zip(self.plans, st))}


class _Merged:
    """Requests of one weight run as one SpMM over their B columns side by
    side: one slab block holds the merged FP32 B (K x ld) and C (M x ld);
    member i owns columns [col[i], col[i] + N_i); the block is freed with the
    last member."""

    def __init__(self, off: int, nbytes: int, b_bytes: int, ld: int, members: List["Request"], cols: List[int]):
        self.off, self.nbytes, self.b_bytes, self.ld = off, nbytes, b_bytes, ld
        self.col = {id(r): c for r, c in zip(members, cols)}
        self.refs = len(members)


class RegTileBackend:
    """Register-tiled sparse panels (``RegTileWeights``): BF16 B widened to
    FP32 on the device, one persistent grouped executor launch per batch."""

    kind = "regtile"
    prob_dtype = EXEC_PROBLEM_DTYPE
    V = 4

    def __init__(self, weights: RegTileWeights, grid: int, merge: bool = True, narrow_cols: int = 64,
                 mid_cols: int = 0):
        self.weights = weights
        self.plans = weights.plans
        self.grid = int(grid)
        self.merge = bool(merge)
        # SpMMs at most this wide run in a second persistent launch at one
        # float per lane (a wave then spans 64 columns, not 256: a narrow
        # request's B gathers and FMAs shrink 4x); 0 = one launch
        self.narrow_cols = int(narrow_cols)
        # SpMMs wider than narrow_cols and at most mid_cols wide run in a third
        # launch at two floats per lane: one pass over the weight's packed
        # columns, half the B bytes a 4-float pass gathers.  0 = no third
        # launch (the default: on the 48-weight mix a third launch at 128
        # saturated at 62.8k requests/s against 64.8k with two)
        self.mid_cols = int(mid_cols)
        self.wM = np.array([p.m for p in self.plans], dtype=np.int64)
        self.wK = np.array([p.k for p in self.plans], dtype=np.int64)

    def weight_records(self, base: int) -> None:
        n = len(self.plans)
        self.rec = np.zeros(n, dtype=EXEC_PROBLEM_DTYPE)
        for i, p in enumerate(self.plans):
            _lib.call("srt_exec_problem_fill", p._h, None, self.V, None, self.V, 0, self.V, 0,
                      self.rec[i:i + 1].ctypes.data_as(ct.c_void_p))
        self.wblocks = self.rec["n_wblocks"].astype(np.int64)

    @staticmethod
    def result_ld(N: int) -> int:
        """Row stride of B and C: N rounded up to the lane width (columns
        past N are zero in B and scratch in C)."""
        return -(-int(N) // RegTileBackend.V) * RegTileBackend.V

    def operand_bytes(self, wi: int, N: int) -> int:
        return _align(self.result_ld(N) * int(self.wK[wi]) * 4)

    def resident_bytes(self) -> int:
        return self.weights.resident_bytes()

    def fill_batch(self, server: "Server", slot: "_Slot", batch: List["Request"]) -> None:
        """Problem tables of one batch.  Requests of the same weight are
        merged (``merge``): one slab block takes their B columns side by side
        and their results, and one SpMM serves them all -- a single pass over
        the weight's packed columns instead of one per request.  Each column's
        result is the same bits as alone.  Members give back their own blocks
        (first, when the budget is full); when the merged block does not fit
        even then, they run alone."""
        if not self.merge:
            return self.fill(slot, *server._batch_arrays(batch))
        groups: Dict[int, List[Request]] = {}
        for r in batch:
            groups.setdefault(r.weight, []).append(r)
        pool = server.pool
        rows = []     # (weight, N, ld, B addr, C addr) per SpMM
        prep = []     # (src, dst, ldk, pad_to, K, N) per request
        for w, members in groups.items():
            K, M = int(self.wK[w]), int(self.wM[w])
            if len(members) > 1:
                lds = [self.result_ld(r.N) for r in members]
                ld = sum(lds)
                b_bytes = _align(K * ld * 4)
                nbytes = b_bytes + M * ld * 4
                off = pool.alloc(nbytes)
                if off is None:
                    # a full budget: the members' own blocks (nothing written
                    # to them yet) make room; if the merged block still does
                    # not fit, the members take blocks of the same classes back
                    for r in members:
                        pool.free(r._off, r._bytes)
                        r._off = -1
                    off = pool.alloc(nbytes)
                    if off is None:
                        for r in members:
                            if not server._place(r):
                                raise RuntimeError("serving budget: a freed request block could not be re-placed")
                if off is not None:
                    cols = list(np.cumsum(lds) - lds)
                    g = _Merged(off, nbytes, b_bytes, ld, members, [int(c) for c in cols])
                    for r, c, l in zip(members, g.col.values(), lds):
                        if r._off >= 0:
                            pool.free(r._off, r._bytes)
                        r._off, r._group = -1, g
                        prep.append((r.B.data_ptr(), pool.base + off + 4 * c, ld, l, K, r.N))
                    rows.append((w, ld, ld, pool.base + off, pool.base + off + b_bytes))
                    continue
            for r in members:
                ld = self.result_ld(r.N)
                prep.append((r.B.data_ptr(), pool.base + r._bt_off, ld, ld, K, r.N))
                rows.append((w, r.N, ld, pool.base + r._bt_off, pool.base + r._c_off))
        n = len(rows)
        wi = np.fromiter((x[0] for x in rows), dtype=np.int64, count=n)
        N = np.fromiter((x[1] for x in rows), dtype=np.int64, count=n)
        ld = np.fromiter((x[2] for x in rows), dtype=np.int64, count=n)
        bt = np.fromiter((x[3] for x in rows), dtype=np.int64, count=n)
        cp = np.fromiter((x[4] for x in rows), dtype=np.int64, count=n)
        spmm = self._spmm_tables(slot, wi, N, ld, bt, cp)
        m = len(prep)
        pK = np.fromiter((x[4] for x in prep), dtype=np.int64, count=m)
        pad = np.fromiter((x[3] for x in prep), dtype=np.int64, count=m)
        wtiles = -(-(pK * pad) // 1024)   # srt_widen_problem_tiles: K x pad_to walked flat
        pp = slot.prep[:m]
        pp["src"] = np.fromiter((x[0] for x in prep), dtype=np.int64, count=m)
        pp["dst"] = np.fromiter((x[1] for x in prep), dtype=np.int64, count=m)
        pp["cscale"] = 0
        pp["ld_src"] = np.fromiter((x[5] for x in prep), dtype=np.int64, count=m)
        pp["ldk"] = np.fromiter((x[2] for x in prep), dtype=np.int64, count=m)
        pp["K"] = pK
        pp["N"] = pp["ld_src"]
        pp["tile_begin"] = np.cumsum(wtiles) - wtiles
        pp["src_bf16"] = 1
        pp["pad_to"] = pad
        slot.hdr[:] = (m, int(wtiles.sum())) + spmm

    def _spmm_tables(self, slot: "_Slot", wi, N, ld, bt, cp):
        """SpMM records of a batch: at most narrow_cols wide into slot.prob2
        (1 float per lane), wider and at most mid_cols into slot.prob3 (2
        floats per lane), the rest into slot.prob (4 floats per lane); returns
        the three headers {problems, workgroups}: wide, narrow, middle."""
        none = np.zeros(len(N), dtype=bool)
        narrow = N <= self.narrow_cols if self.narrow_cols > 0 else none
        mid = ~narrow & (N <= self.mid_cols) if self.mid_cols > 0 else none
        out = ()
        for table, sel, V in ((slot.prob, ~(narrow | mid), self.V), (slot.prob2, narrow, 1), (slot.prob3, mid, 2)):
            if table is None:
                out += (0, 0)
                continue
            w, n_, l_ = wi[sel], N[sel], ld[sel]
            k = len(w)
            blocks = self.wblocks[w] * -(-n_ // (64 * V))
            pr = table[:k]
            pr[:] = self.rec[w]
            pr["B"] = bt[sel]
            pr["C"] = cp[sel]
            pr["ldb"] = l_
            pr["ldc"] = l_
            pr["N"] = n_
            pr["b_bytes"] = np.maximum(self.wK[w] * l_ * 4, 4)
            pr["block_begin"] = np.cumsum(blocks) - blocks
            out += (k, int(blocks.sum()))
        return out

    def fill(self, slot: "_Slot", wi, N, bt, cp, bp) -> None:
        n = len(wi)
        K = self.wK[wi]
        ld = -(-N // self.V) * self.V
        wtiles = -(-(K * ld) // 1024)   # srt_widen_problem_tiles: the K x ld output walked flat
        spmm = self._spmm_tables(slot, wi, N, ld, bt, cp)
        pp = slot.prep[:n]
        pp["src"] = bp
        pp["dst"] = bt
        pp["cscale"] = 0
        pp["ld_src"] = N
        pp["ldk"] = ld
        pp["K"] = K
        pp["N"] = N
        pp["tile_begin"] = np.cumsum(wtiles) - wtiles
        pp["src_bf16"] = 1
        pp["pad_to"] = 0
        slot.hdr[:] = (n, int(wtiles.sum())) + spmm

    def enqueue(self, slot: "_Slot", s: ct.c_void_p) -> None:
        _lib.call("srt_exec_grouped_b_widen", ct.c_void_p(slot.d_prep.data_ptr()),
                  ct.c_void_p(slot.d_hdr.data_ptr()), self.grid, s)
        _lib.call("srt_exec_persistent_execute", self.plans[0]._h, self.V, ct.c_void_p(slot.d_prob.data_ptr()),
                  ct.c_void_p(slot.d_hdr.data_ptr() + 8), self.grid, s)
        if slot.prob2 is not None:
            _lib.call("srt_exec_persistent_execute", self.plans[0]._h, 1, ct.c_void_p(slot.d_prob2.data_ptr()),
                      ct.c_void_p(slot.d_hdr.data_ptr() + 16), self.grid, s)
        if slot.prob3 is not None:
            _lib.call("srt_exec_persistent_execute", self.plans[0]._h, 2, ct.c_void_p(slot.d_prob3.data_ptr()),
                      ct.c_void_p(slot.d_hdr.data_ptr() + 24), self.grid, s)


class _Slot:
    def __init__(self, max_batch: int, device, backend):
        self.stream = torch.cuda.Stream(device=device)
        self.backend = backend
        pdt = backend.prob_dtype
        self.h_prob = torch.zeros(max_batch * pdt.itemsize, dtype=torch.uint8).pin_memory()
        self.h_prep = torch.zeros(max_batch * BPREP_DTYPE.itemsize, dtype=torch.uint8).pin_memory()
        # headers: prep {n, tiles}, spmm {n, tiles}, narrow spmm {n, tiles},
        # middle spmm {n, tiles}
        self.h_hdr = torch.zeros(8, dtype=torch.int32).pin_memory()
        self.d_prob = torch.zeros_like(self.h_prob, device=device)
        self.d_prep = torch.zeros_like(self.h_prep, device=device)
        self.d_hdr = torch.zeros(8, dtype=torch.int32, device=device)
        self.prob = np.frombuffer(self.h_prob.numpy(), dtype=pdt)
        self.h_prob2 = self.d_prob2 = self.prob2 = None
        if getattr(backend, "narrow_cols", 0) > 0:  # a second SpMM table (narrow problems)
            self.h_prob2 = torch.zeros_like(self.h_prob).pin_memory()
            self.d_prob2 = torch.zeros_like(self.h_prob, device=device)
            self.prob2 = np.frombuffer(self.h_prob2.numpy(), dtype=pdt)
        self.h_prob3 = self.d_prob3 = self.prob3 = None
        if getattr(backend, "mid_cols", 0) > 0:  # a third SpMM table (middle widths)
            self.h_prob3 = torch.zeros_like(self.h_prob).pin_memory()
            self.d_prob3 = torch.zeros_like(self.h_prob, device=device)
            self.prob3 = np.frombuffer(self.h_prob3.numpy(), dtype=pdt)
        self.prep = np.frombuffer(self.h_prep.numpy(), dtype=BPREP_DTYPE)
        self.hdr = self.h_hdr.numpy()
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self.batch: List[Request] = []
        self.end = torch.cuda.Event(enable_timing=True)

    def _enqueue(self) -> None:
        s = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
        self.d_prob.copy_(self.h_prob, non_blocking=True)
        if self.d_prob2 is not None:
            self.d_prob2.copy_(self.h_prob2, non_blocking=True)
        if self.d_prob3 is not None:
            self.d_prob3.copy_(self.h_prob3, non_blocking=True)
        self.d_prep.copy_(self.h_prep, non_blocking=True)
        self.d_hdr.copy_(self.h_hdr, non_blocking=True)
        self.backend.enqueue(self, s)

    def capture(self) -> None:
        """Capture upload + grouped prep + grouped SpMM once; replays reuse it."""
        self.hdr[:] = 0
        with torch.cuda.stream(self.stream):
            self._enqueue()  # warm-up outside capture (an empty batch)
        self.stream.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.stream):
            self._enqueue()
        self.graph = g

    def launch(self, use_graph: bool = True) -> None:
        with torch.cuda.stream(self.stream):
            if use_graph and self.graph is not None:
                self.graph.replay()
            else:
                self._enqueue()
            self.end.record(self.stream)


class Server:
    """Resident weights + admission + continuous batching (see module doc).

    ``weights``: a ``RegTileWeights`` (register-tiled sparse panels) or a
    list of ``PanelPlan`` (MFMA panels).  ``grid``: workgroups of the
    persistent launches; ``shape``: MFMA launch shape (MFMA weights only);
    ``narrow_cols`` (register-tiled weights): SpMMs at most this wide run in
    a second persistent launch at one float per lane (0 = none);
    ``mid_cols``: wider SpMMs up to this width run in a third at two floats
    per lane (0 = none)."""

    def __init__(self, weights, budget_bytes: int = 256 << 20, max_batch: int = 64,
                 slots: int = 3, shape: int = 6, grid: int = 1024, use_graphs: bool = True,
                 merge_same_weight: bool = True, narrow_cols: int = 64, mid_cols: int = 0):
        if weights is None or len(weights) == 0:
            raise ValueError("no weights")
        if isinstance(weights, RegTileWeights):
            self.backend = RegTileBackend(weights, grid, merge=merge_same_weight, narrow_cols=narrow_cols,
                                          mid_cols=mid_cols)
            self.dtype = "bf16"
        else:
            self.backend = MfmaBackend(weights, shape, grid)
            self.dtype = self.backend.dtype
        self.weights = weights
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.pool = SlabPool(budget_bytes, self.device)
        self.max_batch = int(max_batch)
        self.shape = int(shape)
        self.grid = int(grid)
        self.use_graphs = use_graphs
        self.slots = [_Slot(self.max_batch, self.device, self.backend) for _ in range(slots)]
        if use_graphs:
            for sl in self.slots:
                sl.capture()
        torch.cuda.synchronize()
        self.backend.weight_records(self.pool.base)
        self.stats: Dict[str, float] = {}

    def resident_bytes(self) -> int:
        return self.backend.resident_bytes()

    # -- admission ------------------------------------------------------------
    def request_bytes(self, req: Request) -> int:
        return (self.backend.operand_bytes(req.weight, req.N)
                + int(self.backend.wM[req.weight]) * self.backend.result_ld(req.N) * 4)

    def admit(self, req: Request) -> bool:
        """Reserve the request's operand and result in the budget (False = wait)."""
        K = int(self.backend.wK[req.weight])
        if req.B.shape[0] != K or req.B.dtype != torch.bfloat16 or not req.B.is_contiguous():
            raise ValueError("request B must be a contiguous K x N BF16 tensor for weight %d" % req.weight)
        return self._place(req)

    def _place(self, req: Request) -> bool:
        nbytes = self.request_bytes(req)
        off = self.pool.alloc(nbytes)
        if off is None:
            return False
        req._off, req._bytes = off, nbytes
        req._bt_off = off
        req._c_off = off + self.backend.operand_bytes(req.weight, req.N)
        return True

    def result_view(self, req: Request) -> torch.Tensor:
        M = int(self.backend.wM[req.weight])
        g = req._group
        if g is not None:   # merged: the request's columns of the block's C
            c = g.col[id(req)]
            return self.pool.view(g.off + g.b_bytes, (M, g.ld), torch.float32)[:, c:c + req.N]
        ld = self.backend.result_ld(req.N)
        return self.pool.view(req._c_off, (M, ld), torch.float32)[:, :req.N]

    def release(self, req: Request, keep_result: bool = False, stream: Optional[torch.cuda.Stream] = None) -> None:
        """Return the request's memory to the budget.  With keep_result the
        result is first copied out on `stream` (and waited for), since the
        freed region may be reused by the next batch on another stream."""
        if keep_result:
            with torch.cuda.stream(stream or torch.cuda.current_stream()):
                req.C = self.result_view(req).clone()
            (stream or torch.cuda.current_stream()).synchronize()
        g = req._group
        if g is not None:
            req._group = None
            g.refs -= 1
            if g.refs == 0:
                self.pool.free(g.off, g.nbytes)
        else:
            self.pool.free(req._off, req._bytes)

    # -- batches ------------------------------------------------------------
    def _batch_arrays(self, batch: List[Request]):
        n = len(batch)
        wi = np.fromiter((r.weight for r in batch), dtype=np.int64, count=n)
        N = np.fromiter((r.N for r in batch), dtype=np.int64, count=n)
        bt = self.pool.base + np.fromiter((r._bt_off for r in batch), dtype=np.int64, count=n)
        cp = self.pool.base + np.fromiter((r._c_off for r in batch), dtype=np.int64, count=n)
        bp = np.fromiter((r.B.data_ptr() for r in batch), dtype=np.int64, count=n)
        return wi, N, bt, cp, bp

    def _fill(self, slot: _Slot, batch: List[Request]) -> None:
        """Problem tables of one batch, vectorised over its requests."""
        if hasattr(self.backend, "fill_batch"):
            self.backend.fill_batch(self, slot, batch)
        else:
            self.backend.fill(slot, *self._batch_arrays(batch))

    def run_batch(self, batch: List[Request], slot_index: int = 0) -> None:
        """Admit, run and release one batch synchronously; results stay in
        req.C.  Raises if the batch does not fit the budget."""
        for i, r in enumerate(batch):
            if not self.admit(r):
                for q in batch[:i]:
                    self.release(q)
                raise MemoryError("batch does not fit the HBM budget")
        slot = self.slots[slot_index]
        self._fill(slot, batch)
        slot.launch(self.use_graphs)
        slot.stream.synchronize()
        for r in batch:
            self.release(r, keep_result=True, stream=slot.stream)

    def serve(self, requests: List[Request], keep_results: bool = False) -> Dict[str, float]:
        """Open-loop serving of requests (arrival times in seconds from start).
        Returns latency percentiles and throughput; fills req.done (and
        req.C with keep_results, which waits for each copy-out)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        ref = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ref.record()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nxt = 0
        waiting: Deque[Request] = deque()
        admitted: List[Request] = []
        busy: Dict[int, List[Request]] = {}
        done_count = 0
        max_inflight = batches = 0
        while done_count < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                waiting.append(reqs[nxt])
                nxt += 1
            # completions
            for si in list(busy):
                sl = self.slots[si]
                if sl.end.query():
                    t_end = ref.elapsed_time(sl.end) * 1e-3
                    for r in busy.pop(si):
                        r.done = t_end
                        self.release(r, keep_results, sl.stream)
                        done_count += 1
            # admission by budget, arrival order
            while waiting and self.admit(waiting[0]):
                admitted.append(waiting.popleft())
            if waiting and not busy and not admitted and self.pool.used == 0:
                raise MemoryError("request %d needs %d bytes, more than the whole budget"
                                  % (waiting[0].rid, self.request_bytes(waiting[0])))
            # dispatch to free slots
            for si, sl in enumerate(self.slots):
                if not admitted:
                    break
                if si in busy:
                    continue
                batch, admitted = admitted[:self.max_batch], admitted[self.max_batch:]
                self._fill(sl, batch)
                sl.launch(self.use_graphs)
                busy[si] = batch
                batches += 1
                max_inflight = max(max_inflight, len(busy))
        lat = np.array([r.latency for r in reqs])
        span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
        self.stats = {"requests": len(reqs), "batches": batches, "max_inflight_batches": max_inflight,
                      "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                      "mean_ms": float(lat.mean() * 1e3), "req_per_s": len(reqs) / span if span > 0 else float("inf")}
        return self.stats


def serve_plain(plans: Sequence, requests: List[Request], Bs32: Sequence[torch.Tensor]) -> Dict[str, float]:
    """The baseline: one plain FP32 launch per request (``MatMulSpecialized``
    per weight), in arrival order on one stream, each as soon as it arrived."""
    reqs = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
    ref = torch.cuda.Event(enable_timing=True)
    outs = [torch.empty((plans[requests[i].weight].m, requests[i].N), device="cuda") for i in range(len(requests))]
    events = [torch.cuda.Event(enable_timing=True) for _ in requests]
    torch.cuda.synchronize()
    ref.record()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in reqs:
        r = requests[i]
        while time.perf_counter() - t0 < r.arrival:
            pass
        plans[r.weight](outs[i], Bs32[i])
        events[i].record()
    torch.cuda.synchronize()
    for i in reqs:
        requests[i].done = ref.elapsed_time(events[i]) * 1e-3
        requests[i].C = outs[i]
    lat = np.array([requests[i].latency for i in reqs])
    span = max(r.done for r in requests) - min(r.arrival for r in requests)
    return {"requests": len(requests), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
            "req_per_s": len(requests) / span if span > 0 else float("inf")}


# ----------------------------------------------------------------- workload
SERVING_SHAPES = [(1024, 1024), (1024, 1024), (4096, 1024), (1024, 4096)]


def serving_weights(layers: int = 12, sparsity: float = 0.9, seed: int = 0):
    """48 weight matrices: per layer 1024x1024 (x2), 4096x1024, 1024x4096,
    magnitude-pruned (log-normal row scales) to `sparsity`."""
    from .workloads import magnitude_pruned
    out = []
    for layer in range(layers):
        for j, (M, K) in enumerate(SERVING_SHAPES):
            out.append(magnitude_pruned(M, K, sparsity, seed=seed + 97 * layer + j))
    return out


def poisson_requests(num: int, qps: float, weight_shapes, seed: int = 1, n_max: int = 512,
                     device="cuda") -> List[Request]:
    """Poisson arrivals at `qps`; each request picks a uniform random weight
    and N = round(exp(U(0, ln n_max))) in [1, n_max]; B ~ U[-1, 1] in BF16."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, num)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = []
    for i in range(num):
        w = int(rng.integers(len(weight_shapes)))
        n = int(min(n_max, max(1, round(math.exp(rng.uniform(0.0, math.log(n_max)))))))
        K = weight_shapes[w][1]
        B = torch.from_numpy(rng.uniform(-1.0, 1.0, (K, n)).astype(np.float32)).to(device).to(torch.bfloat16)
        reqs.append(Request(rid=i, weight=w, B=B, arrival=float(arrivals[i])))
    return reqs
