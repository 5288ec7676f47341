"""Precision form: BF16 and FP8-e4m3 SpMM on MFMA over 16-row dense panels.

No counterpart in the reference (its executors are FP32 only,
tools/codegen/gen_vector_specialized.py:266-349).  A is cut into 16-row panels
and 32-column blocks; every block holding a nonzero is stored dense in the
MFMA operand layout (the reference's ILP padding, sbench/ilp_pad/
nano_solver.py:76-116, taken to dense 16-row panels) and empty blocks are
skipped.  Accumulation is FP32.  FP8 uses per-row scales of A and per-column
scales of B, both computed in FP64 on the device:

    C_ij = s_i * t_j * sum_k e4m3(A_ik / s_i) * e4m3(B_kj / t_j)

    plan = PanelPlan.from_csr(A_csr, dtype="bf16")      # or "fp8"
    Bt, t = plan.prepare_b(B)                             # fp32 -> transposed bf16 / e4m3 (+ scales)
    plan(C, Bt, t)                                        # C = A . B  (fp32)
"""
from __future__ import annotations

import ctypes as ct
from typing import Dict, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .native import ptr, stream_handle

DTYPES = {"bf16": _lib.SRT_DT_BF16, "fp8": _lib.SRT_DT_FP8}
# launch shapes: index -> (N tile, waves per workgroup, panels per wave)
SHAPES = {0: (64, 4, 2), 1: (128, 4, 2), 2: (64, 4, 1), 3: (128, 2, 2), 4: (32, 4, 1), 5: (128, 4, 1),
          6: (64, 8, 1), 7: (128, 8, 1), 8: (64, 16, 1), 9: (32, 8, 1)}


def _index_tensor(x, device) -> torch.Tensor:
    t = torch.as_tensor(x, device=device)
    if t.dtype not in (torch.int32, torch.int64):
        t = t.to(torch.int64)
    return t.contiguous()


def default_launch(num_panels: int, k_blocks: int, n: int, target_groups: int = 256) -> Tuple[int, int]:
    """(shape, ksplit) giving about `target_groups` workgroups: 64-column
    tiles of 8 panels (one per wave, 8 waves), the K range split when panels
    x tiles are too few."""
    shape = 6
    nt, wp, pw = SHAPES[shape]
    groups = -(-num_panels // (wp * pw)) * -(-n // nt)
    ks = 1
    while groups * ks * 2 <= target_groups and ks * 2 <= max(1, k_blocks // 4):
        ks *= 2
    return shape, ks


class PanelPlan:
    """Packed 16-row panels of A on the device for BF16 or FP8 MFMA."""

    def __init__(self, m: int, k: int, Lx, Lp, Li, dtype: str = "bf16",
                 stream: Optional[torch.cuda.Stream] = None):
        if dtype not in DTYPES:
            raise ValueError("dtype must be 'bf16' or 'fp8'")
        device = torch.device("cuda", torch.cuda.current_device())
        Lp_t = _index_tensor(Lp, device)
        Li_t = _index_tensor(Li, device).to(Lp_t.dtype)
        Lx_t = torch.as_tensor(Lx, device=device, dtype=torch.float32).contiguous()
        if Lp_t.numel() != m + 1:
            raise ValueError("Lp must have m + 1 = %d entries" % (m + 1))
        if Li_t.numel() != Lx_t.numel():
            raise ValueError("Li and Lx must have the same length")
        self.m, self.k, self.dtype = int(m), int(k), dtype
        h = ct.c_void_p()
        _lib.call("srt_mfma_plan_create", self.m, self.k, ptr(Lx_t), ptr(Lp_t), ptr(Li_t),
                  Lp_t.element_size(), DTYPES[dtype], stream_handle(stream), ct.byref(h))
        self._h = h
        self.device = device
        self._reserved = (0, 1)

    @classmethod
    def from_csr(cls, csr: torch.Tensor, dtype: str = "bf16") -> "PanelPlan":
        csr = csr.to_sparse_csr() if csr.layout != torch.sparse_csr else csr
        return cls(csr.shape[0], csr.shape[1], csr.values(), csr.crow_indices(), csr.col_indices(), dtype)

    @classmethod
    def from_dense(cls, A, dtype: str = "bf16") -> "PanelPlan":
        A = torch.as_tensor(A, dtype=torch.float32).cpu()
        return cls.from_csr(A.to_sparse_csr(), dtype)

    # -- operands -----------------------------------------------------------
    def prepare_b(self, B: torch.Tensor, stream: Optional[torch.cuda.Stream] = None
                  ) -> Tuple[torch.Tensor, Optional[torch.Tensor]]:
        """fp32 device B (K x N) -> (Bt, FP8 column scales or None): Bt is the
        B operand transposed to N x ldk (k contiguous, ldk = K rounded up to
        32, zeros past K) in the plan's dtype -- torch.bfloat16, or raw e4m3
        bytes (uint8)."""
        if not (B.is_cuda and B.dtype == torch.float32 and B.dim() == 2 and B.stride(1) == 1):
            raise ValueError("B must be a row-major 2-D float32 CUDA tensor")
        rows, cols = B.shape
        ldk = -(-rows // 32) * 32
        if self.dtype == "bf16":
            out = torch.empty((cols, ldk), dtype=torch.bfloat16, device=B.device)
            scales = None
        else:
            out = torch.empty((cols, ldk), dtype=torch.uint8, device=B.device)
            scales = torch.empty(cols, dtype=torch.float32, device=B.device)
        _lib.call("srt_mfma_convert_b", DTYPES[self.dtype], rows, cols, ptr(B), B.stride(0), ptr(out),
                  out.stride(0), ptr(scales), stream_handle(stream))
        return out, scales

    def reserve(self, n_max: int, ksplit_max: int) -> None:
        """Workspace for K-split launches (allocates; do it before graph capture)."""
        _lib.call("srt_mfma_plan_reserve", self._h, int(n_max), int(ksplit_max))
        self._reserved = (max(self._reserved[0], n_max), max(self._reserved[1], ksplit_max))

    def __call__(self, C: torch.Tensor, Bq: torch.Tensor, col_scales: Optional[torch.Tensor] = None,
                 n: Optional[int] = None, shape: Optional[int] = None, ksplit: Optional[int] = None,
                 stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
        want = torch.bfloat16 if self.dtype == "bf16" else torch.uint8
        if Bq.dtype != want or not Bq.is_cuda or Bq.stride(1) != 1:
            raise ValueError("B must be the transposed CUDA %s operand from prepare_b" % want)
        if not (C.is_cuda and C.dtype == torch.float32 and C.stride(1) == 1):
            raise ValueError("C must be a row-major float32 CUDA tensor")
        n = Bq.shape[0] if n is None else int(n)
        if Bq.shape[1] < -(-self.k // 32) * 32 or C.shape[0] < self.m or C.shape[1] < n or Bq.shape[0] < n:
            raise ValueError("shape mismatch")
        st = self.stats()
        dshape, dks = default_launch(st["num_panels"], st["k_blocks"], n)
        shape = dshape if shape is None else shape
        ksplit = dks if ksplit is None else ksplit
        if ksplit > 1 and (self._reserved[0] < n or self._reserved[1] < ksplit):
            self.reserve(max(n, self._reserved[0]), max(ksplit, self._reserved[1]))
        _lib.call("srt_mfma_plan_execute", self._h, ptr(C), C.stride(0), ptr(Bq), Bq.stride(0),
                  ptr(col_scales), n, int(shape), int(ksplit), stream_handle(stream))
        return C

    # -- inspection -----------------------------------------------------------
    def stats(self) -> Dict[str, int]:
        s = _lib.MfmaStatsC()
        _lib.call("srt_mfma_plan_stats_get", self._h, ct.byref(s))
        return {f[0]: getattr(s, f[0]) for f in s._fields_}

    def row_scales(self) -> np.ndarray:
        out = np.zeros(max(1, self.m), dtype=np.float32)
        _lib.call("srt_mfma_plan_row_scales", self._h, out.ctypes.data_as(ct.c_void_p))
        return out[:self.m]

    def blocks(self) -> Dict[str, np.ndarray]:
        """Host copies: slot [P, KB] (-1 = empty block), blk_kb, blocks [n, 512]
        (uint16 BF16 bits or uint8 e4m3 codes, MFMA operand order)."""
        st = self.stats()
        P, KB, nb = st["num_panels"], st["k_blocks"], st["stored_blocks"]
        slot = np.zeros(max(1, P * KB), dtype=np.int32)
        kb = np.zeros(max(1, nb), dtype=np.int32)
        blk = np.zeros((max(1, nb), 512), dtype=np.uint16 if self.dtype == "bf16" else np.uint8)
        _lib.call("srt_mfma_plan_copy_blocks", self._h, slot.ctypes.data_as(ct.c_void_p),
                  kb.ctypes.data_as(ct.c_void_p), blk.ctypes.data_as(ct.c_void_p))
        return {"slot": slot[:P * KB].reshape(P, KB), "blk_kb": kb[:nb], "blocks": blk[:nb]}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().srt_mfma_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spmm(A_csr: torch.Tensor, B: torch.Tensor, dtype: str = "bf16") -> torch.Tensor:
    """One-shot C = A . B in the precision form (packs, converts B, runs)."""
    plan = PanelPlan.from_csr(A_csr, dtype)
    Bq, t = plan.prepare_b(B)
    C = torch.empty((plan.m, B.shape[1]), dtype=torch.float32, device=B.device)
    return plan(C, Bq, t)
