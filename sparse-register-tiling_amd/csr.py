"""HIP CSR SpMM baseline.

``executor(csr, B, num_runs, n_tile) -> (ns_per_run, C)`` mirrors the
reference's ``csr.executor`` (sbench/SOP/_C/pytorch_csr_wrapper.cpp:314-361):
the CSR tensor's int32 or int64 indices are used as they are, the kernel is
timed (median over blocks of ``num_runs`` runs) and C = A . B is returned.
``spmm`` is the plain call.
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch

from . import _lib
from .native import ptr, require_cuda_f32, stream_handle
from .sop import MEASURED_ITERATIONS, _time_runs


def _parts(csr: torch.Tensor):
    if csr.layout != torch.sparse_csr:
        raise ValueError("expected a torch.sparse_csr tensor")
    if csr.dim() != 2:
        raise ValueError("expected a 2-D CSR matrix")
    rp = csr.crow_indices().cuda().contiguous()
    ci = csr.col_indices().cuda().contiguous()
    if ci.dtype != rp.dtype:
        ci = ci.to(rp.dtype)
    vals = csr.values().to(device="cuda", dtype=torch.float32).contiguous()
    return rp, ci, vals


def spmm(csr: torch.Tensor, B: torch.Tensor, C: Optional[torch.Tensor] = None, n_tile: int = 64,
         stream: Optional[torch.cuda.Stream] = None, _parts_cache=None) -> torch.Tensor:
    """C = A_csr . B with the HIP CSR kernel."""
    require_cuda_f32(B, "B")
    m, k = csr.shape
    if B.shape[0] != k:
        raise ValueError("B has %d rows, A has %d columns" % (B.shape[0], k))
    n = B.shape[1]
    if C is None:
        C = torch.empty((m, n), dtype=torch.float32, device=B.device)
    require_cuda_f32(C, "C")
    rp, ci, vals = _parts_cache or _parts(csr)
    _lib.call("srt_csr_spmm", m, k, n, ci.numel(), ptr(rp), ptr(ci), rp.element_size(), ptr(vals),
              ptr(B), B.stride(0), ptr(C), C.stride(0), int(n_tile), stream_handle(stream))
    return C


def executor(csr: torch.Tensor, B: torch.Tensor, num_runs: int = 1, n_tile: int = 64,
             measured_iterations: int = MEASURED_ITERATIONS) -> Tuple[float, torch.Tensor]:
    """(median ns per run, C) like the reference's csr.executor."""
    parts = _parts(csr)
    B = B.cuda().contiguous() if not B.is_cuda else B
    C = torch.zeros((csr.shape[0], B.shape[1]), dtype=torch.float32, device=B.device)
    ns = _time_runs(lambda: spmm(csr, B, C, n_tile, _parts_cache=parts), num_runs, measured_iterations) \
        if num_runs > 0 else float("nan")
    C.zero_()
    spmm(csr, B, C, n_tile, _parts_cache=parts)
    torch.cuda.current_stream().synchronize()
    return ns, C
