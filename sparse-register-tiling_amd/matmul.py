"""MatMulSpecialized: the scheduler/executor pair of the reference, on the GPU.

Mirrors ``sop::MatMulSpecialized<KernelDesc, DataTransform>`` as its callers
use it (cpp_testbed/src/spmm/nano/SpMM_SOP.h:64,89,117-134,155-164,178):

    mm = MatMulSpecialized(m, k, n, Lx, Lp, Li, tile_config, nThreads,
                           executor_id, mapping_id, datatransform)
    tile_config = mm.get_config()
    mm.allocate_executor(n)
    mm(C, B)                  # C = A . B
    mm.require_storage        # bytes of packed A

Packing (pattern codes, mapping, stable sort by id, value compression)
happens on the device in the constructor.  ``SpMM_SOP`` is the testbed
adapter around it (SpMM_SOP.h:57-193): ``set_config`` / ``setup`` /
``__call__`` / ``log_extra_info``.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import asdict, dataclass, fields
from typing import Dict, Optional, Sequence, Union

import numpy as np
import torch

from . import _lib
from .mapping import Mapping
from .native import NativeMapping, ptr, require_cuda_f32, resolve_mapping, stream_handle

# tools/paper/configs.py:9-25
SCHEDULES = {"KNM": 0, "NKM": 1, "MNK": 2, "NMK": 3, "KMN": 4, "MKN": 5, "KM": 6, "MK": 7,
             "KN": 8, "NK": 9, "MN": 10, "NM": 11, "K": 12, "N": 13, "M": 14}


@dataclass
class TileConfig:
    """sop::TileConfig fields set at SpMM_SOP.h:107-114 (defaults of SOPConfig,
    :33-54) plus the GPU realisation (0 / -1 = let the library choose):
    vec_width (fp32 columns per lane), waves_per_block (panels per
    workgroup), xcd_remap, unroll (packed columns in flight per wave),
    nsplit (waves sharing one panel; 1 keeps the reference's single chain),
    chunk_cols (target packed columns per wave), b_staging (1: B tile
    staged in LDS, 2: B rows read from global memory), layout (1: panels
    laid into workgroups in order; 2: panels binned heaviest-first into
    whole rounds of the chip's CUs, slots given to the longest chunks) and
    matrix_cores (1: one v_mfma_f32_4x4x1_16b outer product per packed
    column; needs the M_r = 4 "full" mapping, whose only kernel pattern is
    0b1111, and K <= 640; same bits as the register tile)."""
    M_c: int = 16
    K_c: int = 0
    N_c: int = 64
    tiling_strategy: int = 0
    beta: float = 1.0
    sparse_a: int = 1
    max_tlb_entries: int = 64
    tlb_page_size: int = 4096
    runtimeSchedule: int = 0
    vec_width: int = 0
    waves_per_block: int = 0
    xcd_remap: int = -1
    unroll: int = 0
    nsplit: int = 0
    chunk_cols: int = 0
    b_staging: int = 0
    layout: int = 0
    matrix_cores: int = 0

    def to_c(self) -> _lib.TileConfigC:
        c = _lib.TileConfigC()
        for f in fields(self):
            name = "runtime_schedule" if f.name == "runtimeSchedule" else f.name
            setattr(c, name, getattr(self, f.name))
        return c

    @classmethod
    def from_c(cls, c: _lib.TileConfigC) -> "TileConfig":
        kw = {}
        for f in fields(cls):
            name = "runtime_schedule" if f.name == "runtimeSchedule" else f.name
            kw[f.name] = getattr(c, name)
        return cls(**kw)


def get_executor_id(mapping_id: Union[str, Mapping], arch: str = "CDNA4", vec_width_bits: int = 2048,
                    nr: int = -1, search_dirs: Optional[Sequence[str]] = None) -> str:
    """Executor id for a mapping's pattern set (the reference's
    get_executor_id(mapping_id, arch, vec_width_bits, nr), count_nkerns.cpp:50).
    One CDNA4 executor serves every lane width, so arch/vec_width_bits/nr only
    have to name this target."""
    if arch not in ("CDNA4", "gfx950"):
        raise ValueError("this build targets CDNA4 (gfx950) only, not %r" % arch)
    return resolve_mapping(mapping_id, search_dirs=search_dirs).executor_id()


def _index_tensor(x, device) -> torch.Tensor:
    t = torch.as_tensor(x, device=device)
    if t.dtype not in (torch.int32, torch.int64):
        t = t.to(torch.int64)
    return t.contiguous()


class MatMulSpecialized:
    """Packed A on the device plus its CDNA4 executor."""

    def __init__(self, m: int, k: int, n: int, Lx, Lp, Li,
                 tile_config: Optional[TileConfig] = None, nThreads: int = 1,
                 executor_id: Optional[str] = None,
                 mapping_id: Union[str, Mapping, NativeMapping] = "400fa",
                 datatransform: bool = True, mapping_search_dirs: Optional[Sequence[str]] = None,
                 stream: Optional[torch.cuda.Stream] = None, dense_threshold: Optional[float] = None):
        """``dense_threshold``: 16-row panels at least this dense run on FP32
        matrix cores (a hybrid plan, srt_plan_create_hybrid); None = every
        panel register-tiled."""
        device = torch.device("cuda", torch.cuda.current_device())
        self.mapping = resolve_mapping(mapping_id, search_dirs=mapping_search_dirs)
        Lp_t = _index_tensor(Lp, device)
        Li_t = _index_tensor(Li, device)
        if Li_t.dtype != Lp_t.dtype:
            Li_t = Li_t.to(Lp_t.dtype)
        Lx_t = torch.as_tensor(Lx, device=device, dtype=torch.float32).contiguous()
        if Lp_t.numel() != m + 1:
            raise ValueError("Lp must have m + 1 = %d entries, got %d" % (m + 1, Lp_t.numel()))
        if Li_t.numel() != Lx_t.numel():
            raise ValueError("Li and Lx must have the same length")
        self.m, self.k, self.n = int(m), int(k), int(n)
        self.datatransform = datatransform
        self.nThreads = nThreads
        cfg = (tile_config or TileConfig()).to_c()
        h = ct.c_void_p()
        eid = executor_id.encode() if executor_id else None
        if dense_threshold is None:
            _lib.call("srt_plan_create", self.m, self.k, self.n, ptr(Lx_t), ptr(Lp_t), ptr(Li_t),
                      Lp_t.element_size(), ct.byref(cfg), int(nThreads), eid, self.mapping.handle,
                      stream_handle(stream), ct.byref(h))
        else:
            if executor_id:
                raise ValueError("a hybrid plan derives its executor from the mapping")
            _lib.call("srt_plan_create_hybrid", self.m, self.k, self.n, ptr(Lx_t), ptr(Lp_t), ptr(Li_t),
                      Lp_t.element_size(), ct.byref(cfg), float(dense_threshold), self.mapping.handle,
                      stream_handle(stream), ct.byref(h))
        self.dense_threshold = dense_threshold
        self._h = h
        self.device = device

    # -- reference surface -------------------------------------------------
    def get_config(self) -> TileConfig:
        c = _lib.TileConfigC()
        _lib.call("srt_plan_get_config", self._h, ct.byref(c))
        return TileConfig.from_c(c)

    def allocate_executor(self, n: int) -> None:
        _lib.call("srt_plan_allocate_executor", self._h, int(n))

    @property
    def require_storage(self) -> int:
        v = ct.c_size_t()
        _lib.call("srt_plan_required_storage", self._h, ct.byref(v))
        return v.value

    def dense_panels(self) -> Dict[str, object]:
        """Dense 16-row panels of a hybrid plan: count, first rows, device bytes."""
        n = ct.c_int()
        b = ct.c_size_t()
        _lib.call("srt_plan_dense_panels", self._h, ct.byref(n), None, ct.byref(b))
        rows = np.zeros(max(1, n.value), dtype=np.int32)
        _lib.call("srt_plan_dense_panels", self._h, None, rows.ctypes.data_as(ct.c_void_p), None)
        return {"count": n.value, "rows": rows[:n.value], "bytes": b.value}

    def release_columns(self) -> None:
        """Free the packed column indices (execution reads the batch records);
        the schedule then cannot be rebuilt for another width."""
        _lib.call("srt_plan_release_columns", self._h)

    def schedule_bytes(self) -> int:
        """Device bytes of the current launch schedule (wave descriptors and
        batch records) beside the packed data."""
        v = ct.c_size_t()
        _lib.call("srt_plan_schedule_bytes", self._h, ct.byref(v))
        return v.value

    def __call__(self, C: torch.Tensor, B: torch.Tensor, n: Optional[int] = None,
                 stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
        """C[:m, :n] = A . B[:k, :n] on ``stream`` (default: the current stream)."""
        require_cuda_f32(C, "C")
        require_cuda_f32(B, "B")
        n = B.shape[1] if n is None else int(n)
        if B.shape[0] < self.k or C.shape[0] < self.m or C.shape[1] < n or B.shape[1] < n:
            raise ValueError("shape mismatch: A is %dx%d, B %s, C %s, n=%d"
                             % (self.m, self.k, tuple(B.shape), tuple(C.shape), n))
        _lib.call("srt_plan_execute", self._h, ptr(C), C.stride(0), ptr(B), B.stride(0), n,
                  stream_handle(stream))
        return C

    # -- inspection ---------------------------------------------------------
    def stats(self) -> Dict[str, int]:
        s = _lib.PlanStatsC()
        _lib.call("srt_plan_stats_get", self._h, ct.byref(s))
        return {f[0]: getattr(s, f[0]) for f in s._fields_}

    def split_counts(self) -> np.ndarray:
        """Waves each panel is split over in the current schedule (the chunk
        order C = chain_0 + chain_1 + ... is reproducible from it)."""
        kp = np.zeros(max(1, self.stats()["num_panels"]), dtype=np.int32)
        _lib.call("srt_plan_split_counts", self._h, kp.ctypes.data_as(ct.c_void_p))
        return kp[:self.stats()["num_panels"]]

    def packed(self) -> Dict[str, np.ndarray]:
        """Packed arrays copied to the host: counts [P, NK] (nkern_counts;
        [P, k_tiles, NK] for a K-tiled plan), col_ptr / val_ptr [P+1], cols,
        vals."""
        st = self.stats()
        P, NK, KT = st["num_panels"], st["num_patterns"], st["k_tiles"]
        counts = np.zeros(max(1, P * KT * NK), dtype=np.int32)
        col_ptr = np.zeros(P + 1, dtype=np.int64)
        val_ptr = np.zeros(P + 1, dtype=np.int64)
        cols = np.zeros(max(1, st["packed_cols"]), dtype=np.int32)
        vals = np.zeros(max(1, st["packed_values"]), dtype=np.float32)
        _lib.call("srt_plan_copy_packed", self._h, counts.ctypes.data_as(ct.c_void_p),
                  col_ptr.ctypes.data_as(ct.c_void_p), val_ptr.ctypes.data_as(ct.c_void_p),
                  cols.ctypes.data_as(ct.c_void_p), vals.ctypes.data_as(ct.c_void_p))
        counts = counts[:P * KT * NK].reshape(P, KT, NK) if KT > 1 else counts[:P * NK].reshape(P, NK)
        return {"counts": counts, "col_ptr": col_ptr, "val_ptr": val_ptr,
                "cols": cols[:st["packed_cols"]], "vals": vals[:st["packed_values"]],
                "patterns": list(self.mapping.patterns)}

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.lib().srt_plan_destroy(h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- constructors --------------------------------------------------------
    @classmethod
    def from_dense(cls, A: torch.Tensor, n: int, **kw) -> "MatMulSpecialized":
        """Pack a dense matrix (structural nonzeros = entries != 0)."""
        A = torch.as_tensor(A, dtype=torch.float32)
        csr = A.to_sparse_csr()
        return cls(A.shape[0], A.shape[1], n, csr.values(), csr.crow_indices(), csr.col_indices(), **kw)

    @classmethod
    def from_csr(cls, csr: torch.Tensor, n: int, **kw) -> "MatMulSpecialized":
        if csr.layout != torch.sparse_csr:
            raise ValueError("expected a torch.sparse_csr tensor")
        return cls(csr.shape[0], csr.shape[1], n, csr.values(), csr.crow_indices(), csr.col_indices(), **kw)


class SpMM_SOP:
    """The testbed adapter (cpp_testbed/src/spmm/nano/SpMM_SOP.h:57-193) over a
    CSR task: ``set_config`` re-packs when a tiling field changes, ``setup``
    packs once, ``__call__`` runs ``C = A . B`` on the task's tensors."""

    TILING_KEYS = ("m_tile", "k_tile", "n_tile", "tiling_strategy", "beta_10x", "sparse_a",
                   "max_tlb_entries", "tlb_page_size")

    def __init__(self, executor_id: Optional[str], mapping_id: Union[str, Mapping], task: dict,
                 schedule: int = 0):
        # task: {"A": torch sparse_csr (cuda or cpu), "B": cuda fp32 [k, n], "C": cuda fp32 [m, n]}
        self.executor_id = executor_id
        self.mapping_id = mapping_id
        self.task = task
        self.config = {"m_tile": 16, "k_tile": 256, "n_tile": 64, "tiling_strategy": 0,
                       "beta_10x": 10, "sparse_a": 1, "max_tlb_entries": 64, "tlb_page_size": 4096}
        self.schedule = schedule
        self.sop_matmul: Optional[MatMulSpecialized] = None

    def _tile_config(self) -> TileConfig:
        c = self.config
        return TileConfig(M_c=c["m_tile"], K_c=c["k_tile"], N_c=c["n_tile"],
                          tiling_strategy=c["tiling_strategy"], beta=c["beta_10x"] / 10.0,
                          sparse_a=c["sparse_a"], max_tlb_entries=c["max_tlb_entries"],
                          tlb_page_size=c["tlb_page_size"], runtimeSchedule=self.schedule,
                          **{k: c[k] for k in ("vec_width", "waves_per_block", "xcd_remap", "unroll",
                                               "nsplit") if k in c})

    def _build(self) -> None:
        A = self.task["A"]
        n = self.task["B"].shape[1]
        self.sop_matmul = MatMulSpecialized.from_csr(A, n, tile_config=self._tile_config(),
                                                     executor_id=self.executor_id,
                                                     mapping_id=self.mapping_id)
        real = self.sop_matmul.get_config()
        self.config.update(n_tile=real.N_c, m_tile=real.M_c, k_tile=real.K_c,
                           tiling_strategy=real.tiling_strategy)
        self.sop_matmul.allocate_executor(n)

    def set_config(self, new_config: Dict[str, int]) -> bool:
        old = dict(self.config)
        self.config.update(new_config)
        if any(old.get(k) != self.config.get(k) for k in set(old) | set(new_config)):
            self._build()
        return True

    def setup(self) -> None:
        if self.sop_matmul is None:
            self._build()

    def get_config_rep(self) -> str:
        return ",".join("%s:%s" % kv for kv in sorted(self.config.items()))

    def log_extra_info(self, row: dict) -> None:
        row["required_storage"] = self.sop_matmul.require_storage

    def __call__(self) -> None:
        if self.sop_matmul is None:
            raise RuntimeError("sop_matmul is null, not setup")
        self.sop_matmul(self.task["C"], self.task["B"])
