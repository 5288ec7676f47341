"""SOPTile torch binding, on the GPU.

Mirrors the reference's Python path (sbench/SOP/sop_driver.py:38-129 and the
torch extension sbench/SOP/_C/pytorch_wrapper_intrin.cpp.template:396-409):

    module = make_sop_module(Acc(M_r, N), patterns, pattern_mapping)
    tile   = module.make_sop_tile(dense_tile)          # SOPTile, packed
    ns, C  = module.execute_tile(tile, B, num_runs)    # C = tile . B
    sort_and_encode_patterns(pat_codes) -> (encoded, order)

Packing runs on the device (pack_panel per panel, pack_tile), execution is
the generated CDNA4 executor for the pattern set.  Errors raise
``SrtError`` (the reference raises on TORCH_CHECK and exits on an
unsupported N tail; here tails are masked).
"""
from __future__ import annotations

import ctypes as ct
import math
import statistics
from dataclasses import dataclass
from typing import Callable, Dict, Iterable, List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import _lib
from .mapping import Mapping, mapping_table
from .matmul import TileConfig
from .native import NativeMapping, ptr, require_cuda_f32, resolve_mapping, stream_handle

MEASURED_ITERATIONS = 150  # template:317


@dataclass
class Acc:
    """Register tile: M rows per panel (M_r) and N accumulator width
    (sbench/SOP/sop_utils.py:5-8).  On the GPU a wavefront covers 64*V
    columns; N is kept as a hint for V when it is 1, 2 or 4."""
    M: int = 4
    N: int = 4


def pattern_code(vec) -> int:
    """Bit idx set iff vec[idx] != 0 (sbench/SOP/sop_utils.py:25-30 for 0/1 masks)."""
    v = torch.as_tensor(vec).reshape(-1).cpu()
    code = 0
    for idx, x in enumerate(v.tolist()):
        if x != 0:
            code |= 1 << idx
    return code


def panel_pattern_codes(tile: torch.Tensor, M_r: int) -> torch.Tensor:
    """Codes of every (panel, column) of a dense tile, int64 [P, K], on the tile's device."""
    M, K = tile.shape
    P = (M + M_r - 1) // M_r
    nz = torch.zeros((P * M_r, K), dtype=torch.int64, device=tile.device)
    nz[:M] = (tile != 0).to(torch.int64)
    w = (torch.ones(M_r, dtype=torch.int64, device=tile.device) << torch.arange(M_r, device=tile.device))
    return (nz.view(P, M_r, K) * w.view(1, M_r, 1)).sum(dim=1)


def sort_and_encode_patterns(pat_codes: torch.Tensor, mapping: Union[NativeMapping, Mapping, str],
                             stream: Optional[torch.cuda.Stream] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """(encoded ids sorted ascending with the zero pattern last, stable order)
    (template:75-98).  ``pat_codes`` are kernel pattern bitmasks."""
    nm = resolve_mapping(mapping)
    pc = torch.as_tensor(pat_codes, dtype=torch.int32, device="cuda").contiguous()
    enc = torch.empty_like(pc)
    order = torch.empty_like(pc)
    _lib.call("srt_sort_and_encode_patterns", nm.handle, ptr(pc), pc.numel(), ptr(enc), ptr(order),
              stream_handle(stream))
    return enc, order


class SOPTile:
    """SOPTile(rows, cols, num_panels) (template:101-262), packed on the device."""

    def __init__(self, rows: int, cols: int, num_panels: int, mapping: Union[NativeMapping, Mapping, str]):
        self.mapping = resolve_mapping(mapping)
        self.rows, self.cols, self.num_panels = int(rows), int(cols), int(num_panels)
        h = ct.c_void_p()
        _lib.call("srt_tile_create", self.mapping.handle, self.rows, self.cols, self.num_panels, ct.byref(h))
        self._h = h
        self.tile_packed = False

    def pack_panel(self, panel_id: int, col_offset: int, pat_codes, col_indices,
                   dense_panel: torch.Tensor, stream: Optional[torch.cuda.Stream] = None) -> None:
        """Pack one panel from kernel patterns, their columns and the K x M_r panel."""
        pc = torch.as_tensor(pat_codes, dtype=torch.int32, device="cuda").contiguous()
        ci = torch.as_tensor(col_indices, dtype=torch.int32, device="cuda").contiguous()
        dp = torch.as_tensor(dense_panel, dtype=torch.float32, device="cuda").contiguous()
        if dp.dim() != 2 or dp.shape[1] != self.mapping.M_r:
            raise ValueError("dense_panel must be K x M_r = K x %d" % self.mapping.M_r)
        if pc.numel() != ci.numel():
            raise ValueError("pat_codes and col_indices differ in length")
        _lib.call("srt_tile_pack_panel", self._h, int(panel_id), int(col_offset), ptr(pc), ptr(ci),
                  pc.numel(), ptr(dp), dp.shape[0], stream_handle(stream))

    def pack_tile(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        _lib.call("srt_tile_pack_tile", self._h, stream_handle(stream))
        self.tile_packed = True

    def padding(self) -> int:
        v = ct.c_int64()
        _lib.call("srt_tile_padding", self._h, ct.byref(v))
        return v.value

    def buffer(self) -> Dict[str, object]:
        """The packed tile in the reference's buffer layout (template:219-261):
        ``required_bytes`` (template:222-229), ``used_bytes``, per panel the
        byte offsets of its values / col_indices / nkern_counts segments
        (``offsets`` [P, 3]) and a host copy of the buffer (``data``, uint8)."""
        dev = ct.c_void_p()
        req, used = ct.c_size_t(), ct.c_size_t()
        offs = np.zeros(max(1, 3 * self.num_panels), dtype=np.int64)
        _lib.call("srt_tile_buffer", self._h, ct.byref(dev), ct.byref(req), ct.byref(used),
                  offs.ctypes.data_as(ct.POINTER(ct.c_int64)))
        data = np.zeros(max(1, req.value), dtype=np.uint8)
        _lib.call("srt_tile_copy_buffer", self._h, data.ctypes.data_as(ct.c_void_p), req.value)
        return {"required_bytes": req.value, "used_bytes": used.value,
                "offsets": offs[:3 * self.num_panels].reshape(self.num_panels, 3), "data": data[:req.value]}

    def execute(self, N_c: int, B: torch.Tensor, C: torch.Tensor, load_c: bool,
                tile_config: Optional[TileConfig] = None, stream: Optional[torch.cuda.Stream] = None) -> None:
        cfg = (tile_config or TileConfig()).to_c()
        _lib.call("srt_tile_execute", self._h, int(N_c), ptr(B), B.stride(0), ptr(C), C.stride(0),
                  1 if load_c else 0, ct.byref(cfg), stream_handle(stream))

    def split_counts(self) -> np.ndarray:
        """Waves per panel of the last executed configuration."""
        kp = np.zeros(max(1, self.num_panels), dtype=np.int32)
        _lib.call("srt_tile_split_counts", self._h, kp.ctypes.data_as(ct.c_void_p))
        return kp[:self.num_panels]

    def packed(self) -> Dict[str, np.ndarray]:
        """Packed arrays on the host: per-panel nkern_counts, col/val offsets,
        col_indices, values (the MicroKernelPackedData fields, template:169-212)."""
        nc = ct.c_int64()
        nv = ct.c_int64()
        _lib.call("srt_tile_copy_packed", self._h, None, None, None, None, None, ct.byref(nc), ct.byref(nv))
        P, NK = self.num_panels, len(self.mapping.patterns)
        counts = np.zeros(max(1, P * NK), dtype=np.int32)
        col_ptr = np.zeros(P + 1, dtype=np.int64)
        val_ptr = np.zeros(P + 1, dtype=np.int64)
        cols = np.zeros(max(1, nc.value), dtype=np.int32)
        vals = np.zeros(max(1, nv.value), dtype=np.float32)
        _lib.call("srt_tile_copy_packed", self._h, counts.ctypes.data_as(ct.c_void_p),
                  col_ptr.ctypes.data_as(ct.c_void_p), val_ptr.ctypes.data_as(ct.c_void_p),
                  cols.ctypes.data_as(ct.c_void_p), vals.ctypes.data_as(ct.c_void_p), None, None)
        return {"counts": counts[:P * NK].reshape(P, NK), "col_ptr": col_ptr, "val_ptr": val_ptr,
                "cols": cols[:nc.value], "vals": vals[:nv.value], "patterns": list(self.mapping.patterns)}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().srt_tile_destroy(h)
            except Exception:
                pass
            self._h = ct.c_void_p()


def _time_runs(fn, num_runs: int, iterations: int) -> float:
    """Median ns per run over ``iterations`` blocks of ``num_runs`` runs,
    timed with events on the current stream (template:308-349)."""
    for _ in range(num_runs):
        fn()
    times = []
    for _ in range(iterations):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(num_runs):
            fn()
        e.record()
        e.synchronize()
        times.append(s.elapsed_time(e) * 1e6)
    return statistics.median(times) / float(num_runs)


def execute_tile(N_c: int, tile: SOPTile, B: torch.Tensor, num_runs: int = 1,
                 measured_iterations: int = MEASURED_ITERATIONS,
                 tile_config: Optional[TileConfig] = None) -> Tuple[float, torch.Tensor]:
    """(median ns per run, C = tile . B[:, :N_c]) (template:351-392)."""
    if not tile.tile_packed:
        raise RuntimeError("pack_tile has not been called")
    require_cuda_f32(B, "B")
    if B.shape[0] != tile.cols:
        raise ValueError("B has %d rows, tile has %d columns" % (B.shape[0], tile.cols))
    C = torch.zeros((tile.rows, B.shape[1]), dtype=torch.float32, device=B.device)
    ns = _time_runs(lambda: tile.execute(N_c, B, C, True, tile_config), num_runs, measured_iterations) \
        if num_runs > 0 else float("nan")
    C.zero_()
    tile.execute(N_c, B, C, True, tile_config)
    torch.cuda.current_stream().synchronize()
    return ns, C


class SOPModule:
    """sbench/SOP/sop_driver.py:38-85 on the GPU."""

    def __init__(self, acc: Acc, patterns: Sequence[int], pattern_mapping: Union[Callable, Mapping],
                 native: NativeMapping, kernel_id: str):
        self.acc = acc
        self.patterns = list(patterns)
        self.pattern_mapping = pattern_mapping
        self.native = native
        self.kernel_id = kernel_id

    def SOPTile(self, rows: int, cols: int, num_panels: int) -> SOPTile:
        return SOPTile(rows, cols, num_panels, self.native)

    def make_sop_tile(self, tile: torch.Tensor, col_offset: int = 0) -> SOPTile:
        """Pack a dense tile panel by panel (sop_driver.py:47-74): per column the
        code is split into kernel patterns by the mapping, then pack_panel."""
        M_r = self.acc.M
        t = torch.as_tensor(tile, dtype=torch.float32, device="cuda")
        rows, K = t.shape
        P = math.ceil(rows / M_r)
        codes = panel_pattern_codes(t, M_r)                                   # [P, K]
        # LUT code -> padded list of kernel patterns
        table = self.native.mapping.table
        maxlen = max([len(v) for v in table.values()] + [1])
        lut = torch.zeros((1 << M_r, maxlen), dtype=torch.int32)
        for code, pats in table.items():
            if pats:
                lut[code, :len(pats)] = torch.tensor(pats, dtype=torch.int32)
        lut = lut.cuda()
        padded = torch.zeros((P * M_r, K), dtype=torch.float32, device=t.device)
        padded[:rows] = t
        sop_tile = self.SOPTile(rows, K, P)
        cols_all = torch.arange(K, dtype=torch.int32, device=t.device)
        for p in range(P):
            ent = lut[codes[p]]                                               # [K, maxlen]
            keep = ent != 0
            pat_codes = ent[keep]
            col_idx = cols_all.unsqueeze(1).expand(K, maxlen)[keep]
            dense_panel = padded[p * M_r:(p + 1) * M_r].t().contiguous()     # K x M_r
            sop_tile.pack_panel(p, col_offset, pat_codes, col_idx, dense_panel)
        sop_tile.pack_tile()
        return sop_tile

    def execute_tile(self, tile: SOPTile, B: torch.Tensor, num_runs: int = 1024, N_c: Optional[int] = None,
                     measured_iterations: int = MEASURED_ITERATIONS,
                     tile_config: Optional[TileConfig] = None) -> Tuple[float, torch.Tensor]:
        N_c = B.shape[-1] if N_c is None else N_c
        if tile_config is None:
            tile_config = TileConfig(vec_width=self.acc.N if self.acc.N in (1, 2, 4) else 0)
        return execute_tile(N_c, tile, B, num_runs, measured_iterations, tile_config)


def make_sop_module(acc: Acc, patterns: Iterable[int], pattern_mapping: Union[Callable, Mapping],
                    regen: bool = False) -> SOPModule:
    """Module for a pattern set (sop_driver.py:88-129).  The executor for the
    pattern set must have been generated (libsrt ships executors for every
    mapping under mappings/; ``jit.ensure_executor`` builds others)."""
    patterns = list(patterns)
    if 0 in patterns:
        raise ValueError("Zero is not a valid pattern")
    # the generator receives list(patterns) (sop_driver.py:93): that list,
    # stable-sorted by popcount, fixes the kernel ids
    if isinstance(pattern_mapping, Mapping):
        table = pattern_mapping.table
    else:
        table = mapping_table(acc.M, lambda c: [p for p in pattern_mapping(c) if p])
    extra = {p for v in table.values() for p in v if p} - set(patterns)
    if extra:
        raise ValueError("mapping uses patterns %s that are not in the pattern set" % sorted(extra))
    m = Mapping(acc.M, table, kernel_list=patterns)
    native = NativeMapping.from_mapping(m)
    from . import jit
    kernel_id = jit.ensure_executor(native, regen=regen)
    return SOPModule(acc, patterns, m, native, kernel_id)
