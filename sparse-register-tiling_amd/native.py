"""Thin object layer over the C ABI: device tensors, streams, native mappings."""
from __future__ import annotations

import ctypes as ct
import os
from typing import Optional, Sequence, Union

import torch

from . import _lib
from .mapping import (MAPPINGS_DIR, STRATEGY_M_R, Mapping, find_mapping, named_mapping, read_mapping_file,
                      strategy_for)


def stream_handle(stream: Optional[torch.cuda.Stream] = None) -> ct.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ct.c_void_p(s.cuda_stream)


def require_cuda_f32(t: torch.Tensor, name: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch.Tensor" % name)
    if not t.is_cuda:
        raise ValueError("%s must be a CUDA (HIP) tensor" % name)
    if t.dtype != torch.float32:
        raise TypeError("%s must be float32, got %s" % (name, t.dtype))
    if t.dim() != 2:
        raise ValueError("%s must be 2-D" % name)
    if t.stride(1) != 1:
        raise ValueError("%s must have unit column stride (row-major)" % name)
    return t


def ptr(t: Optional[torch.Tensor]) -> ct.c_void_p:
    return ct.c_void_p(t.data_ptr() if t is not None else 0)


class NativeMapping:
    """An srt_mapping handle.  Kernel ids are the stable popcount sort of the
    mapping's kernel list (``Mapping.kernel_list``)."""

    def __init__(self, handle: ct.c_void_p, mapping: Optional[Mapping] = None):
        self._h = handle
        self.mapping = mapping
        M_r = ct.c_int()
        nk = ct.c_int()
        _lib.call("srt_mapping_info", self._h, ct.byref(M_r), ct.byref(nk))
        self.M_r = M_r.value
        arr = (ct.c_int32 * max(1, nk.value))()
        _lib.call("srt_mapping_pattern_table", self._h, arr)
        self.patterns = [arr[i] for i in range(nk.value)]

    @classmethod
    def from_mapping(cls, m: Mapping) -> "NativeMapping":
        code_ptr, flat = m.code_arrays()
        h = ct.c_void_p()
        _lib.call("srt_mapping_create_ordered", m.M_r, _lib.int32_array(code_ptr), _lib.int32_array(flat),
                  _lib.int32_array(m.kernel_list or [0]), len(m.kernel_list), ct.byref(h))
        native = cls(h, m)
        if native.patterns != m.patterns:
            raise RuntimeError("kernel id order differs between the C library and the mapping")
        return native

    @classmethod
    def from_id(cls, mapping_id: str, search_dirs: Optional[Sequence[str]] = None) -> "NativeMapping":
        """Read mapping_<id>.txt through the C reader (srt_mapping_read): the
        file's kernel list, as the reference's load_mapping builds it."""
        dirs = [os.fsencode(d) for d in list(search_dirs or []) + [MAPPINGS_DIR]]
        arr = (ct.c_char_p * len(dirs))(*dirs)
        h = ct.c_void_p()
        _lib.call("srt_mapping_read", mapping_id.encode(), arr, len(dirs), ct.byref(h))
        return cls(h, read_mapping_file(find_mapping(mapping_id, search_dirs)))

    def executor_id(self) -> str:
        buf = ct.create_string_buffer(128)
        _lib.call("srt_executor_id", self._h, buf, 128)
        return buf.value.decode()

    @property
    def handle(self) -> ct.c_void_p:
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().srt_mapping_destroy(h)
            except Exception:
                pass
            self._h = ct.c_void_p()


def resolve_mapping(mapping: Union[str, Mapping, NativeMapping], M_r: Optional[int] = None,
                    search_dirs: Optional[Sequence[str]] = None) -> NativeMapping:
    """A NativeMapping from an id ("400fa"), a config name with M_r ("orig"),
    a Mapping, or a NativeMapping.  Names and ids of the shipped strategies
    carry the strategy's kernel list (``named_mapping``); other ids are read
    by the C reader."""
    if isinstance(mapping, NativeMapping):
        return mapping
    if isinstance(mapping, Mapping):
        return NativeMapping.from_mapping(mapping)
    if M_r is not None:
        return NativeMapping.from_mapping(named_mapping(M_r, mapping, search_dirs))
    for mr in sorted(set(STRATEGY_M_R.values())):
        if strategy_for(mr, mapping) is not None:
            return NativeMapping.from_mapping(named_mapping(mr, mapping, search_dirs))
    return NativeMapping.from_id(mapping, search_dirs)


def executors():
    """(executor id, M_r, patterns) of every generated executor."""
    L = _lib.lib()
    out = []
    for i in range(L.srt_executor_count()):
        M_r = ct.c_int()
        nk = ct.c_int()
        _lib.call("srt_executor_describe", i, ct.byref(M_r), ct.byref(nk), None, None, 0)
        arr = (ct.c_int32 * max(1, nk.value))()
        buf = ct.create_string_buffer(128)
        _lib.call("srt_executor_describe", i, None, None, arr, buf, 128)
        out.append((buf.value.decode(), M_r.value, [arr[j] for j in range(nk.value)]))
    return out
