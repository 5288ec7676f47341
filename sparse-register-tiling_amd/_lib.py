"""ctypes binding of libsrt.so (the C ABI declared in include/srt.h).

The library is loaded from this package's directory.  If it has not been
built the import of this module fails loudly: there is no fallback path.
"""
from __future__ import annotations

import ctypes as ct
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libsrt.so")

SRT_OK = 0
SRT_ERR_INVALID = 1
SRT_ERR_UNSUPPORTED = 2
SRT_ERR_HIP = 3
SRT_ERR_IO = 4
SRT_ERR_NOMEM = 5
ZERO_PATTERN_ID = 32768
MAX_MR = 16


class SrtError(RuntimeError):
    """A libsrt call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__("libsrt error %d: %s" % (code, msg))
        self.code = code


class LibraryNotBuilt(ImportError):
    pass


class TileConfigC(ct.Structure):
    _fields_ = [
        ("M_c", ct.c_int), ("K_c", ct.c_int), ("N_c", ct.c_int),
        ("tiling_strategy", ct.c_int), ("beta", ct.c_float), ("sparse_a", ct.c_int),
        ("max_tlb_entries", ct.c_int), ("tlb_page_size", ct.c_int),
        ("runtime_schedule", ct.c_int), ("vec_width", ct.c_int),
        ("waves_per_block", ct.c_int), ("xcd_remap", ct.c_int), ("unroll", ct.c_int),
        ("nsplit", ct.c_int), ("chunk_cols", ct.c_int), ("b_staging", ct.c_int),
        ("layout", ct.c_int), ("matrix_cores", ct.c_int),
    ]


class MfmaStatsC(ct.Structure):
    _fields_ = [("M", ct.c_int), ("K", ct.c_int), ("num_panels", ct.c_int), ("k_blocks", ct.c_int),
                ("dtype", ct.c_int), ("nnz", ct.c_int64), ("stored_blocks", ct.c_int64),
                ("nonzero_blocks", ct.c_int64)]


SRT_DT_BF16 = 1
SRT_DT_FP8 = 2


class PlanStatsC(ct.Structure):
    _fields_ = [
        ("M", ct.c_int), ("K", ct.c_int), ("M_r", ct.c_int), ("num_panels", ct.c_int),
        ("num_patterns", ct.c_int), ("nnz", ct.c_int64), ("packed_cols", ct.c_int64),
        ("packed_values", ct.c_int64), ("padding", ct.c_int64), ("k_tile", ct.c_int), ("k_tiles", ct.c_int),
    ]


_vp = ct.c_void_p
_i = ct.c_int
_i64 = ct.c_int64
_p_i32 = ct.POINTER(ct.c_int32)
_p_i64 = ct.POINTER(ct.c_int64)

# name -> (restype, argtypes)
SIGNATURES = {
    "srt_last_error": (ct.c_char_p, []),
    "srt_version": (_i, []),
    "srt_tile_config_default": (None, [ct.POINTER(TileConfigC)]),
    "srt_mapping_read": (_i, [ct.c_char_p, ct.POINTER(ct.c_char_p), _i, ct.POINTER(_vp)]),
    "srt_mapping_create": (_i, [_i, _p_i32, _p_i32, ct.POINTER(_vp)]),
    "srt_mapping_create_ordered": (_i, [_i, _p_i32, _p_i32, _p_i32, _i, ct.POINTER(_vp)]),
    "srt_mapping_info": (_i, [_vp, ct.POINTER(_i), ct.POINTER(_i)]),
    "srt_mapping_pattern_table": (_i, [_vp, _p_i32]),
    "srt_mapping_destroy": (None, [_vp]),
    "srt_executor_id": (_i, [_vp, ct.c_char_p, ct.c_size_t]),
    "srt_executor_count": (_i, []),
    "srt_executor_describe": (_i, [_i, ct.POINTER(_i), ct.POINTER(_i), _p_i32, ct.c_char_p, ct.c_size_t]),
    "srt_plan_create": (_i, [_i, _i, _i, _vp, _vp, _vp, _i, ct.POINTER(TileConfigC), _i,
                             ct.c_char_p, _vp, _vp, ct.POINTER(_vp)]),
    "srt_plan_create_hybrid": (_i, [_i, _i, _i, _vp, _vp, _vp, _i, ct.POINTER(TileConfigC), ct.c_float, _vp, _vp,
                                    ct.POINTER(_vp)]),
    "srt_plan_dense_panels": (_i, [_vp, ct.POINTER(_i), _vp, ct.POINTER(ct.c_size_t)]),
    "srt_plan_get_config": (_i, [_vp, ct.POINTER(TileConfigC)]),
    "srt_plan_required_storage": (_i, [_vp, ct.POINTER(ct.c_size_t)]),
    "srt_plan_allocate_executor": (_i, [_vp, _i]),
    "srt_plan_execute": (_i, [_vp, _vp, _i64, _vp, _i64, _i, _vp]),
    "srt_plan_stats_get": (_i, [_vp, ct.POINTER(PlanStatsC)]),
    "srt_plan_copy_packed": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "srt_plan_destroy": (None, [_vp]),
    "srt_group_create": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "srt_group_rebind": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "srt_group_execute": (_i, [_vp, _vp]),
    "srt_group_info": (_i, [_vp, _vp, _vp, _vp]),
    "srt_group_destroy": (None, [_vp]),
    "srt_plan_release_columns": (_i, [_vp]),
    "srt_plan_schedule_bytes": (_i, [_vp, ct.POINTER(ct.c_size_t)]),
    "srt_plan_launch_shape": (_i, [_vp, _i, ct.POINTER(_i), ct.POINTER(_i)]),
    "srt_probe_launch_floor": (_i, [_i, _i, _i, _vp, _i64, _i, _vp, _vp]),
    "srt_plan_split_counts": (_i, [_vp, _vp]),
    "srt_tile_split_counts": (_i, [_vp, _vp]),
    "srt_sort_and_encode_patterns": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "srt_tile_create": (_i, [_vp, _i64, _i64, _i64, ct.POINTER(_vp)]),
    "srt_tile_pack_panel": (_i, [_vp, _i64, _i64, _vp, _vp, _i, _vp, _i, _vp]),
    "srt_tile_pack_tile": (_i, [_vp, _vp]),
    "srt_tile_buffer": (_i, [_vp, ct.POINTER(_vp), ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_size_t), _p_i64]),
    "srt_tile_copy_buffer": (_i, [_vp, _vp, ct.c_size_t]),
    "srt_tile_padding": (_i, [_vp, _p_i64]),
    "srt_tile_execute": (_i, [_vp, _i, _vp, _i64, _vp, _i64, _i, ct.POINTER(TileConfigC), _vp]),
    "srt_tile_copy_packed": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _p_i64, _p_i64]),
    "srt_tile_destroy": (None, [_vp]),
    "srt_exec_problem_fill": (_i, [_vp, _vp, _i64, _vp, _i64, _i, _i, _i, _vp]),
    "srt_exec_persistent_execute": (_i, [_vp, _i, _vp, _vp, _i, _vp]),
    "srt_widen_problem_tiles": (_i, [_vp]),
    "srt_exec_grouped_b_widen": (_i, [_vp, _vp, _i, _vp]),
    "srt_csr_spmm": (_i, [_i, _i, _i, _i64, _vp, _vp, _i, _vp, _vp, _i64, _vp, _i64, _i, _vp]),
    "srt_mfma_plan_create": (_i, [_i, _i, _vp, _vp, _vp, _i, _i, _vp, ct.POINTER(_vp)]),
    "srt_mfma_plan_reserve": (_i, [_vp, _i, _i]),
    "srt_mfma_convert_b": (_i, [_i, _i64, _i, _vp, _i64, _vp, _i64, _vp, _vp]),
    "srt_mfma_plan_execute": (_i, [_vp, _vp, _i64, _vp, _i64, _vp, _i, _i, _i, _vp]),
    "srt_mfma_plan_execute_noalloc": (_i, [_vp, _vp, _i64, _vp, _i64, _vp, _i, _i, _i, _vp]),
    "srt_mfma_plan_stats_get": (_i, [_vp, ct.POINTER(MfmaStatsC)]),
    "srt_mfma_plan_row_scales": (_i, [_vp, _vp]),
    "srt_mfma_plan_copy_blocks": (_i, [_vp, _vp, _vp, _vp]),
    "srt_mfma_plan_destroy": (None, [_vp]),
    "srt_mfma_problem_fill": (_i, [_vp, _vp, _i64, _vp, _vp, _i64, _i, _i, _vp]),
    "srt_mfma_problem_tiles": (_i, [_vp, _i]),
    "srt_mfma_grouped_execute": (_i, [_i, _i, _vp, _vp, _i, _vp]),
    "srt_bprep_problem_tiles": (_i, [_vp]),
    "srt_mfma_grouped_b_prep": (_i, [_i, _vp, _vp, _i, _vp]),
}

_lib = None


def lib() -> ct.CDLL:
    """The loaded libsrt (raises LibraryNotBuilt when libsrt.so is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryNotBuilt(
            "%s is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python sparse-register-tiling_amd/build_lib.py`" % LIB_PATH)
    L = ct.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    msg = lib().srt_last_error()
    return msg.decode() if msg else ""


def check(code: int) -> None:
    if code != SRT_OK:
        raise SrtError(code, last_error())


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def int32_array(values):
    arr = (ct.c_int32 * max(1, len(values)))()
    for i, v in enumerate(values):
        arr[i] = int(v)
    return arr
