"""ctypes front end of the C oracle (oracle/srt_oracle.c).

TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py cpu_baseline).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
from typing import Dict, List

import numpy as np

from . import sop_oracle as so

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libsrt_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_SO):
        build()
    L = ct.CDLL(_SO)
    L.srt_oracle_pack_sizes.argtypes = [ct.c_int, ct.c_int, _f32p, ct.c_longlong, ct.c_int,
                                        _i32p, _i32p, ct.c_int, _i32p, _i64p, _i64p]
    L.srt_oracle_pack_sizes.restype = ct.c_int
    L.srt_oracle_pack_fill.argtypes = [ct.c_int, ct.c_int, _f32p, ct.c_longlong, ct.c_int,
                                       _i32p, _i32p, ct.c_int, _i32p, ct.c_int, _i64p, _i64p,
                                       _i32p, _i32p, _f32p]
    L.srt_oracle_pack_fill.restype = ct.c_longlong
    L.srt_oracle_exec.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, _i32p, _i32p, _i64p, _i64p,
                                  _i32p, _f32p, _f32p, ct.c_longlong, _f32p, ct.c_longlong,
                                  ct.c_int, _i32p, ct.c_int]
    L.srt_oracle_exec_kt.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _i32p, _i32p, _i64p,
                                     _i64p, _i32p, _f32p, _f32p, ct.c_longlong, _f32p, ct.c_longlong,
                                     ct.c_int, _i32p, ct.c_int]
    L.srt_oracle_exec_kt.restype = ct.c_int
    L.srt_oracle_isa.restype = ct.c_int
    L.srt_oracle_exec.restype = ct.c_int
    L.srt_oracle_csr.argtypes = [ct.c_int, ct.c_int, _i64p, _i32p, _f32p, _f32p, ct.c_longlong,
                                 _f32p, ct.c_longlong, ct.c_int]
    L.srt_oracle_csr.restype = ct.c_int
    L.srt_oracle_dense_f64.argtypes = [ct.c_int, ct.c_int, ct.c_int, _f32p, ct.c_longlong, _f32p,
                                       ct.c_longlong, _f64p, ct.c_longlong, ct.c_int]
    L.srt_oracle_dense_f64.restype = ct.c_int
    _lib = L
    return L


def build_lut(M_r: int, table: Dict[int, List[int]], patterns: List[int]):
    """Flatten a mapping table into (lut_ptr[2^M_r+1], lut_ids[]) of kernel ids."""
    enc = so.encoder(patterns)
    ptr = np.zeros(2 ** M_r + 1, dtype=np.int32)
    ids: List[int] = []
    for code in range(2 ** M_r):
        ptr[code] = len(ids)
        if code:
            for p in table.get(code, []):
                if p:
                    ids.append(enc(int(p)))
    ptr[2 ** M_r] = len(ids)
    return ptr, np.array(ids if ids else [0], dtype=np.int32)


def pack_dense_ktiled(A: np.ndarray, M_r: int, table: Dict[int, List[int]], k_tile: int, kernel_list=None):
    """K-tiled packing (count_nkerns.cpp:52-126: A cut into K_c-column tiles,
    each panel packed per tile): every K slice is packed on its own with its
    column offset and the slices' packed columns are laid one after another
    per panel; counts become [P, KT, NK]."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    M, K = A.shape
    if not (0 < k_tile < K):
        return pack_dense(A, M_r, table, kernel_list=kernel_list)
    slices = [pack_dense(A[:, k0:min(K, k0 + k_tile)], M_r, table, col_offset=k0, kernel_list=kernel_list)
              for k0 in range(0, K, k_tile)]
    P, NK = slices[0]["P"], slices[0]["NK"]
    cols, vals = [], []
    col_ptr = np.zeros(P + 1, dtype=np.int64)
    val_ptr = np.zeros(P + 1, dtype=np.int64)
    for p in range(P):
        for sl in slices:
            cols.append(sl["cols"][sl["col_ptr"][p]:sl["col_ptr"][p + 1]])
            vals.append(sl["vals"][sl["val_ptr"][p]:sl["val_ptr"][p + 1]])
        col_ptr[p + 1] = col_ptr[p] + sum(sl["col_ptr"][p + 1] - sl["col_ptr"][p] for sl in slices)
        val_ptr[p + 1] = val_ptr[p] + sum(sl["val_ptr"][p + 1] - sl["val_ptr"][p] for sl in slices)
    counts = np.stack([sl["counts"] for sl in slices], axis=1) if NK else np.zeros((P, len(slices), 0), np.int32)
    return {
        "M": M, "K": K, "M_r": M_r, "P": P, "NK": NK, "KT": len(slices), "patterns": slices[0]["patterns"],
        "counts": np.ascontiguousarray(counts, dtype=np.int32),
        "col_ptr": col_ptr, "val_ptr": val_ptr,
        "cols": np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32),
        "vals": np.concatenate(vals).astype(np.float32) if vals else np.zeros(0, np.float32),
        "padding": sum(sl["padding"] for sl in slices),
    }


def pack_dense(A: np.ndarray, M_r: int, table: Dict[int, List[int]], col_offset: int = 0, kernel_list=None):
    """Pack a dense row-major fp32 matrix with the mapping ``table``; kernel
    ids from ``kernel_list`` (default: the table's own, see
    ``sop_oracle.kernel_pattern_table``)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    M, K = A.shape
    patterns = so.kernel_pattern_table(table, kernel_list)
    NK = len(patterns)
    lut_ptr, lut_ids = build_lut(M_r, table, patterns)
    id_pat = np.array(patterns, dtype=np.int32)
    P = (M + M_r - 1) // M_r
    col_ptr = np.zeros(P + 1, dtype=np.int64)
    val_ptr = np.zeros(P + 1, dtype=np.int64)
    L = lib()
    if L.srt_oracle_pack_sizes(M, K, A, K, M_r, lut_ptr, lut_ids, NK, id_pat, col_ptr, val_ptr):
        raise ValueError("mapping refers to an id outside the kernel pattern table")
    counts = np.zeros(max(1, P * NK), dtype=np.int32)
    cols = np.zeros(max(1, int(col_ptr[-1])), dtype=np.int32)
    vals = np.zeros(max(1, int(val_ptr[-1])), dtype=np.float32)
    padding = L.srt_oracle_pack_fill(M, K, A, K, M_r, lut_ptr, lut_ids, NK, id_pat, col_offset,
                                     col_ptr, val_ptr, counts, cols, vals)
    return {
        "M": M, "K": K, "M_r": M_r, "P": P, "NK": NK, "patterns": patterns,
        "counts": counts[:P * NK].reshape(P, NK) if NK else counts[:0],
        "col_ptr": col_ptr, "val_ptr": val_ptr,
        "cols": cols[:int(col_ptr[-1])], "vals": vals[:int(val_ptr[-1])],
        "padding": int(padding),
    }


def exec_packed(packed, B: np.ndarray, nthreads: int = 0, C: np.ndarray | None = None,
                load_c: bool = False, nsplit=1) -> np.ndarray:
    """Run the restated executor over a ``pack_dense`` result.  ``nsplit`` is
    the number of chunks per panel, one int for every panel or an array with
    one entry per panel: 1 is the reference's single chain per panel; k > 1
    sums the chains of k contiguous chunks left to right (the GPU executor's
    split order)."""
    B = np.ascontiguousarray(B, dtype=np.float32)
    M, M_r, NK = packed["M"], packed["M_r"], packed["NK"]
    N = B.shape[1]
    kp = np.ascontiguousarray(np.broadcast_to(np.asarray(nsplit, dtype=np.int32), (max(1, packed["P"]),)))
    if C is None:
        C = np.zeros((M, N), dtype=np.float32)
    counts = np.ascontiguousarray(packed["counts"], dtype=np.int32).reshape(-1)
    if counts.size == 0:
        counts = np.zeros(1, dtype=np.int32)
    cols = packed["cols"] if packed["cols"].size else np.zeros(1, dtype=np.int32)
    vals = packed["vals"] if packed["vals"].size else np.zeros(1, dtype=np.float32)
    id_pat = np.array(packed["patterns"] or [0], dtype=np.int32)
    rc = lib().srt_oracle_exec_kt(M, N, M_r, NK, packed.get("KT", 1), id_pat, counts, packed["col_ptr"],
                                  packed["val_ptr"], np.ascontiguousarray(cols, dtype=np.int32),
                                  np.ascontiguousarray(vals, dtype=np.float32), B, B.shape[1], C, N,
                                  1 if load_c else 0, kp, nthreads)
    if rc:
        raise MemoryError("oracle executor failed (%d)" % rc)
    return C


def isa() -> str:
    """Inner-loop variant the C oracle runs on this host."""
    return {2: "avx512", 1: "avx2+fma", 0: "x86-64"}[lib().srt_oracle_isa()]


def csr_exec(rowptr, colind, vals, M: int, B: np.ndarray, nthreads: int = 0) -> np.ndarray:
    """Restated CSR executor (one fmaf per stored entry, storage order)."""
    B = np.ascontiguousarray(B, dtype=np.float32)
    N = B.shape[1]
    C = np.zeros((M, N), dtype=np.float32)
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(colind, dtype=np.int32)
    vv = np.ascontiguousarray(vals, dtype=np.float32)
    if ci.size == 0:
        ci = np.zeros(1, dtype=np.int32)
        vv = np.zeros(1, dtype=np.float32)
    lib().srt_oracle_csr(M, N, rp, ci, vv, B, N, C, N, nthreads)
    return C


def dense_f64(A: np.ndarray, B: np.ndarray, nthreads: int = 0) -> np.ndarray:
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N), dtype=np.float64)
    lib().srt_oracle_dense_f64(M, K, N, A, K, B, N, C, N, nthreads)
    return C
