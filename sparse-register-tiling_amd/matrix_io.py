"""On-disk sparse matrix formats the reference's benchmarks read.

  * SMTX (the DLMC pruned-model format; cpp_testbed/include/ParseSMTX.h:15-55,
    sbench/cpp/smtx_loader.cpp:5-45): line 1 "rows, cols, nnz" (commas or
    spaces), then the rows + 1 row pointers, then on the next line the nnz
    column indices, all whitespace separated.  The file holds no values:
    every stored entry is 1.0.
  * MatrixMarket coordinate (cpp_testbed/include/ParseMatrixMarket.h:48-126):
    banner "%%MatrixMarket matrix coordinate <field> <symmetry>", comment lines
    starting with '%', a size line "rows cols nnz", then one 1-based
    "row col value" triplet per line.  A symmetric file stores one triangle;
    every off-diagonal entry is mirrored.  Fields real / integer / pattern
    are read (pattern entries are 1.0; the reference requires a value
    column, so pattern files are an extension here).
  * File lists (tools/filelists/*.txt): one matrix path per line, optionally
    followed by ",<H*W>" (the activation size paired with a 1x1-conv weight,
    tools/filelists/mobilenet_70.txt:1-11).

Matrices come back as CSR arrays (rows, cols, Lp int32 [rows + 1], Li int32
[nnz], Lx float32 [nnz]) with the column indices of a row ascending.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np


@dataclass
class CSRArrays:
    rows: int
    cols: int
    Lp: np.ndarray
    Li: np.ndarray
    Lx: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.Li.size)

    def to_dense(self) -> np.ndarray:
        A = np.zeros((self.rows, self.cols), dtype=np.float32)
        r = np.repeat(np.arange(self.rows), np.diff(self.Lp))
        np.add.at(A, (r, self.Li), self.Lx)
        return A

    def to_torch(self):
        import torch
        return torch.sparse_csr_tensor(torch.from_numpy(self.Lp.astype(np.int64)),
                                       torch.from_numpy(self.Li.astype(np.int64)),
                                       torch.from_numpy(self.Lx), size=(self.rows, self.cols))


def _check_csr(rows: int, cols: int, Lp: np.ndarray, Li: np.ndarray, what: str) -> None:
    if rows < 0 or cols < 0:
        raise ValueError("%s: negative dimension" % what)
    if Lp.size != rows + 1 or Lp[0] != 0 or np.any(np.diff(Lp) < 0) or Lp[-1] != Li.size:
        raise ValueError("%s: row pointers are not a CSR offset array for %d entries" % (what, Li.size))
    if Li.size and (Li.min() < 0 or Li.max() >= cols):
        raise ValueError("%s: column index out of range [0, %d)" % (what, cols))


# ---------------------------------------------------------------- SMTX
def read_smtx(path: str) -> CSRArrays:
    """Read an SMTX file (values 1.0)."""
    with open(path) as f:
        header = f.readline().replace(",", " ").split()
        if len(header) < 3:
            raise ValueError("%s: header must be 'rows, cols, nnz'" % path)
        rows, cols, nnz = (int(x) for x in header[:3])
        rest = f.read().split()
    if len(rest) < rows + 1 + nnz:
        raise ValueError("%s: expected %d row pointers and %d column indices, found %d numbers"
                         % (path, rows + 1, nnz, len(rest)))
    Lp = np.array(rest[:rows + 1], dtype=np.int64)
    Li = np.array(rest[rows + 1:rows + 1 + nnz], dtype=np.int64)
    _check_csr(rows, cols, Lp, Li, path)
    Li32 = Li.astype(np.int32)
    Lp32 = Lp.astype(np.int32)
    _sort_rows(Lp32, Li32, None)  # DLMC files are already sorted
    return CSRArrays(rows, cols, Lp32, Li32, np.ones(nnz, dtype=np.float32))


def write_smtx(path: str, rows: int, cols: int, Lp, Li) -> None:
    """Write the sparsity pattern in SMTX form (values are not stored)."""
    Lp = np.asarray(Lp, dtype=np.int64)
    Li = np.asarray(Li, dtype=np.int64)
    _check_csr(rows, cols, Lp, Li, path)
    with open(path, "w") as f:
        f.write("%d, %d, %d\n" % (rows, cols, Li.size))
        f.write(" ".join(str(int(x)) for x in Lp) + "\n")
        f.write(" ".join(str(int(x)) for x in Li) + "\n")


# ---------------------------------------------------------------- MatrixMarket
def read_mtx(path: str) -> CSRArrays:
    """Read a MatrixMarket coordinate file into CSR (duplicates are kept as
    separate entries, each row's entries ordered by column, stably)."""
    with open(path) as f:
        banner = f.readline().split()
        if len(banner) < 5 or banner[0].lower() != "%%matrixmarket" or banner[1].lower() != "matrix":
            raise ValueError("%s: not a MatrixMarket matrix file" % path)
        if banner[2].lower() != "coordinate":
            raise ValueError("%s: only coordinate (sparse) MatrixMarket files are read" % path)
        field, sym = banner[3].lower(), banner[4].lower()
        if field not in ("real", "integer", "pattern", "double"):
            raise ValueError("%s: field %r is not supported" % (path, field))
        if sym not in ("general", "symmetric"):
            raise ValueError("%s: symmetry %r is not supported" % (path, sym))
        line = f.readline()
        while line and (line.startswith("%") or not line.strip()):
            line = f.readline()
        rows, cols, nnz = (int(x) for x in line.split()[:3])
        ncol = 2 if field == "pattern" else 3
        data = np.array(f.read().split()[:nnz * ncol], dtype=np.float64)
    if data.size != nnz * ncol:
        raise ValueError("%s: expected %d entries" % (path, nnz))
    data = data.reshape(nnz, ncol)
    r = data[:, 0].astype(np.int64) - 1
    c = data[:, 1].astype(np.int64) - 1
    v = data[:, 2] if ncol == 3 else np.ones(nnz)
    if nnz and (r.min() < 0 or r.max() >= rows or c.min() < 0 or c.max() >= cols):
        raise ValueError("%s: entry outside the %d x %d matrix" % (path, rows, cols))
    if sym == "symmetric":
        off = r != c
        r, c, v = np.concatenate([r, c[off]]), np.concatenate([c, r[off]]), np.concatenate([v, v[off]])
    order = np.lexsort((c, r))  # by row, then column; stable for duplicates
    r, c, v = r[order], c[order], v[order]
    Lp = np.zeros(rows + 1, dtype=np.int32)
    np.add.at(Lp, r + 1, 1)
    Lp = np.cumsum(Lp).astype(np.int32)
    return CSRArrays(rows, cols, Lp, c.astype(np.int32), v.astype(np.float32))


def write_mtx(path: str, csr: CSRArrays) -> None:
    """Write CSR as a general real coordinate MatrixMarket file."""
    r = np.repeat(np.arange(csr.rows), np.diff(csr.Lp))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n%\n")
        f.write("%d %d %d\n" % (csr.rows, csr.cols, csr.nnz))
        for i, j, x in zip(r, csr.Li, csr.Lx):
            f.write("%d %d %.9e\n" % (i + 1, j + 1, x))


def _sort_rows(Lp: np.ndarray, Li: np.ndarray, Lx: Optional[np.ndarray]) -> None:
    """Order each row's entries by column (stable), in place."""
    rows = Lp.size - 1
    if Li.size == 0:
        return
    r = np.repeat(np.arange(rows), np.diff(Lp))
    order = np.lexsort((Li, r))
    Li[:] = Li[order]
    if Lx is not None:
        Lx[:] = Lx[order]


# ---------------------------------------------------------------- file lists
def read_filelist(path: str, root: Optional[str] = None) -> List[Tuple[str, Optional[int]]]:
    """(matrix path, H*W or None) per non-empty line; relative paths are
    joined to `root` (default: the list's directory)."""
    base = root if root is not None else os.path.dirname(os.path.abspath(path))
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            parts = [p.strip() for p in line.split(",")]
            p = parts[0] if os.path.isabs(parts[0]) else os.path.join(base, parts[0])
            out.append((p, int(parts[1]) if len(parts) > 1 and parts[1] else None))
    return out


def read_matrix(path: str) -> CSRArrays:
    """Dispatch on the extension: .smtx or .mtx."""
    ext = os.path.splitext(path)[1].lower()
    if ext == ".smtx":
        return read_smtx(path)
    if ext == ".mtx":
        return read_mtx(path)
    raise ValueError("%s: unknown matrix format (expected .smtx or .mtx)" % path)
