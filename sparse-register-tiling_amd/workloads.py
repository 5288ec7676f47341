"""Seeded synthetic inputs for the benchmark configurations.

All generators draw from ``numpy.random.default_rng(seed)`` in the order
stated, so every caller (tests, bench, smoke) sees the same matrices.

* ``random_sparse``      independent Bernoulli(1 - sparsity) mask x N(0,1)
                         values (the CPU configuration: 2048 x 512 at p=0.2).
* ``magnitude_pruned``   W = N(0,1)[M x K] * lognormal(0, 0.5)[M x 1], keep
                         the (1 - s)*M*K largest |W| (uneven row densities).
* ``dense_b``            B ~ U[-1, 1].
"""
from __future__ import annotations

from typing import List, NamedTuple, Tuple

import numpy as np


class SpMMCase(NamedTuple):
    M: int
    K: int
    sparsity: float
    N: int

    @property
    def name(self) -> str:
        return "%dx%d_s%02d_N%d" % (self.M, self.K, round(self.sparsity * 100), self.N)


# Pruned-FFN-like GPU cases: shapes x sparsities x N.
GPU_SHAPES = [(512, 512), (2048, 512), (512, 2048), (1024, 4096)]
GPU_SPARSITIES = [0.70, 0.80, 0.90, 0.95]
GPU_NS = [128, 256, 512]
# The headline case: 2048 x 512 at 90% sparsity times 512 x 512.
PRUNED_2048_512_S90_N512 = SpMMCase(2048, 512, 0.90, 512)
CPU_CASE = SpMMCase(2048, 512, 0.80, 256)


def gpu_cases() -> List[SpMMCase]:
    return [SpMMCase(M, K, s, N) for (M, K) in GPU_SHAPES for s in GPU_SPARSITIES for N in GPU_NS]


def random_sparse(M: int, K: int, density: float, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    mask = rng.random((M, K)) < density
    vals = rng.standard_normal((M, K))
    return (mask * vals).astype(np.float32)


def magnitude_pruned(M: int, K: int, sparsity: float, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    W = rng.standard_normal((M, K)) * rng.lognormal(0.0, 0.5, (M, 1))
    keep = int(round((1.0 - sparsity) * M * K))
    A = np.zeros((M, K), dtype=np.float32)
    if keep > 0:
        flat = np.abs(W).ravel()
        idx = np.argpartition(flat, flat.size - keep)[flat.size - keep:]
        A.ravel()[idx] = W.ravel()[idx].astype(np.float32)
    return A


def dense_b(K: int, N: int, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, (K, N)).astype(np.float32)


def cpu_case(seed: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """A 2048 x 512 at density 0.2 with N(0,1) values, B 512 x 256 U[-1,1];
    one generator, mask then values then B."""
    rng = np.random.default_rng(seed)
    mask = rng.random((2048, 512)) < 0.2
    vals = rng.standard_normal((2048, 512))
    A = (mask * vals).astype(np.float32)
    B = rng.uniform(-1.0, 1.0, (512, 256)).astype(np.float32)
    return A, B


def case_inputs(case: SpMMCase, seed: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    A = magnitude_pruned(case.M, case.K, case.sparsity, seed)
    B = dense_b(case.K, case.N, seed + 1)
    return A, B


def to_csr(A: np.ndarray, index_dtype=np.int64):
    """(rowptr, colind, vals) of a dense matrix, row-major, columns ascending."""
    nz = A != 0
    rowptr = np.zeros(A.shape[0] + 1, dtype=index_dtype)
    np.cumsum(nz.sum(axis=1), out=rowptr[1:])
    r, c = np.nonzero(nz)
    return rowptr, c.astype(index_dtype), A[r, c].astype(np.float32)


def flops(nnz: int, N: int) -> float:
    """Algorithmic flops of one SpMM: 2 * nnz * N (padding excluded)."""
    return 2.0 * nnz * N


def bytes_alg(nnz: int, M: int, K: int, N: int, s_a: int = 4, s_b: int = 4, s_c: int = 4,
              scale_bytes: int = 0) -> float:
    """Format-independent lower bound of HBM bytes of one SpMM: every true
    nonzero once, B once, C written once (+ per-row scales)."""
    return float(nnz * s_a + K * N * s_b + M * N * s_c + scale_bytes)
