"""Pins the C oracle's executor order against an independent restatement.

The restatement below walks the packed data in Python and performs every
fused multiply-add exactly (rational arithmetic, one rounding to float32,
ties to even), so it pins the oracle bit for bit:
  * nsplit = 1: one chain per output element in packed (id, column) order,
    the order of the reference's executor (template:281-297,
    gen_vector_specialized.py:266-296);
  * nsplit > 1: chunk w = packed columns [floor(w*n/nsplit),
    floor((w+1)*n/nsplit)) of the panel runs its own chain, chunks added
    left to right.
Inputs are drawn so that a different order would round differently.
"""
from fractions import Fraction

import numpy as np
import pytest


def _round_f32(x: Fraction) -> np.float32:
    """Nearest float32 to an exact rational (ties to even)."""
    f = np.float32(float(x))
    best = f
    for cand in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        if not np.isfinite(cand):
            continue
        d = abs(Fraction(float(cand)) - x)
        db = abs(Fraction(float(best)) - x)
        if d < db or (d == db and (int(cand.view(np.uint32)) & 1) == 0 and cand != best):
            best = cand
    return np.float32(best)


def _fma(a, b, c) -> np.float32:
    return _round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def _add(a, b) -> np.float32:
    return _round_f32(Fraction(float(a)) + Fraction(float(b)))


def _exec_exact(pk, B, nsplit):
    M, M_r = pk["M"], pk["M_r"]
    N = B.shape[1]
    C = np.zeros((M, N), dtype=np.float32)
    for p in range(pk["P"]):
        c0, c1 = int(pk["col_ptr"][p]), int(pk["col_ptr"][p + 1])
        ncols = c1 - c0
        part = np.zeros((nsplit, M_r, N), dtype=np.float32)
        c = 0
        v = int(pk["val_ptr"][p])
        for kid, cnt in enumerate(pk["counts"][p]):
            pat = pk["patterns"][kid]
            for _ in range(int(cnt)):
                w = 0
                while w < nsplit - 1 and c >= (w + 1) * ncols // nsplit:
                    w += 1
                b = B[pk["cols"][c0 + c]]
                r, q = 0, pat
                while q:
                    if q & 1:
                        a = pk["vals"][v]
                        v += 1
                        for j in range(N):
                            part[w, r, j] = _fma(a, b[j], part[w, r, j])
                    q >>= 1
                    r += 1
                c += 1
        for r in range(M_r):
            row = p * M_r + r
            if row >= M:
                continue
            for j in range(N):
                s = part[0, r, j]
                for w in range(1, nsplit):
                    s = _add(s, part[w, r, j])
                C[row, j] = s
    return C


@pytest.mark.parametrize("nsplit", [1, 2, 4, 8])
def test_oracle_exec_order(oracle, nsplit):
    so, cb = oracle
    rng = np.random.default_rng(21)
    M, K, N = 19, 40, 3
    mask = rng.random((M, K)) < 0.45
    # mixed magnitudes so that summation order changes the rounded result
    A = (mask * rng.standard_normal((M, K)) * 10.0 ** rng.integers(-3, 4, (M, K))).astype(np.float32)
    B = (rng.uniform(-1, 1, (K, N)) * 10.0 ** rng.integers(-3, 4, (K, N))).astype(np.float32)
    table = so.mapping_table(8, so.orig8_strategy().map_pattern)
    pk = cb.pack_dense(A, 8, table)
    got = cb.exec_packed(pk, B, nsplit=nsplit)
    want = _exec_exact(pk, B, nsplit)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_split_changes_rounding_but_not_value(oracle):
    """The split order is a different rounding of the same sum: within the
    FP32 bound of the FP64 product, and (on these inputs) not bit-identical to
    the single chain - so the parity tests do distinguish the orders."""
    so, cb = oracle
    rng = np.random.default_rng(5)
    A = ((rng.random((64, 512)) < 0.5) * rng.standard_normal((64, 512))).astype(np.float32)
    B = rng.uniform(-1, 1, (512, 32)).astype(np.float32)
    pk = cb.pack_dense(A, 8, so.mapping_table(8, so.identity_strategy(8).map_pattern))
    c1 = cb.exec_packed(pk, B, nsplit=1)
    c8 = cb.exec_packed(pk, B, nsplit=8)
    bound = so.fp64_error_bound(A, B)
    C64 = so.dense_f64(A, B)
    assert np.all(np.abs(c1 - C64) <= bound) and np.all(np.abs(c8 - C64) <= bound)
    assert not np.array_equal(c1, c8)
