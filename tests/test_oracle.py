"""Pins the oracle (oracle/) before anything is checked against it.

What each test pins:
  * mapping generator and id rule - the reference's committed mapping files
    (tests/golden/mappings, copied from sbench/SOP/*mappings) regenerate byte
    for byte, and every file's name is the md5 id of its content;
  * pattern codes - closed form vs explicit bit loop vs brute force;
  * stable sort by id - a brute-force bucket walk;
  * packing - two independent restatements (Python literal loops, C counting
    sort) agree exactly, and the reference's invariant num_nnz == padding +
    nnz(panel) holds;
  * executors - against an FP64 dense product within the FP32 summation
    bound, and the reference's tolerance rule on its positive-input protocol;
  * planted mistakes in the oracle are caught by those same checks.
"""
import glob
import hashlib
import itertools
import os

import numpy as np
import pytest

from conftest import GOLDEN_MAPPINGS


def _golden(mid):
    return open(os.path.join(GOLDEN_MAPPINGS, "mapping_%s.txt" % mid)).read()


@pytest.mark.parametrize("mid,M_r,strategy", [
    ("61fee", 4, "identity4"), ("161fb", 8, "identity8"), ("4af51", 2, "identity2"),
    ("001e3", 6, "identity6"), ("400fa", 8, "orig8"), ("747f9", 8, "alt8")])
def test_mapping_files_regenerate(oracle, mid, M_r, strategy):
    so, _ = oracle
    strat = {"identity4": so.identity_strategy(4), "identity8": so.identity_strategy(8),
             "identity2": so.identity_strategy(2), "identity6": so.identity_strategy(6),
             "orig8": so.orig8_strategy(), "alt8": so.alt8_strategy()}[strategy]
    table = so.mapping_table(M_r, strat.map_pattern)
    assert so.mapping_id(table) == mid
    assert so.mapping_file_text(M_r, table) == _golden(mid)


def test_da01e_id(oracle):
    """M_r=4 "orig" (tools/paper/configs.py:86) is not committed; its id is."""
    so, _ = oracle
    t = so.mapping_table(4, so.orig4_strategy().map_pattern)
    assert so.mapping_id(t) == "da01e"
    # The first-minimum tie rule matters: another merge order gives another id.
    t2 = so.mapping_table(4, so.orig4_strategy((7, 11, 13, 14)).map_pattern)
    assert so.mapping_id(t2) != "da01e"


# Kernel id tables written out (tools/codegen/gen_vector_specialized.py:113
# stable popcount sort of the strategy's all_patterns(), merge then split:
# sbench/SOP/sop_hand_configs.py:16-17,99-103,160-169,209-219).
ID_KATS = {(8, "orig"): [1, 2, 4, 8, 16, 32, 64, 128, 85, 170, 195, 60, 15, 240, 252, 243, 207, 63, 255],
           (8, "alt"): [1, 2, 4, 8, 16, 32, 64, 128, 85, 170, 195, 60, 15, 240, 252, 243, 207, 63, 255],
           (4, "orig"): [1, 2, 4, 8, 14, 13, 11, 7, 15],
           (4, "identity"): [1, 2, 4, 8, 3, 5, 6, 9, 10, 12, 7, 11, 13, 14, 15],
           (2, "identity"): [1, 2, 3]}


@pytest.mark.parametrize("key", sorted(ID_KATS))
def test_kernel_id_kats(oracle, key):
    so, _ = oracle
    table = so.named_table(*key)
    assert so.kernel_pattern_table(table) == ID_KATS[key]
    # the table itself is the committed / published mapping
    mid = {(8, "orig"): "400fa", (8, "alt"): "747f9", (4, "identity"): "61fee", (2, "identity"): "4af51"}.get(key)
    if mid:
        assert so.mapping_file_text(key[0], table) == _golden(mid)
    else:
        assert so.mapping_id(table) == "da01e"


def test_file_kernel_ids(oracle):
    """A mapping read from a file: list(set) of its patterns in line order
    (load_mapping), then the stable popcount sort."""
    so, _ = oracle
    M_r, table = so.read_mapping(os.path.join(GOLDEN_MAPPINGS, "mapping_400fa.txt"))
    assert so.kernel_pattern_table(table) == [128, 1, 2, 4, 8, 16, 32, 64, 15, 170, 60, 195, 85, 240,
                                              63, 207, 243, 252, 255]
    # identity files: the set of 1..2^M_r-1 iterates ascending, so ids are (popcount, value)
    M_r, table = so.read_mapping(os.path.join(GOLDEN_MAPPINGS, "mapping_161fb.txt"))
    assert so.kernel_pattern_table(table) == sorted(range(1, 256), key=lambda p: (bin(p).count("1"), p))
    # a pattern the table uses must be listed
    with pytest.raises(ValueError):
        so.kernel_pattern_table(table, kernel_list=[1, 2, 3])


def test_every_golden_file_id(oracle):
    so, _ = oracle
    files = sorted(glob.glob(os.path.join(GOLDEN_MAPPINGS, "mapping_*.txt")))
    assert len(files) == 25
    for f in files:
        mid = os.path.basename(f)[len("mapping_"):-4]
        M_r, table = so.read_mapping(f)
        as_lists = so.mapping_id(table)
        # sbench/ilp_pad/nano_solver_best.py:27 hashes set-valued dicts.
        as_sets = hashlib.md5(str({k: set(v) for k, v in table.items()}).encode()).hexdigest()[-5:]
        assert mid in (as_lists, as_sets), f
        assert so.mapping_file_text(M_r, table) == open(f).read()
        # every code covered exactly once
        for code, pats in table.items():
            u = 0
            for p in pats:
                assert p & u == 0
                u |= p
            assert u & code == code


def test_pattern_code(oracle):
    so, _ = oracle
    rng = np.random.default_rng(3)
    for M_r in (2, 4, 8, 16):
        A = (rng.random((3 * M_r + 1, 17)) < 0.4) * rng.standard_normal((3 * M_r + 1, 17))
        codes = so.panel_codes(A, M_r)
        for p in range(codes.shape[0]):
            for c in range(A.shape[1]):
                rows = A[p * M_r:(p + 1) * M_r, c]
                assert codes[p, c] == so.pattern_code(rows)
                assert codes[p, c] == sum(1 << r for r in range(len(rows)) if rows[r] != 0)
    # brute force over every 4-bit column
    for bits in itertools.product([0.0, -2.5], repeat=4):
        assert so.pattern_code(bits) == sum(1 << i for i, b in enumerate(bits) if b)


def test_sort_and_encode_stable(oracle):
    so, _ = oracle
    rng = np.random.default_rng(5)
    patterns = [1, 2, 4, 8, 15, 7, 3]
    pats = sorted(patterns, key=lambda p: (bin(p).count("1"), p))
    codes = rng.choice(pats + [0, 0], size=200).tolist()
    enc, order = so.sort_and_encode_patterns(codes, pats)
    ids = {p: i for i, p in enumerate(pats)}
    expect = []
    for i in range(len(pats)):
        expect += [j for j, c in enumerate(codes) if c and ids[c] == i]
    expect += [j for j, c in enumerate(codes) if c == 0]
    assert order == expect
    assert enc == [ids[codes[j]] if codes[j] else so.ZERO_PATTERN_ID for j in order]


MAPPING_CASES = [("61fee", 4), ("da01e", 4), ("400fa", 8), ("747f9", 8), ("161fb", 8), ("4af51", 2),
                 ("001e3", 6), ("b7844", 8), ("9b680", 4)]


def _table(so, mid, M_r):
    if mid == "da01e":
        return so.mapping_table(4, so.orig4_strategy().map_pattern)
    return so.read_mapping(os.path.join(GOLDEN_MAPPINGS, "mapping_%s.txt" % mid))[1]


@pytest.mark.parametrize("mid,M_r", MAPPING_CASES)
def test_pack_python_equals_c(oracle, mid, M_r):
    so, cb = oracle
    rng = np.random.default_rng(11)
    table = _table(so, mid, M_r)
    pats = so.kernel_pattern_table(table)
    for (M, K, d) in [(37, 50, 0.3), (64, 96, 0.1), (19, 8, 0.9)]:
        A = ((rng.random((M, K)) < d) * rng.standard_normal((M, K))).astype(np.float32)
        panels = so.make_sop_tile(A, M_r, so.table_mapping(table), pats)
        pk = cb.pack_dense(A, M_r, table)
        assert pk["patterns"] == pats
        for p, pd in enumerate(panels):
            assert np.array_equal(pd["nkern_counts"], pk["counts"][p])
            assert np.array_equal(pd["col_indices"], pk["cols"][pk["col_ptr"][p]:pk["col_ptr"][p + 1]])
            assert np.array_equal(pd["values"], pk["vals"][pk["val_ptr"][p]:pk["val_ptr"][p + 1]])
            assert pd["num_nnz"] == pd["padding"] + np.count_nonzero(A[p * M_r:(p + 1) * M_r])
        assert pk["padding"] == sum(pd["padding"] for pd in panels)
        if "identity" in mid or mid == "61fee":
            assert pk["padding"] == 0


def test_executor_matches_fp64(oracle):
    so, cb = oracle
    rng = np.random.default_rng(7)
    for mid, M_r in MAPPING_CASES[:5]:
        table = _table(so, mid, M_r)
        A = ((rng.random((120, 300)) < 0.15) * rng.standard_normal((120, 300))).astype(np.float32)
        B = rng.uniform(-1, 1, (300, 72)).astype(np.float32)
        pk = cb.pack_dense(A, M_r, table)
        C = cb.exec_packed(pk, B)
        C64 = so.dense_f64(A, B)
        assert np.all(np.abs(C - C64) <= so.fp64_error_bound(A, B))
        # float64 walk of the packed data equals the dense FP64 product
        pats = so.kernel_pattern_table(table)
        panels = so.make_sop_tile(A, M_r, so.table_mapping(table), pats)
        assert np.allclose(so.executor_f64(panels, pats, M_r, A.shape[0], B), C64, rtol=0, atol=1e-12)


def test_reference_tolerance_protocol(oracle):
    """Positive U[0,1] inputs as SPMM_demo.cpp:446-450 sets them, checked with
    is_within_tol (SPMM_demo.cpp:105-112) against the FP64 product."""
    so, cb = oracle
    rng = np.random.default_rng(9)
    mask = rng.random((256, 128)) < 0.2
    A = (mask * rng.random((256, 128))).astype(np.float32)
    B = rng.random((128, 64)).astype(np.float32)
    pk = cb.pack_dense(A, 8, _table(so, "400fa", 8))
    C = cb.exec_packed(pk, B)
    assert so.is_within_tol(so.dense_f64(A, B), C).all()


def test_load_c_accumulates(oracle):
    so, cb = oracle
    rng = np.random.default_rng(4)
    A = ((rng.random((40, 30)) < 0.3) * rng.standard_normal((40, 30))).astype(np.float32)
    B = rng.uniform(-1, 1, (30, 16)).astype(np.float32)
    pk = cb.pack_dense(A, 4, _table(so, "61fee", 4))
    C0 = rng.uniform(-1, 1, (40, 16)).astype(np.float32)
    C = cb.exec_packed(pk, B, C=C0.copy(), load_c=True)
    assert np.all(np.abs(C - (C0 + so.dense_f64(A, B))) <= so.fp64_error_bound(A, B) + 1e-6)


def test_csr_oracle(oracle):
    so, cb = oracle
    rng = np.random.default_rng(8)
    A = ((rng.random((77, 90)) < 0.2) * rng.standard_normal((77, 90))).astype(np.float32)
    B = rng.uniform(-1, 1, (90, 33)).astype(np.float32)
    nz = A != 0
    rp = np.concatenate([[0], np.cumsum(nz.sum(1))])
    r, c = np.nonzero(nz)
    C = cb.csr_exec(rp, c, A[r, c], 77, B)
    assert np.all(np.abs(C - so.dense_f64(A, B)) <= so.fp64_error_bound(A, B))


def test_full_pattern_chain_is_the_csr_chain(oracle):
    """Under the M_r = 4 "full" mapping every packed column carries all four
    rows, zeros included.  fmaf(0, b, acc) = acc for finite b when acc is not
    -0 (a chain that starts at +0 never reaches -0 in round-to-nearest), so
    per output element the packed chain is the CSR chain over the stored
    entries in column order: the two executors agree bit for bit."""
    so, cb = oracle
    rng = np.random.default_rng(21)
    table = so.named_table(4, "full")
    assert so.kernel_pattern_table(table) == [15] and all(v == [15] for v in table.values())
    for (M, K, d) in [(61, 140, 0.1), (128, 64, 0.5), (9, 300, 0.02)]:
        A = ((rng.random((M, K)) < d) * rng.standard_normal((M, K))).astype(np.float32)
        A[A.shape[0] // 2] = 0.0  # an empty row
        B = rng.uniform(-1, 1, (K, 40)).astype(np.float32)
        pk = cb.pack_dense(A, 4, table)
        assert pk["vals"].size == 4 * pk["cols"].size   # every packed column full
        nz = A != 0
        rp = np.concatenate([[0], np.cumsum(nz.sum(1))])
        r, c = np.nonzero(nz)
        C_csr = cb.csr_exec(rp, c, A[r, c], M, B)
        assert np.array_equal(cb.exec_packed(pk, B).view(np.uint32), C_csr.view(np.uint32))


# ---------------------------------------------------------- planted mistakes
def test_planted_mistakes_are_caught(oracle):
    """Mutations of the oracle that a plausible bug would make must fail the
    checks above (dropped term, reversed row order, transposed operand)."""
    so, cb = oracle
    rng = np.random.default_rng(12)
    A = ((rng.random((64, 80)) < 0.25) * rng.standard_normal((64, 80))).astype(np.float32)
    B = rng.uniform(-1, 1, (80, 40)).astype(np.float32)
    table = _table(so, "400fa", 8)
    pk = cb.pack_dense(A, 8, table)
    bound = so.fp64_error_bound(A, B)
    C64 = so.dense_f64(A, B)
    assert np.all(np.abs(cb.exec_packed(pk, B) - C64) <= bound)

    # 1) values of a split column within a pattern in reversed row order
    bad = dict(pk)
    vals = pk["vals"].copy()
    pats = pk["patterns"]
    v = 0
    for p in range(pk["P"]):
        for kid, cnt in enumerate(pk["counts"][p]):
            pc = bin(pats[kid]).count("1")
            for _ in range(cnt):
                vals[v:v + pc] = vals[v:v + pc][::-1].copy()
                v += pc
    bad["vals"] = vals
    assert not np.all(np.abs(cb.exec_packed(bad, B) - C64) <= bound)

    # 2) a dropped packed column
    bad = dict(pk)
    cols = pk["cols"].copy()
    cols[len(cols) // 2] = (cols[len(cols) // 2] + 1) % A.shape[1]
    bad["cols"] = cols
    assert not np.all(np.abs(cb.exec_packed(bad, B) - C64) <= bound)

    # 3) a mapping entry that does not cover its code (value silently lost)
    t2 = dict(table)
    t2[3] = [1]
    A3 = np.zeros_like(A)
    A3[0, 0] = 1.0
    A3[1, 0] = 2.0
    pk3 = cb.pack_dense(A3, 8, t2)
    assert not np.all(np.abs(cb.exec_packed(pk3, B) - so.dense_f64(A3, B)) <= so.fp64_error_bound(A3, B))
    assert pk3["val_ptr"][-1] - pk3["padding"] != np.count_nonzero(A3)  # invariant catches it

    # 4) transposed B (square case)
    Bs = rng.uniform(-1, 1, (80, 80)).astype(np.float32)
    assert not np.all(np.abs(cb.exec_packed(pk, Bs.T.copy()) - so.dense_f64(A, Bs))
                      <= so.fp64_error_bound(A, Bs))
