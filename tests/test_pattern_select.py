"""Set-cover pattern selection (the reference's ILP padding model) on the CPU.

The greedy + local-search solver is held against exhaustive enumeration of
every partition on small pattern sets; the cost model against a hand
evaluation of the reference's formula; the column histogram against the
oracle's per-column pattern codes."""
import itertools

import numpy as np
import pytest


def _ps(srt):
    return srt.pattern_select


def test_cost_model_by_hand(srt):
    ps = _ps(srt)
    f = np.zeros(8)
    f[1], f[2], f[3], f[5] = 10.0, 4.0, 2.0, 1.0
    # group {1, 2}: kernel pattern 3 (2 bits): 2.0 + 14 * (1.15 + 0.2 * 2)
    assert ps.group_cost(14.0, 3) == pytest.approx(2.0 + 14 * (1.15 + 0.4))
    part = [[1, 2], [3], [5]]
    want = (2.0 + 14 * 1.55) + (2.0 + 2 * 1.55) + (2.0 + 1 * 1.55)
    assert ps.partition_cost(part, f) == pytest.approx(want)
    # merging 5 into {1, 2} widens the kernel pattern to 7 (3 bits)
    assert ps.partition_cost([[1, 2, 5], [3]], f) == pytest.approx((2.0 + 15 * 1.75) + (2.0 + 2 * 1.55))


def test_greedy_close_to_exact(srt):
    ps = _ps(srt)
    rng = np.random.default_rng(0)
    ratios = []
    for trial in range(40):
        if trial % 2:
            d = ps.random_distribution(3, rng.uniform(0.05, 0.6))
        else:
            d = rng.dirichlet(np.ones(8) * rng.uniform(0.2, 2))
        for mg in (1, 2, 3, 7):
            g, c = ps.select_patterns(d, 3, mg)
            e, ce = ps.exact_select(d, 3, mg)
            assert len(g) <= mg and sorted(itertools.chain(*g)) == list(range(1, 8))
            assert c >= ce - 1e-9  # enumeration is optimal
            ratios.append(c / ce)
    ratios = np.array(ratios)
    assert ratios.max() <= 1.005
    assert np.mean(ratios <= 1 + 1e-9) >= 0.9


def test_exact_enumerates_all_partitions(srt):
    ps = _ps(srt)
    parts = list(ps._partitions([1, 2, 3, 4]))
    assert len(parts) == 15  # Bell(4)
    assert len({tuple(sorted(tuple(sorted(g)) for g in p)) for p in parts}) == 15


def test_column_distribution_matches_oracle_codes(srt, oracle):
    so, _ = oracle
    ps = _ps(srt)
    rng = np.random.default_rng(3)
    A = (rng.random((13, 9)) < 0.3) * rng.standard_normal((13, 9))
    d = ps.column_distribution(A, 4)
    codes = so.panel_codes(A[:12], 4)  # three full panels
    want = np.bincount(np.asarray(codes).ravel(), minlength=16) / codes.size
    assert np.allclose(d, want)
    assert d.sum() == pytest.approx(1.0)


def test_mapping_from_groups(srt):
    ps = _ps(srt)
    A = srt.workloads.magnitude_pruned(256, 128, 0.8, seed=1)
    m, cost = ps.specialized_mapping(A, 4, max_groups=6)
    assert m.M_r == 4 and 1 <= len(m.patterns) <= 6
    for q in range(1, 16):
        (k,) = m(q)
        assert k & q == q  # padded up, never dropped
    # the dominant pattern is never padded: its group's kernel pattern is itself
    d = np.zeros(16)
    d[5] = 0.9
    d[1:5] += 0.025
    groups, _ = ps.select_patterns(d, 4, 16)
    (g5,) = [g for g in groups if 5 in g]
    assert np.bitwise_or.reduce(g5) == 5
