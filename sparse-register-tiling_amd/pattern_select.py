"""Pattern-set selection by set cover: which kernel patterns an executor is
generated for, and which kernel pattern each column pattern is padded to.

The reference solves this as an integer program with MOSEK
(sbench/ilp_pad/nano_solver.py).  Its model, restated:

  * Q = every column pattern q in [1, 2^M_r); f(q) = its expected count in a
    K_c = 128 column tile: f(q) = P(q) * K_c, with P(q) the measured
    frequency of q among a matrix's panel columns (d_fQ, :88-90,
    extract_distribution in sbench/SOP/distribution.py:43-52) or the
    Bernoulli(d) model (1 - d)^(M_r - |q|) * d^|q| (p_fQ, :81-85).
  * A kernel group g is a set of column patterns executed by one kernel
    pattern k(g) = OR of its members (every member padded up to k(g)); its
    cost is  nCheck + F(g) * (nBase + nCnz * |k(g)|)  with F(g) the sum of
    f over its members and nCheck = 2.0, nBase = 1.15, nCnz = 0.2
    (subset_cost / ncost / fS, :66-116).
  * Choose a partition of Q into at most max_groups groups (every pattern
    covered exactly once, :247-253; the MERGE_ONLY universe, :187-190,
    makes each subset a union of whole patterns) minimising the total cost.

This module solves the same model without an ILP solver: greedy
agglomerative merging (repeatedly merge the pair of groups whose union
lowers the total cost most, and keep merging while more than max_groups
remain), then single-pattern moves between groups while any lowers the
cost.  `exact_select` enumerates every partition for small pattern counts
and is the check the greedy result is measured against.
"""
from __future__ import annotations

from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from .mapping import Mapping, mapping_table, popcount

N_CHECK = 2.0
N_BASE = 1.15
N_CNZ = 0.2
K_C = 128


def column_distribution(A: np.ndarray, M_r: int) -> np.ndarray:
    """Frequency of every column pattern (index = code, bit r = row r of the
    panel) over all (panel, column) pairs of A (compute_histogram,
    sbench/SOP/distribution.py:18-27).  Rows past the last full panel are
    ignored, as the reference's loop over matrix[i:i + M_r] with a short
    last slice would give them codes of fewer bits."""
    A = np.asarray(A)
    M, K = A.shape
    P = M // M_r
    nz = (A[:P * M_r] != 0).reshape(P, M_r, K)
    codes = np.zeros((P, K), dtype=np.int64)
    for r in range(M_r):
        codes |= nz[:, r, :].astype(np.int64) << r
    hist = np.bincount(codes.ravel(), minlength=1 << M_r).astype(np.float64)
    total = max(1, P * K)
    return hist / total


def random_distribution(M_r: int, density: float) -> np.ndarray:
    """P(q) of independent Bernoulli(density) entries (p_fQ's model)."""
    q = np.arange(1 << M_r)
    nnz = np.array([popcount(int(x)) for x in q])
    return (1 - density) ** (M_r - nnz) * density ** nnz


def group_cost(F: float, k: int) -> float:
    return N_CHECK + F * (N_BASE + N_CNZ * popcount(k))


def partition_cost(groups: Iterable[Sequence[int]], f: np.ndarray) -> float:
    total = 0.0
    for g in groups:
        k = 0
        for q in g:
            k |= q
        total += group_cost(float(sum(f[q] for q in g)), k)
    return total


def _greedy(f: np.ndarray, patterns: List[int], max_groups: int) -> List[List[int]]:
    groups = [[q] for q in patterns]
    F = np.array([f[q] for q in patterns], dtype=np.float64)
    K = np.array(patterns, dtype=np.int64)
    bits = np.array([popcount(int(x)) for x in range(int(K.max()) * 2 + 2 if len(K) else 2)])
    while len(groups) > 1:
        # delta of merging i and j: -nCheck + nCnz * ((Fi+Fj)|ki|kj| - Fi|ki| - Fj|kj|)
        Kor = K[:, None] | K[None, :]
        cost_ij = N_CHECK + (F[:, None] + F[None, :]) * (N_BASE + N_CNZ * bits[Kor])
        own = N_CHECK + F * (N_BASE + N_CNZ * bits[K])
        delta = cost_ij - own[:, None] - own[None, :]
        np.fill_diagonal(delta, np.inf)
        i, j = np.unravel_index(int(np.argmin(delta)), delta.shape)
        if delta[i, j] >= 0 and len(groups) <= max_groups:
            break
        i, j = min(i, j), max(i, j)
        groups[i] = groups[i] + groups[j]
        F[i] += F[j]
        K[i] |= K[j]
        del groups[j]
        F = np.delete(F, j)
        K = np.delete(K, j)
    return groups


def _improve(groups: List[List[int]], f: np.ndarray, max_groups: int) -> List[List[int]]:
    """Local search: move one pattern to another group or to a new group
    (while the limit allows), or swap two patterns of different groups,
    while any such step lowers the cost."""
    def gcost(g):
        return group_cost(float(sum(f[q] for q in g)), _or(g)) if g else 0.0

    groups = [list(g) for g in groups]
    improved = True
    while improved:
        improved = False
        costs = [gcost(g) for g in groups]
        best = (-1e-12, None)
        for a in range(len(groups)):
            for q in groups[a]:
                src = [x for x in groups[a] if x != q]
                src_cost = gcost(src)
                if len(groups) < max_groups and src:  # q alone in a new group
                    d = src_cost + gcost([q]) - costs[a]
                    if d < best[0]:
                        best = (d, ("new", a, q))
                for b in range(len(groups)):
                    if b == a:
                        continue
                    d = src_cost + gcost(groups[b] + [q]) - costs[a] - costs[b]
                    if d < best[0]:
                        best = (d, ("move", a, q, b))
                    for r in groups[b]:
                        if r < q and b < a:
                            continue  # each unordered swap once
                        dst = [x for x in groups[b] if x != r]
                        d = gcost(src + [r]) + gcost(dst + [q]) - costs[a] - costs[b]
                        if d < best[0]:
                            best = (d, ("swap", a, q, b, r))
        step = best[1]
        if step is not None:
            improved = True
            if step[0] == "new":
                _, a, q = step
                groups[a].remove(q)
                groups.append([q])
            elif step[0] == "move":
                _, a, q, b = step
                groups[a].remove(q)
                groups[b].append(q)
            else:
                _, a, q, b, r = step
                groups[a].remove(q)
                groups[b].remove(r)
                groups[a].append(r)
                groups[b].append(q)
            groups = [g for g in groups if g]
    return [sorted(g) for g in groups]


def select_patterns(dist: np.ndarray, M_r: int, max_groups: int = 16,
                    patterns: Optional[Sequence[int]] = None) -> Tuple[List[List[int]], float]:
    """Greedy solution of the set-cover model: (groups, total cost).
    `dist` is P(q) indexed by code (dist[0] ignored); `patterns` restricts Q
    (default: every nonzero code, as generate_Q_for)."""
    if max_groups < 1:
        raise ValueError("max_groups must be >= 1")
    f = np.asarray(dist, dtype=np.float64) * K_C
    if f.size != 1 << M_r:
        raise ValueError("dist must have 2^M_r entries")
    Q = list(patterns) if patterns is not None else list(range(1, 1 << M_r))
    # two starts: agglomerative merging, and one group per popcount layer
    # folded to the limit; the local search polishes both, the cheaper wins
    layers: Dict[int, List[int]] = {}
    for q in Q:
        layers.setdefault(popcount(q), []).append(q)
    start2 = [g for _, g in sorted(layers.items())]
    while len(start2) > max_groups:
        start2[-2:] = [start2[-2] + start2[-1]]
    cands = [_improve(_greedy(f, Q, max_groups), f, max_groups), _improve(start2, f, max_groups)]
    groups = min(cands, key=lambda g: partition_cost(g, f))
    groups.sort(key=lambda g: (popcount(_or(g)), _or(g)))
    return groups, partition_cost(groups, f)


def _or(g: Iterable[int]) -> int:
    k = 0
    for q in g:
        k |= q
    return k


def _partitions(items: List[int]):
    if not items:
        yield []
        return
    first, rest = items[0], items[1:]
    for part in _partitions(rest):
        for i in range(len(part)):
            yield part[:i] + [[first] + part[i]] + part[i + 1:]
        yield [[first]] + part


def exact_select(dist: np.ndarray, M_r: int, max_groups: int = 16,
                 patterns: Optional[Sequence[int]] = None) -> Tuple[List[List[int]], float]:
    """Optimal partition by enumeration (Bell-number many; small Q only)."""
    f = np.asarray(dist, dtype=np.float64) * K_C
    Q = list(patterns) if patterns is not None else list(range(1, 1 << M_r))
    if len(Q) > 10:
        raise ValueError("exact enumeration is for at most 10 patterns")
    best, best_cost = None, np.inf
    for part in _partitions(Q):
        if len(part) > max_groups:
            continue
        c = partition_cost(part, f)
        if c < best_cost - 1e-12:
            best, best_cost = part, c
    return [sorted(g) for g in best], best_cost


def groups_to_mapping(groups: Sequence[Sequence[int]], M_r: int) -> Mapping:
    """Mapping q -> [k(group of q)] (the ILP's mapping[q] = {k(s)},
    nano_solver.py:268-271); codes outside every group map to the full
    pattern."""
    owner: Dict[int, int] = {}
    for g in groups:
        k = _or(g)
        for q in g:
            owner[q] = k
    full = (1 << M_r) - 1
    # kernel list in group order, as the ILP's output file lists its kernels
    # one per line (read by sbench/SOP/sop_bench_ilp.py:98-111)
    kernels = [_or(g) for g in groups]
    if any(q not in owner for q in range(1, full + 1)) and full not in kernels:
        kernels.append(full)
    return Mapping(M_r, mapping_table(M_r, lambda q: [owner.get(q, full)]), kernel_list=kernels)


def specialized_mapping(A: np.ndarray, M_r: int, max_groups: int = 16) -> Tuple[Mapping, float]:
    """Pattern set and mapping specialised to one matrix's column patterns."""
    groups, cost = select_patterns(column_distribution(A, M_r), M_r, max_groups)
    return groups_to_mapping(groups, M_r), cost
