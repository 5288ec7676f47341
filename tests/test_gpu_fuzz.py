"""Seeded random configurations of the register-tiled executor on the GPU:
matrix shape, density and row skew, mapping, lane width, unroll, workgroup
size, split, chunk length, K tile, workgroup layout, XCD remap and B path,
each result bit-identical to the oracle's chain over the same packing with
the plan's chunk splits, and within the FP64 bound."""
import sys

import numpy as np
import pytest

from test_gpu_parity import _oracle_table, _plan, _run, _bitexact

pytestmark = pytest.mark.gpu

MAPPINGS = [(4, "orig"), (4, "identity"), (8, "orig"), (8, "alt"), (2, "identity"), (6, "identity")]


def _case(i):
    rng = np.random.default_rng(1000 + i)
    M = int(rng.integers(1, 700))
    K = int(rng.integers(1, 800))
    N = int(rng.integers(1, 320))
    density = float(rng.choice([0.01, 0.05, 0.1, 0.3, 0.6]))
    skew = float(rng.choice([0.0, 1.0]))
    scale = rng.lognormal(0.0, skew, (M, 1)) if skew else np.ones((M, 1))
    p = np.clip(density * scale / scale.mean(), 0.0, 1.0)
    A = np.where(rng.random((M, K)) < p, rng.standard_normal((M, K)), 0.0).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    W = int(rng.choice([1, 2, 4, 8, 16]))
    cfg = dict(vec_width=int(rng.choice([1, 2, 4])), unroll=int(rng.choice([1, 2, 4, 8])), waves_per_block=W,
               nsplit=int(rng.integers(1, W + 1)), chunk_cols=int(rng.choice([0, 4, 16, 64])),
               K_c=int(rng.choice([0, 0, 32, 128, 256])), layout=int(rng.choice([1, 2])),
               xcd_remap=int(rng.choice([0, 1])), b_staging=int(rng.choice([1, 2, 2])))
    return A, B, MAPPINGS[i % len(MAPPINGS)], cfg


@pytest.mark.parametrize("i", range(60))
def test_random_configuration(srt, oracle, i):
    import torch
    so, cb = oracle
    A, B, (M_r, name), cfg = _case(i)
    print("fuzz", i, A.shape, B.shape[1], M_r, name, cfg, flush=True)
    sys.stdout.flush()
    plan = _plan(srt, torch, A, B.shape[1], M_r, name, **cfg)
    real = plan.get_config()
    table = _oracle_table(so, M_r, name)
    K = A.shape[1]
    pk = cb.pack_dense_ktiled(A, M_r, table, real.K_c) if 0 < real.K_c < K else cb.pack_dense(A, M_r, table)
    C = _run(srt, torch, plan, B)
    _bitexact(C, cb.exec_packed(pk, B, nsplit=plan.split_counts()))
    assert np.all(np.abs(C - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B))
    plan.close()


def _hybrid_case(i):
    """Panels of mixed density (each 16-row panel dense with probability
    0.4), ragged M, K and N, one of the mappings, a random threshold."""
    rng = np.random.default_rng(5000 + i)
    M = int(rng.integers(1, 600))
    K = int(rng.integers(1, 1500))
    N = int(rng.integers(1, 300))
    dens = np.where(rng.random(-(-M // 16)) < 0.4, rng.uniform(0.3, 0.9), rng.uniform(0.0, 0.15))
    p = np.repeat(dens, 16)[:M, None]
    A = np.where(rng.random((M, K)) < p, rng.standard_normal((M, K)), 0.0).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    tau = float(rng.choice([0.05, 0.2, 0.3, 0.5]))
    cfg = dict(vec_width=int(rng.choice([1, 2, 4])), waves_per_block=int(rng.choice([4, 8, 16])),
               layout=int(rng.choice([1, 2])), K_c=int(rng.choice([0, 0, 128, 256])))
    cfg["nsplit"] = int(rng.integers(1, cfg["waves_per_block"] + 1))
    return A, B, MAPPINGS[i % len(MAPPINGS)], tau, cfg


@pytest.mark.parametrize("i", range(32))
def test_random_hybrid(srt, oracle, i):
    """Hybrid plans at random: rows of register-tiled panels bit-identical
    to the oracle's chain over the matrix with the dense panels emptied, the
    dense panels' rows within the FP64 bound, every row written."""
    import torch
    so, cb = oracle
    A, B, (M_r, name), tau, cfg = _hybrid_case(i)
    M, K = A.shape
    N = B.shape[1]
    print("hybrid fuzz", i, A.shape, N, M_r, name, tau, cfg, flush=True)
    rp, ci, v = srt.workloads.to_csr(A)
    plan = srt.MatMulSpecialized(M, K, N, torch.from_numpy(v), torch.from_numpy(rp), torch.from_numpy(ci),
                                 mapping_id=srt.named_mapping(M_r, name), tile_config=srt.TileConfig(**cfg),
                                 dense_threshold=tau)
    rows = np.zeros(M, dtype=bool)
    for r0 in plan.dense_panels()["rows"]:
        nz = np.count_nonzero(A[r0:r0 + 16])
        assert nz and nz / (min(16, M - r0) * K) >= tau
        rows[r0:r0 + 16] = True
    C = _run(srt, torch, plan, B)
    assert not np.isnan(C).any()
    As = A.copy()
    As[rows] = 0
    table = _oracle_table(so, M_r, name)
    kc = plan.get_config().K_c
    pk = cb.pack_dense_ktiled(As, M_r, table, kc) if 0 < kc < K else cb.pack_dense(As, M_r, table)
    ref = cb.exec_packed(pk, B, nsplit=plan.split_counts())
    _bitexact(np.ascontiguousarray(C[~rows]), np.ascontiguousarray(ref[~rows]))
    assert np.all(np.abs(C - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B))
    plan.close()
