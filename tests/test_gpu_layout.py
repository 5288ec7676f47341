"""GPU parity of the balanced workgroup layout (TileConfig.layout = 2: panels
binned heaviest-first into whole rounds of the chip's CUs, wave slots given
to the longest chunks, so slots no longer hold the panels in order): C is
bit-identical to the oracle's chain over the same packed data with the
plan's per-panel chunk splits, on uneven rows, empty panels, row and column
tails, K tiles, every lane width and workgroup size, and after the schedule
is rebuilt for another B width."""
import sys

import numpy as np
import pytest

from test_gpu_parity import _oracle_table, _plan, _run, _bitexact

pytestmark = pytest.mark.gpu


def _uneven(M, K, density, seed):
    """Rows of very different lengths and a block of empty rows (whole empty
    panels), values N(0, 1)."""
    rng = np.random.default_rng(seed)
    scale = rng.lognormal(0.0, 1.0, (M, 1))
    p = np.clip(density * scale / scale.mean(), 0.0, 1.0)
    A = np.where(rng.random((M, K)) < p, rng.standard_normal((M, K)), 0.0).astype(np.float32)
    A[M // 3:M // 3 + 24] = 0.0
    return A


def _check(srt, oracle, A, B, M_r, name, layout=2, **cfg):
    import torch
    so, cb = oracle
    plan = _plan(srt, torch, A, B.shape[1], M_r, name, layout=layout, xcd_remap=1, **cfg)
    real = plan.get_config()
    assert real.layout == (layout or 2)
    pk = cb.pack_dense_ktiled(A, M_r, _oracle_table(so, M_r, name), real.K_c) if real.K_c < A.shape[1] \
        else cb.pack_dense(A, M_r, _oracle_table(so, M_r, name))
    C = _run(srt, torch, plan, B)
    kp = plan.split_counts()
    _bitexact(C, cb.exec_packed(pk, B, nsplit=kp))
    assert np.all(np.abs(C - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B))
    return plan, kp


@pytest.mark.parametrize("W,S", [(16, 16), (16, 8), (8, 8), (4, 4), (8, 1)])
@pytest.mark.parametrize("V", [1, 2, 4])
@pytest.mark.parametrize("U", [2, 4])
def test_balanced_layout_uneven(srt, oracle, W, S, V, U):
    print("layout2 uneven", W, S, V, U, flush=True)
    sys.stdout.flush()
    A = _uneven(770, 300, 0.12, seed=5)
    B = srt.workloads.dense_b(300, 264, seed=6)
    plan, kp = _check(srt, oracle, A, B, 4, "orig", waves_per_block=W, nsplit=S, vec_width=V, unroll=U,
                      chunk_cols=16)
    assert kp.max() <= S
    plan.close()


@pytest.mark.parametrize("W,S", [(16, 16), (8, 8), (4, 2)])
@pytest.mark.parametrize("V", [1, 4])
def test_panel_order_layout_uneven(srt, oracle, W, S, V):
    """layout = 1: panels laid into workgroups in panel order."""
    print("layout1 uneven", W, S, V, flush=True)
    sys.stdout.flush()
    A = _uneven(770, 300, 0.12, seed=5)
    B = srt.workloads.dense_b(300, 264, seed=6)
    plan, kp = _check(srt, oracle, A, B, 4, "orig", layout=1, waves_per_block=W, nsplit=S, vec_width=V,
                      chunk_cols=16)
    assert kp.max() <= S
    plan.close()


def test_default_layout_is_balanced(srt, oracle):
    A = _uneven(300, 128, 0.1, seed=4)
    B = srt.workloads.dense_b(128, 96, seed=4)
    plan, _ = _check(srt, oracle, A, B, 4, "orig", layout=0)
    plan.close()


@pytest.mark.parametrize("M_r,name", [(4, "identity"), (8, "orig"), (2, "identity")])
@pytest.mark.parametrize("k_tile", [0, 128])
def test_balanced_layout_mappings_and_k_tiles(srt, oracle, M_r, name, k_tile):
    print("layout2", M_r, name, k_tile, flush=True)
    sys.stdout.flush()
    A = _uneven(1030, 520, 0.08, seed=7)
    B = srt.workloads.dense_b(520, 200, seed=8)
    plan, _ = _check(srt, oracle, A, B, M_r, name, K_c=k_tile, waves_per_block=16, nsplit=16)
    plan.close()


def test_balanced_layout_2048x512_workload(srt, oracle):
    """2048 x 512 (90 %) x 512 with the balanced layout on 16-wave workgroups."""
    A, B = srt.workloads.case_inputs(srt.workloads.PRUNED_2048_512_S90_N512, seed=0)
    plan, kp = _check(srt, oracle, A, B, 4, "orig", K_c=256, waves_per_block=16, nsplit=16, vec_width=4,
                      unroll=2, chunk_cols=24)
    assert kp.max() > 1
    plan.close()


def test_balanced_layout_rebuilt_for_another_width(srt, oracle):
    """allocate_executor with a wider B rebuilds the balanced schedule (its
    workgroup count depends on the N tiles); results stay exact at both."""
    import torch
    so, cb = oracle
    A = _uneven(600, 256, 0.1, seed=9)
    plan = _plan(srt, torch, A, 64, 4, "orig", layout=2, waves_per_block=8, vec_width=1)
    pk = cb.pack_dense(A, 4, _oracle_table(so, 4, "orig"))
    for n in (64, 448):
        plan.allocate_executor(n)
        B = srt.workloads.dense_b(256, n, seed=n)
        C = _run(srt, torch, plan, B)
        _bitexact(C, cb.exec_packed(pk, B, nsplit=plan.split_counts()))
    plan.close()


def test_balanced_layout_tiny(srt, oracle):
    """Fewer panels than one workgroup's slots, and a single row."""
    for M, K, N in ((9, 5, 3), (1, 7, 64), (40, 33, 130)):
        A = _uneven(M, K, 0.5, seed=M)
        B = srt.workloads.dense_b(K, N, seed=K)
        plan, _ = _check(srt, oracle, A, B, 4, "orig", waves_per_block=16, nsplit=16)
        plan.close()


def test_graph_replays_after_allocate_executor(srt, oracle):
    """A graph captured at one width keeps replaying exactly after
    allocate_executor switched the plan to another schedule (the earlier
    schedule stays alive), and switching back reuses it."""
    import torch
    so, cb = oracle
    A = _uneven(600, 256, 0.1, seed=19)
    plan = _plan(srt, torch, A, 64, 4, "orig", layout=2, waves_per_block=8, vec_width=1)
    pk = cb.pack_dense(A, 4, _oracle_table(so, 4, "orig"))
    plan.allocate_executor(64)
    kp64 = plan.split_counts().copy()
    B = torch.from_numpy(srt.workloads.dense_b(256, 64, seed=3)).cuda()
    ref = cb.exec_packed(pk, B.cpu().numpy(), nsplit=kp64)
    C = torch.full((600, 64), float("nan"), device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan(C, B, stream=s)
    for n in (448, 64, 448):
        plan.allocate_executor(n)          # another width: another schedule becomes current
        B2 = srt.workloads.dense_b(256, n, seed=n)
        C2 = _run(srt, torch, plan, B2)
        _bitexact(C2, cb.exec_packed(pk, B2, nsplit=plan.split_counts()))
        for _ in range(2):
            C.fill_(float("nan"))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            _bitexact(C.cpu().numpy(), ref)
    del g
    plan.close()
