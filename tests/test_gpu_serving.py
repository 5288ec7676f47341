"""GPU tests of the serving form: resident BF16 weights, admission by HBM
budget, continuous batching into grouped persistent launches replayed from
HIP graphs.

Bar: every served result is within the FP32-accumulation bound of the FP64
product of the BF16-quantised operands (oracle.precision_oracle), its
normwise error against the FP64 product of the unquantised inputs is below
1e-2, and each result is bit-identical alone and inside a mixed batch, with
graphs and without.
"""
import sys

import numpy as np
import pytest

from oracle import precision_oracle as po

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import importlib
    import torch
    torch.cuda.init()
    srt = importlib.import_module("sparse-register-tiling_amd")
    shapes = [(256, 128), (128, 256), (300, 96), (64, 520)]
    Ws = [srt.workloads.magnitude_pruned(M, K, 0.9, seed=i) for i, (M, K) in enumerate(shapes)]
    plans = [srt.precision.PanelPlan.from_dense(torch.from_numpy(W), "bf16") for W in Ws]
    return srt, torch, shapes, Ws, plans


def _requests(srt, torch, shapes, num, qps, seed, n_max=96):
    return srt.serving.poisson_requests(num, qps, shapes, seed=seed, n_max=n_max)


def _check(req, W):
    B = req.B.float().cpu().numpy()
    C = req.C.cpu().numpy()
    ref, bound = po.reference_f64(W, B, "bf16")
    assert np.all(np.abs(C - ref) <= bound + 1e-30), "request %d" % req.rid
    assert po.rel_error_vs_fp64(C, W, B) <= 1e-2


@pytest.mark.parametrize("graphs", [True, False])
def test_served_results_and_latency(env, graphs):
    srt, torch, shapes, Ws, plans = env
    _log("serving graphs=%s" % graphs)
    server = srt.serving.Server(plans, budget_bytes=64 << 20, max_batch=8, slots=3, use_graphs=graphs, grid=256)
    reqs = _requests(srt, torch, shapes, 40, qps=20000.0, seed=3)
    st = server.serve(reqs, keep_results=True)
    _log("  ", st)
    assert st["requests"] == 40 and st["batches"] >= 5
    for r in reqs:
        assert r.done >= r.arrival
        _check(r, Ws[r.weight])
    assert server.pool.used == 0


def test_alone_equals_mixed(env):
    srt, torch, shapes, Ws, plans = env
    _log("serving alone vs mixed")
    server = srt.serving.Server(plans, budget_bytes=64 << 20, max_batch=16, slots=2, grid=64)
    reqs = _requests(srt, torch, shapes, 12, qps=1e6, seed=5)
    alone = []
    for r in reqs:
        server.run_batch([r])
        alone.append(r.C.cpu().numpy())
    mixed = [srt.serving.Request(rid=r.rid, weight=r.weight, B=r.B) for r in reqs]
    server.run_batch(list(reversed(mixed)), slot_index=1)
    for r, a in zip(mixed, alone):
        assert np.array_equal(r.C.cpu().numpy().view(np.uint32), a.view(np.uint32)), "request %d" % r.rid


def test_admission_by_budget(env):
    srt, torch, shapes, Ws, plans = env
    _log("serving small budget")
    probe = srt.serving.Request(rid=0, weight=2, B=torch.zeros((96, 96), dtype=torch.bfloat16, device="cuda"))
    server = srt.serving.Server(plans, budget_bytes=256 << 10, max_batch=64, slots=2, grid=64)
    assert server.request_bytes(probe) < server.pool.nbytes < 4 * server.request_bytes(probe)
    reqs = _requests(srt, torch, shapes, 24, qps=1e6, seed=7)  # all arrive at once
    st = server.serve(reqs, keep_results=True)
    _log("  ", st)
    assert st["batches"] > 3          # the budget forced several small batches
    for r in reqs:
        _check(r, Ws[r.weight])
    with pytest.raises(MemoryError):
        big = srt.serving.Request(rid=99, weight=3, B=torch.zeros((520, 512), dtype=torch.bfloat16, device="cuda"))
        server.run_batch([big])


@pytest.fixture(scope="module")
def full_env():
    """The serving workload at its real shapes: 48 weights, 12 layers x
    {1024x1024 x2, 4096x1024, 1024x4096}, 90% magnitude-pruned, BF16."""
    import importlib
    import torch
    srt = importlib.import_module("sparse-register-tiling_amd")
    Ws = srt.serving.serving_weights()
    plans = [srt.precision.PanelPlan.from_dense(torch.from_numpy(W), "bf16") for W in Ws]
    return srt, torch, Ws, plans


def test_full_serving_workload(full_env):
    """48 resident weights, Poisson requests with log-uniform N up to 512:
    every served result within the bound of the FP64 product of the BF16
    operands and rel <= 1e-2 of the FP64 product of the unquantised inputs,
    and each result bit-identical alone and in the mix."""
    srt, torch, Ws, plans = full_env
    shapes = [W.shape for W in Ws]
    assert len(Ws) == 48 and sorted(set(shapes)) == [(1024, 1024), (1024, 4096), (4096, 1024)]
    _log("serving 48 weights at full shapes")
    server = srt.serving.Server(plans, budget_bytes=256 << 20, max_batch=32, slots=3, grid=1024)
    reqs = srt.serving.poisson_requests(48, 50000.0, shapes, seed=11, n_max=512)
    # the sample spans every weight shape and N up to 512
    assert max(r.N for r in reqs) > 256 and {shapes[r.weight] for r in reqs} == set(shapes)
    big = srt.serving.Request(rid=999, weight=2, B=(torch.rand((1024, 512), device="cuda") * 2 - 1).to(torch.bfloat16))
    reqs.append(big)
    st = server.serve(reqs, keep_results=True)
    _log("  ", st)
    assert st["requests"] == len(reqs)
    served = {}
    for r in reqs:
        _check(r, Ws[r.weight])
        served[r.rid] = r.C.cpu().numpy()
    # alone == mixed at these shapes (a batch holding every request, reversed)
    sample = reqs[:24] + [big]
    mixed = [srt.serving.Request(rid=r.rid, weight=r.weight, B=r.B) for r in sample]
    server.run_batch(list(reversed(mixed)), slot_index=1)
    for r in mixed:
        alone = srt.serving.Request(rid=r.rid, weight=r.weight, B=r.B)
        server.run_batch([alone], slot_index=2)
        a = alone.C.cpu().numpy()
        assert np.array_equal(r.C.cpu().numpy().view(np.uint32), a.view(np.uint32)), "request %d" % r.rid
        assert np.array_equal(a.view(np.uint32), served[r.rid].view(np.uint32)), "request %d" % r.rid
    assert server.pool.used == 0


# ------------------------------------------------------------------ register-tiled sparse weights
def _oracle_served(so, cb, W, B_bf16, plan):
    """The oracle's chain over the BF16-valued weight (its own packing with
    the 4-row "orig" strategy) and the BF16 B widened to FP32."""
    import torch
    Wq = torch.from_numpy(W).to(torch.bfloat16).float().numpy()
    pk = cb.pack_dense(Wq, 4, so.named_table(4, "orig"))
    B = B_bf16.float().cpu().numpy()
    return cb.exec_packed(pk, B, nsplit=plan.split_counts())


@pytest.fixture(scope="module")
def rt_env():
    import importlib
    import torch
    torch.cuda.init()
    srt = importlib.import_module("sparse-register-tiling_amd")
    shapes = [(256, 128), (128, 256), (300, 96), (64, 520)]
    Ws = [srt.workloads.magnitude_pruned(M, K, 0.9, seed=i) for i, (M, K) in enumerate(shapes)]
    weights = srt.serving.RegTileWeights(Ws)
    return srt, torch, shapes, Ws, weights


@pytest.mark.parametrize("graphs", [True, False])
def test_regtile_served_results_exact(rt_env, oracle, graphs):
    """Register-tiled sparse serving: every result equals the oracle's chain
    over the BF16 operands bit for bit (and is within 1e-2 of FP64)."""
    srt, torch, shapes, Ws, weights = rt_env
    so, cb = oracle
    _log("regtile serving graphs=%s" % graphs)
    server = srt.serving.Server(weights, budget_bytes=64 << 20, max_batch=8, slots=3, use_graphs=graphs, grid=256)
    reqs = _requests(srt, torch, shapes, 40, qps=20000.0, seed=13)
    st = server.serve(reqs, keep_results=True)
    _log("  ", st)
    assert st["requests"] == 40 and st["batches"] >= 5
    for r in reqs:
        ref = _oracle_served(so, cb, Ws[r.weight], r.B, weights[r.weight])
        got = r.C.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), "request %d" % r.rid
        assert po.rel_error_vs_fp64(got, Ws[r.weight], r.B.float().cpu().numpy()) <= 1e-2
    assert server.pool.used == 0


def test_regtile_one_spmm_in_each_launch(rt_env, oracle):
    """With the middle launch on (mid_cols = 128), a batch with one SpMM per
    table: 20 columns (narrow, 1 float per lane), 100 (middle, 2 floats per
    lane) and 200 (wide, 4 floats per lane), on three weights; every result
    is the oracle's chain bit for bit, and each request alone gives the same
    bits."""
    srt, torch, shapes, Ws, weights = rt_env
    so, cb = oracle
    _log("regtile narrow / middle / wide launches")
    server = srt.serving.Server(weights, budget_bytes=64 << 20, max_batch=8, slots=2, grid=128, mid_cols=128)
    g = torch.Generator().manual_seed(21)
    reqs = [srt.serving.Request(rid=i, weight=w, B=(torch.rand((shapes[w][1], n), generator=g) * 2 - 1)
                                .to(torch.bfloat16).cuda())
            for i, (w, n) in enumerate(((0, 20), (1, 100), (3, 200)))]
    server.run_batch(reqs, slot_index=0)
    hdr = [int(x) for x in server.slots[0].hdr]
    assert (hdr[2], hdr[4], hdr[6]) == (1, 1, 1)
    for r in reqs:
        got = r.C.cpu().numpy()
        ref = _oracle_served(so, cb, Ws[r.weight], r.B, weights[r.weight])
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), "request %d" % r.rid
        alone = srt.serving.Request(rid=r.rid, weight=r.weight, B=r.B)
        server.run_batch([alone], slot_index=1)
        assert np.array_equal(alone.C.cpu().numpy().view(np.uint32), got.view(np.uint32)), "request %d" % r.rid
    assert server.pool.used == 0


def test_regtile_alone_equals_mixed_and_budget(rt_env, oracle):
    srt, torch, shapes, Ws, weights = rt_env
    so, cb = oracle
    server = srt.serving.Server(weights, budget_bytes=64 << 20, max_batch=16, slots=2, grid=64)
    reqs = _requests(srt, torch, shapes, 12, qps=1e6, seed=15)
    alone = []
    for r in reqs:
        server.run_batch([r])
        alone.append(r.C.cpu().numpy())
    mixed = [srt.serving.Request(rid=r.rid, weight=r.weight, B=r.B) for r in reqs]
    server.run_batch(list(reversed(mixed)), slot_index=1)
    # 12 requests over 4 weights: requests of one weight ran merged, one SpMM
    # each (in the wide, narrow or middle table)
    hdr = server.slots[1].hdr
    assert hdr[2] + hdr[4] + hdr[6] == len({r.weight for r in mixed}) < len(mixed)
    assert server.pool.used == 0
    for r, a in zip(mixed, alone):
        assert np.array_equal(r.C.cpu().numpy().view(np.uint32), a.view(np.uint32)), "request %d" % r.rid
    small = srt.serving.Server(weights, budget_bytes=256 << 10, max_batch=64, slots=2, grid=64)
    with pytest.raises(MemoryError):
        big = srt.serving.Request(rid=99, weight=3, B=torch.zeros((520, 512), dtype=torch.bfloat16, device="cuda"))
        small.run_batch([big])
    # a budget this small is full at every batch: merged blocks take the
    # members' own blocks, or the members run alone; the results stay exact
    tight = _requests(srt, torch, shapes, 24, qps=1e6, seed=17)
    st = small.serve(tight, keep_results=True)
    assert st["batches"] > 1
    assert small.pool.used == 0
    for r in tight:
        ref = _oracle_served(so, cb, Ws[r.weight], r.B, weights[r.weight])
        assert np.array_equal(r.C.cpu().numpy().view(np.uint32), ref.view(np.uint32)), "request %d" % r.rid


def test_regtile_full_serving_workload(oracle):
    """The 48 weights at their real shapes on register-tiled sparse panels:
    resident bytes well below the dense BF16 size, every result exact
    against the oracle, alone == mixed."""
    import importlib
    import torch
    srt = importlib.import_module("sparse-register-tiling_amd")
    so, cb = oracle
    Ws = srt.serving.serving_weights()
    weights = srt.serving.RegTileWeights(Ws)
    dense_bf16 = sum(W.size * 2 for W in Ws)
    nnz = sum(int((W != 0).sum()) for W in Ws)
    server = srt.serving.Server(weights, budget_bytes=256 << 20, max_batch=32, slots=3, grid=1024)
    resident = server.resident_bytes()
    _log("regtile 48 weights: resident %.1f MB, dense BF16 %.1f MB, nnz BF16 %.1f MB"
         % (resident / 1e6, dense_bf16 / 1e6, 2 * nnz / 1e6))
    assert resident < 0.7 * dense_bf16
    shapes = [W.shape for W in Ws]
    reqs = srt.serving.poisson_requests(40, 50000.0, shapes, seed=21, n_max=512)
    big = srt.serving.Request(rid=999, weight=3, B=(torch.rand((4096, 512), device="cuda") * 2 - 1).to(torch.bfloat16))
    reqs.append(big)
    st = server.serve(reqs, keep_results=True)
    _log("  ", st)
    for r in reqs:
        ref = _oracle_served(so, cb, Ws[r.weight], r.B, weights[r.weight])
        assert np.array_equal(r.C.cpu().numpy().view(np.uint32), ref.view(np.uint32)), "request %d" % r.rid
    mixed = [srt.serving.Request(rid=r.rid, weight=r.weight, B=r.B) for r in reqs[:16]]
    server.run_batch(list(reversed(mixed)), slot_index=1)
    for r, m in zip(reqs[:16], mixed):
        assert np.array_equal(r.C.cpu().numpy().view(np.uint32), m.C.cpu().numpy().view(np.uint32))
