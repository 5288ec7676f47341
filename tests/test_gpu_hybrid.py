"""Hybrid plans: 16-row panels at or above a density threshold on FP32
matrix cores, the rest register-tiled.

Bar: rows of register-tiled panels bit-identical to the oracle's chain over
the matrix with the dense panels removed; rows of dense panels within the
FP32 summation bound of the FP64 product (and normwise rel <= 1e-5); the
dense-panel list is the one the threshold rule gives; results repeat bit for
bit across launches and graph replays."""
import numpy as np
import pytest

from test_gpu_parity import _oracle_table, _run

pytestmark = pytest.mark.gpu


def _mixed(M, K, seed):
    """Panels alternating between 50 % and 92 % sparsity, plus ragged rows."""
    rng = np.random.default_rng(seed)
    dens = np.where((np.arange(M) // 16) % 3 == 0, 0.5, 0.08)
    A = (rng.random((M, K)) < dens[:, None]) * rng.standard_normal((M, K))
    return A.astype(np.float32)


def _dense_panels(A, tau):
    M, K = A.shape
    out = []
    for r0 in range(0, M, 16):
        r1 = min(M, r0 + 16)
        nz = np.count_nonzero(A[r0:r1])
        if nz and nz / ((r1 - r0) * K) >= tau:
            out.append(r0)
    return out


@pytest.mark.parametrize("M,K,N,tau,mapping", [(500, 1000, 200, 0.3, (4, "orig")), (256, 300, 64, 0.3, (4, "orig")),
                                               (333, 517, 130, 0.05, (4, "orig")), (1024, 4096, 512, 0.3, (4, "orig")),
                                               (500, 1000, 200, 0.3, (8, "orig")),
                                               (500, 1000, 200, 0.3, (6, "identity"))])
def test_hybrid_matches_oracle(srt, oracle, M, K, N, tau, mapping):
    """With M_r = 4 and 8 the register-tiled schedule leaves the dense
    panels' rows out; M_r = 6 panels straddle the 16-row panels, so the
    register tile writes those rows as zero and the dense launch overwrites
    them."""
    import torch
    so, cb = oracle
    print("hybrid %dx%d N=%d tau=%g mapping %s" % (M, K, N, tau, mapping), flush=True)
    A = srt.workloads.magnitude_pruned(M, K, 0.7, seed=1) if M == 1024 else _mixed(M, K, seed=M)
    B = srt.workloads.dense_b(K, N, seed=2)
    rp, ci, v = srt.workloads.to_csr(A)
    plan = srt.MatMulSpecialized(M, K, N, torch.from_numpy(v), torch.from_numpy(rp), torch.from_numpy(ci),
                                 mapping_id=srt.named_mapping(*mapping), dense_threshold=tau)
    want = _dense_panels(A, tau)
    dp = plan.dense_panels()
    assert list(dp["rows"]) == want and dp["count"] == len(want) > 0
    C = _run(srt, torch, plan, B)
    dense_rows = np.zeros(M, dtype=bool)
    for r0 in want:
        dense_rows[r0:r0 + 16] = True
    As = A.copy()
    As[dense_rows] = 0
    ref = cb.exec_packed(cb.pack_dense(As, mapping[0], _oracle_table(so, *mapping)), B, nsplit=plan.split_counts())
    assert np.array_equal(C[~dense_rows].view(np.uint32), ref[~dense_rows].view(np.uint32))
    C64 = cb.dense_f64(A, B)
    assert np.all(np.abs(C - C64) <= so.fp64_error_bound(A, B))
    rel = np.linalg.norm(C[dense_rows] - C64[dense_rows]) / np.linalg.norm(C64[dense_rows])
    assert rel <= 1e-5, rel
    # repeatable, and in a graph
    again = _run(srt, torch, plan, B)
    assert np.array_equal(again.view(np.uint32), C.view(np.uint32))
    Bd = torch.from_numpy(B).cuda()
    Cd = torch.zeros((M, N), device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan(Cd, Bd, stream=s)
    for _ in range(2):
        Cd.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(Cd.cpu().numpy().view(np.uint32), C.view(np.uint32))
    del g
    plan.close()


def test_hybrid_threshold_above_one_is_plain(srt, oracle):
    import torch
    so, cb = oracle
    A = _mixed(100, 90, seed=3)
    B = srt.workloads.dense_b(90, 70, seed=4)
    rp, ci, v = srt.workloads.to_csr(A)
    plan = srt.MatMulSpecialized(100, 90, 70, torch.from_numpy(v), torch.from_numpy(rp), torch.from_numpy(ci),
                                 mapping_id=srt.named_mapping(4, "orig"), dense_threshold=1.5)
    assert plan.dense_panels()["count"] == 0
    C = _run(srt, torch, plan, B)
    ref = cb.exec_packed(cb.pack_dense(A, 4, _oracle_table(so, 4, "orig")), B, nsplit=plan.split_counts())
    assert np.array_equal(C.view(np.uint32), ref.view(np.uint32))
    plan.close()


def test_hybrid_all_dense_split_k_and_width_change(srt, oracle):
    """Every panel dense (the dense launch splits K over several workgroups
    per output tile and adds the partials in split order): within the FP64
    bound, repeatable; a graph captured at one width replays bit-exact after
    the plan was widened (its K-split workspace grew; the old one stays
    alive), and the wide launch matches the bound too."""
    import torch
    so, cb = oracle
    M, K = 200, 3000
    A = srt.workloads.magnitude_pruned(M, K, 0.5, seed=7)
    B = srt.workloads.dense_b(K, 520, seed=8)
    rp, ci, v = srt.workloads.to_csr(A)
    plan = srt.MatMulSpecialized(M, K, 64, torch.from_numpy(v), torch.from_numpy(rp), torch.from_numpy(ci),
                                 mapping_id=srt.named_mapping(4, "orig"), dense_threshold=0.1)
    assert plan.dense_panels()["count"] == -(-M // 16)
    B64 = np.ascontiguousarray(B[:, :64])
    C = _run(srt, torch, plan, B64)
    assert np.all(np.abs(C - cb.dense_f64(A, B64)) <= so.fp64_error_bound(A, B64))
    assert np.array_equal(_run(srt, torch, plan, B64).view(np.uint32), C.view(np.uint32))
    Bd = torch.from_numpy(B64).cuda()
    Cd = torch.zeros((M, 64), device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan(Cd, Bd, stream=s)
    torch.cuda.synchronize()
    plan.allocate_executor(520)
    Cw = _run(srt, torch, plan, B)
    assert np.all(np.abs(Cw - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B))
    assert np.array_equal(np.ascontiguousarray(Cw[:, :64]).view(np.uint32), C.view(np.uint32))
    for _ in range(2):
        Cd.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(Cd.cpu().numpy().view(np.uint32), C.view(np.uint32))
    del g
    plan.close()


def test_tuned_hybrid_chooses_by_measurement(srt, oracle):
    """autotune.tuned_hybrid times the plain plan against its hybrids: at 50 %
    sparsity the dense panels on matrix cores win by a wide margin (about
    47 vs 80 us on this shape), at 95 % no panel reaches a threshold and the
    plain plan is kept; the chosen plan's result is within the FP64 bound."""
    import importlib
    import torch
    so, cb = oracle
    at = importlib.import_module("sparse-register-tiling_amd.autotune")
    A = srt.workloads.magnitude_pruned(1024, 4096, 0.5, seed=3)
    B = srt.workloads.dense_b(4096, 512, seed=4)
    plan, rep = at.tuned_hybrid(torch.from_numpy(A).to_sparse_csr(), 512)
    print("tuned hybrid 50 %:", rep, flush=True)
    assert rep["choice"] is not None and plan.dense_panels()["count"] > 0
    assert min(t for _, _, t in rep["candidates"]) < 0.8 * rep["plain_s"]
    C = _run(srt, torch, plan, B)
    assert np.all(np.abs(C - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B))
    plan.close()
    A2 = srt.workloads.magnitude_pruned(512, 2048, 0.95, seed=5)
    plan2, rep2 = at.tuned_hybrid(torch.from_numpy(A2).to_sparse_csr(), 256)
    assert rep2["choice"] is None and rep2["candidates"] == [] and plan2.dense_panels()["count"] == 0
    plan2.close()
