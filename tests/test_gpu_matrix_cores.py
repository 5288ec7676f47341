"""Matrix-core execution of the M_r = 4 "full" packing (TileConfig.matrix_cores
= 1, one v_mfma_f32_4x4x1_16b outer product per packed column).

Bar: bit-identical to the oracle's chain over the same packing (which the
CPU suite shows is the CSR chain) and to the register tile's executor for
that packing run as one chain per panel; results repeat across launches,
graph replays and width changes; configurations it cannot run are refused
with an error."""
import numpy as np
import pytest

from test_gpu_parity import _bitexact, _run

pytestmark = pytest.mark.gpu

CASES = [
    (2048, 512, 0.9, 512),    # the pruned-FFN headline shape
    (512, 512, 0.7, 512),
    (1000, 640, 0.95, 200),   # K at the LDS limit, N not a multiple of 64
    (37, 90, 0.3, 65),        # ragged panel, one column past a slice
    (4, 1, 1.0, 1),
    (130, 300, 0.0, 64),      # no entries
    (2048, 512, 0.99, 1024),
    (8192, 256, 0.9, 128),    # many panels per workgroup range
]


def _plan(srt, torch, A, N, **cfg):
    rp, ci, v = srt.workloads.to_csr(A)
    return srt.MatMulSpecialized(A.shape[0], A.shape[1], N, torch.from_numpy(v), torch.from_numpy(rp),
                                 torch.from_numpy(ci), tile_config=srt.TileConfig(**cfg),
                                 mapping_id=srt.named_mapping(4, "full"))


def _matrix(srt, M, K, s, seed):
    A = srt.workloads.random_sparse(M, K, 1.0 - s, seed=seed)
    if M > 3:
        A[M // 3] = 0.0  # an empty row
    return A


@pytest.mark.parametrize("M,K,s,N", CASES)
def test_matrix_cores_bitexact(srt, oracle, M, K, s, N):
    import torch
    so, cb = oracle
    print("matrix cores %dx%d s=%g N=%d" % (M, K, s, N), flush=True)
    A = _matrix(srt, M, K, s, seed=M + K)
    B = srt.workloads.dense_b(K, N, seed=3)
    ref = cb.exec_packed(cb.pack_dense(A, 4, so.named_table(4, "full")), B)
    plan = _plan(srt, torch, A, N, matrix_cores=1)
    assert plan.get_config().matrix_cores == 1
    _bitexact(_run(srt, torch, plan, B), ref)
    plan.close()
    # the register tile on the same packing: one chain per panel gives the
    # same bits; split chains give the oracle's split order
    tile = _plan(srt, torch, A, N, nsplit=1)
    assert tile.get_config().matrix_cores == 0
    _bitexact(_run(srt, torch, tile, B), ref)
    tile.close()
    tile = _plan(srt, torch, A, N)
    pk = cb.pack_dense(A, 4, so.named_table(4, "full"))
    _bitexact(_run(srt, torch, tile, B), cb.exec_packed(pk, B, nsplit=tile.split_counts()))
    tile.close()


def test_matrix_cores_graph_and_width_change(srt, oracle):
    """A graph captured at one width replays bit-exact after the plan was
    prepared for another width (the earlier panel ranges stay alive)."""
    import torch
    so, cb = oracle
    M, K = 1024, 384
    A = _matrix(srt, M, K, 0.9, seed=5)
    B = srt.workloads.dense_b(K, 512, seed=6)
    table = so.named_table(4, "full")
    ref = cb.exec_packed(cb.pack_dense(A, 4, table), B)
    plan = _plan(srt, torch, A, 512, matrix_cores=1)
    Bd = torch.from_numpy(B).cuda()
    Cd = torch.zeros((M, 512), device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan(Cd, Bd, stream=s)
    torch.cuda.synchronize()
    plan.allocate_executor(64)
    B64 = np.ascontiguousarray(B[:, :64])
    _bitexact(_run(srt, torch, plan, B64), np.ascontiguousarray(ref[:, :64]))
    for _ in range(3):
        Cd.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _bitexact(Cd.cpu().numpy(), ref)
    del g
    plan.close()


def test_matrix_cores_refused_where_it_cannot_run(srt):
    import torch
    A = _matrix(srt, 64, 96, 0.8, seed=1)
    rp, ci, v = srt.workloads.to_csr(A)
    args = (64, 96, 32, torch.from_numpy(v), torch.from_numpy(rp), torch.from_numpy(ci))
    with pytest.raises(srt.SrtError, match="only kernel pattern is 15"):
        srt.MatMulSpecialized(*args, tile_config=srt.TileConfig(matrix_cores=1),
                              mapping_id=srt.named_mapping(4, "orig"))
    big = _matrix(srt, 16, 2048, 0.9, seed=2)
    rp, ci, v = srt.workloads.to_csr(big)
    with pytest.raises(srt.SrtError, match="K <="):
        srt.MatMulSpecialized(16, 2048, 32, torch.from_numpy(v), torch.from_numpy(rp), torch.from_numpy(ci),
                              tile_config=srt.TileConfig(matrix_cores=1), mapping_id=srt.named_mapping(4, "full"))
    plan = srt.MatMulSpecialized(*args, tile_config=srt.TileConfig(matrix_cores=1),
                                 mapping_id=srt.named_mapping(4, "full"))
    with pytest.raises(srt.SrtError, match="column indices"):
        plan.release_columns()
    plan.close()
