"""GPU parity of the precision form (BF16 / FP8-e4m3 on MFMA 16-row panels).

Bar: the packed blocks, the FP8 row scales and the converted B operand are
bit-identical to the oracle's; C is within the FP32-accumulation bound of the
FP64 product of the quantised operands (oracle.precision_oracle.reference_f64)
for every launch shape and K split; the normwise error against the FP64
product of the unquantised inputs is reported and bounded (BF16 <= 1e-2).
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
def torch_cuda():
    import torch
    torch.cuda.init()
    return torch


def _sparse(M, K, density, seed):
    rng = np.random.default_rng(seed)
    return ((rng.random((M, K)) < density) * rng.standard_normal((M, K))).astype(np.float32)


def _run(srt, torch, A, B, dtype, shape=None, ksplit=None, check_pack=True):
    plan = srt.precision.PanelPlan.from_dense(torch.from_numpy(A), dtype)
    if check_pack:
        slot, kb, blocks = po.pack_blocks(A, dtype)
        got = plan.blocks()
        assert np.array_equal(got["slot"], slot)
        assert np.array_equal(got["blk_kb"], kb)
        assert np.array_equal(got["blocks"], blocks)
        if dtype == "fp8":
            assert np.array_equal(plan.row_scales(), po.row_scales(A))
    Bd = torch.from_numpy(B).cuda()
    Bq, t = plan.prepare_b(Bd)
    if check_pack:
        qb, tref = po.quantize_b(B, dtype)
        gotb = Bq.cpu().view(torch.int16).numpy().view(np.uint16) if dtype == "bf16" else Bq.cpu().numpy()
        K = B.shape[0]
        assert gotb.shape == (B.shape[1], -(-K // 32) * 32)
        assert np.array_equal(gotb[:, :K], qb.T)
        assert not gotb[:, K:].any()
        if dtype == "fp8":
            assert np.array_equal(t.cpu().numpy(), tref)
    C = torch.full((A.shape[0], B.shape[1]), float("nan"), device="cuda")
    plan(C, Bq, t, shape=shape, ksplit=ksplit)
    torch.cuda.synchronize()
    return plan, C.cpu().numpy()


def _within(C, A, B, dtype):
    ref, bound = po.reference_f64(A, B, dtype)
    ok = np.abs(C.astype(np.float64) - ref) <= bound + 1e-30
    assert ok.all(), "max excess %g" % (np.abs(C - ref) - bound).max()


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
@pytest.mark.parametrize("shape_mkn", [(16, 32, 16), (37, 70, 24), (1000, 300, 100), (130, 257, 3), (1, 1, 1)])
def test_small_shapes(srt, torch_cuda, dtype, shape_mkn):
    M, K, N = shape_mkn
    _log("precision", dtype, shape_mkn)
    A = _sparse(M, K, 0.3, M + K)
    B = np.random.default_rng(N).uniform(-1, 1, (K, N)).astype(np.float32)
    _, C = _run(srt, torch_cuda, A, B, dtype)
    _within(C, A, B, dtype)


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
@pytest.mark.parametrize("shape", list(range(10)))
@pytest.mark.parametrize("ksplit", [1, 2, 5])
def test_launch_shapes_and_splits(srt, torch_cuda, dtype, shape, ksplit):
    _log("precision shape", dtype, shape, ksplit)
    A = _sparse(200, 600, 0.4, 7)
    A[32:48, :] = 0.0            # an empty panel
    A[:, 64:128] = 0.0           # empty blocks in every panel
    B = np.random.default_rng(8).uniform(-1, 1, (600, 136)).astype(np.float32)
    _, C = _run(srt, torch_cuda, A, B, dtype, shape=shape, ksplit=ksplit, check_pack=(shape == 0 and ksplit == 1))
    _within(C, A, B, dtype)


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
def test_unaligned_b_and_empty(srt, torch_cuda, dtype):
    torch = torch_cuda
    _log("precision unaligned", dtype)
    A = _sparse(64, 96, 0.5, 9)
    B = np.random.default_rng(10).uniform(-1, 1, (96, 50)).astype(np.float32)
    plan = srt.precision.PanelPlan.from_dense(torch.from_numpy(A), dtype)
    Bq, t = plan.prepare_b(torch.from_numpy(B).cuda())
    big = torch.zeros((50, 99), dtype=Bq.dtype, device="cuda")
    view = big[:, 1:97]           # 2- or 1-byte offset: not 16-byte aligned, odd row stride
    view.copy_(Bq)
    C = torch.empty((64, 50), device="cuda")
    with pytest.raises(srt.SrtError):
        plan(C, view, t)          # the kernel reads 16-byte rows: refused, not misread
    wide = torch.zeros((50, 128), dtype=Bq.dtype, device="cuda")
    wide[:, :96].copy_(Bq)        # a wider, aligned row stride works
    plan(C, wide, t)
    torch.cuda.synchronize()
    _within(C.cpu().numpy(), A, B, dtype)
    Z = np.zeros((40, 30), dtype=np.float32)
    pz = srt.precision.PanelPlan.from_dense(torch.from_numpy(Z), dtype)
    assert pz.stats()["nonzero_blocks"] == 0
    Bz, tz = pz.prepare_b(torch.ones((30, 8), device="cuda"))
    Cz = torch.full((40, 8), 5.0, device="cuda")
    pz(Cz, Bz, tz)
    torch.cuda.synchronize()
    assert (Cz.cpu().numpy() == 0).all()


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
@pytest.mark.parametrize("mk,s", [(mk, s) for mk in ((4096, 1024), (1024, 4096)) for s in (0.5, 0.6, 0.7)])
def test_precision_workloads_error_vs_fp64(srt, torch_cuda, dtype, mk, s):
    """All six precision configurations (magnitude-pruned, N = 512)."""
    M, K = mk
    _log("precision workload", dtype, mk, s)
    case = srt.workloads.SpMMCase(M, K, s, 512)
    A, B = srt.workloads.case_inputs(case, seed=0)
    _, C = _run(srt, torch_cuda, A, B, dtype, check_pack=False)
    _within(C, A, B, dtype)
    err = po.rel_error_vs_fp64(C, A, B)
    _log("  rel error vs FP64 (unquantised inputs): %.3e" % err)
    assert err <= (1e-2 if dtype == "bf16" else 1e-1)


def test_invalid_inputs(srt, torch_cuda):
    torch = torch_cuda
    with pytest.raises(srt.SrtError):
        srt.precision.PanelPlan(2, 4, torch.ones(1), torch.tensor([0, 1, 1]), torch.tensor([7]), "bf16")
    with pytest.raises(srt.SrtError):
        srt.precision.PanelPlan(2, 4, torch.ones(2), torch.tensor([0, 2, 1]), torch.tensor([0, 1]), "bf16")
    with pytest.raises(ValueError):
        srt.precision.PanelPlan(2, 4, torch.ones(1), torch.tensor([0, 1, 1]), torch.tensor([0]), "fp16")


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
@pytest.mark.parametrize("shape,ksplit", [(6, 4), (0, 2), (4, 5), (8, 3)])
def test_k_split_sum_is_deterministic(srt, torch_cuda, dtype, shape, ksplit):
    """The K splits' partial tiles are added in split order: repeated launches
    (and a graph replaying them) give the same bits."""
    torch = torch_cuda
    _log("precision split determinism", dtype, shape, ksplit)
    A = _sparse(1000, 1300, 0.3, 21)
    B = np.random.default_rng(22).uniform(-1, 1, (1300, 200)).astype(np.float32)
    plan = srt.precision.PanelPlan.from_dense(torch.from_numpy(A), dtype)
    Bq, t = plan.prepare_b(torch.from_numpy(B).cuda())
    plan.reserve(200, ksplit)
    C = torch.full((1000, 200), float("nan"), device="cuda")
    plan(C, Bq, t, shape=shape, ksplit=ksplit)
    torch.cuda.synchronize()
    first = C.clone()
    _within(first.cpu().numpy(), A, B, dtype)
    for _ in range(5):
        C.fill_(float("nan"))
        plan(C, Bq, t, shape=shape, ksplit=ksplit)
    torch.cuda.synchronize()
    assert torch.equal(C.view(torch.int32), first.view(torch.int32))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan(C, Bq, t, shape=shape, ksplit=ksplit)
    for _ in range(3):
        C.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(C.view(torch.int32), first.view(torch.int32))


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
@pytest.mark.parametrize("shape", [0, 3, 8])
def test_compact_blocks_alone_and_grouped(srt, torch_cuda, dtype, shape):
    """Only blocks holding a nonzero are stored: a block-sparse matrix keeps a
    fraction of the dense block count, and its results -- alone, with a K
    split, and in a grouped launch beside a plan with every block stored --
    match the oracle's bound and each other bit for bit."""
    torch = torch_cuda
    _log("precision compact", dtype, shape)
    rng = np.random.default_rng(31)
    M, K = 300, 700
    P, KB = -(-M // 16), -(-K // 32)
    keep = rng.random((P, KB)) < 0.15          # about 15 % of the 16 x 32 blocks occupied
    keep[3] = False                            # an empty panel
    mask = np.repeat(np.repeat(keep, 16, axis=0), 32, axis=1)[:M, :K]
    A = (mask * (rng.random((M, K)) < 0.5) * rng.standard_normal((M, K))).astype(np.float32)
    B = rng.uniform(-1, 1, (K, 96)).astype(np.float32)
    plan, C = _run(srt, torch, A, B, dtype, shape=shape, ksplit=1)
    st = plan.stats()
    slot, _, blocks = po.pack_blocks(A, dtype)
    assert st["stored_blocks"] == st["nonzero_blocks"] == len(blocks) < P * KB // 4
    _within(C, A, B, dtype)
    Bq, t = plan.prepare_b(torch.from_numpy(B).cuda())
    C2 = torch.full((M, 96), float("nan"), device="cuda")
    plan.reserve(96, 3)
    plan(C2, Bq, t, shape=shape, ksplit=3)
    torch.cuda.synchronize()
    _within(C2.cpu().numpy(), A, B, dtype)
    # grouped: the compact plan beside a plan with every block present
    D = _sparse(200, 700, 0.5, 32)
    dense = srt.precision.PanelPlan.from_dense(torch.from_numpy(D), dtype)
    assert dense.stats()["stored_blocks"] == 13 * 22
    Dq, td = dense.prepare_b(torch.from_numpy(B).cuda())
    Cd = torch.full((200, 96), float("nan"), device="cuda")
    dense(Cd, Dq, td, shape=shape, ksplit=1)
    Cg = [torch.full((M, 96), float("nan"), device="cuda"), torch.full((200, 96), float("nan"), device="cuda")]
    g = srt.cnn.GroupedPanels([plan, dense], [Bq, Dq], Cg, col_scales=[t, td], shape=shape, grid=37)
    for _ in range(2):
        g()
        torch.cuda.synchronize()
        assert np.array_equal(Cg[0].cpu().numpy().view(np.uint32), C.view(np.uint32))
        assert np.array_equal(Cg[1].cpu().numpy().view(np.uint32), Cd.cpu().numpy().view(np.uint32))
