"""GPU tests of the CNN batched form: the 16 ResNet-50-style 1x1-conv layers
as one grouped launch.

Bar (FP32): every layer's result from the grouped launch is bit-identical to
the oracle's sequential fmaf chain over the same packed data (and so to the
layer's own MatMulSpecialized launch), and within the FP32 summation bound
of the FP64 dense product.  Bar (BF16 / FP8): the grouped MFMA launch is
bit-identical to each layer's own PanelPlan launch and within the
accumulation bound of the FP64 product of the quantised operands
(oracle.precision_oracle).  Graph replays equal the eager launch.
"""
import sys

import numpy as np
import pytest

from oracle import precision_oracle as po
from test_gpu_parity import _oracle_table

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
    layers = srt.cnn.cnn_layers(seed=0)
    return srt, torch, layers


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


@pytest.mark.parametrize("batch", [1, 4, 8])
@pytest.mark.parametrize("mapping,shape", [((4, "orig"), dict(vec_width=4, waves_per_block=8, nsplit=1)),
                                           ((8, "orig"), dict(vec_width=4, waves_per_block=8, nsplit=8)),
                                           ((4, "identity"), dict(vec_width=2, waves_per_block=16, nsplit=16))])
def test_grouped_fp32_matches_oracle(env, oracle, batch, mapping, shape):
    srt, torch, layers = env
    so, cb = oracle
    _log("cnn fp32 batch=%d mapping=%s shape=%s" % (batch, mapping, shape))
    X = srt.cnn.activations(layers, batch, seed=5)
    net = srt.cnn.CnnFp32(layers, batch, X, mapping=mapping, shape=shape)
    for y in net.Y:
        y.fill_(float("nan"))
    net()
    torch.cuda.synchronize()
    info = net.group.info()
    # N = 49 at batch 1 allows only one float per lane for the whole group.
    assert info["vec_width"] == (1 if batch == 1 else shape["vec_width"])
    table = _oracle_table(so, *mapping)
    for l, p, x, y in zip(layers, net.plans, X, net.Y):
        A = l.A.to_dense().numpy()
        B = x.cpu().numpy()
        C = y.cpu().numpy()
        pk = cb.pack_dense(A, mapping[0], table)
        ref = cb.exec_packed(pk, B, nsplit=p.split_counts())
        assert np.array_equal(_bits(C), _bits(ref)), "layer %s max |diff| %g" % (l.name, np.abs(C - ref).max())
        assert np.all(np.abs(C - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B)), l.name
    # one launch per layer gives the same bits
    got = [y.clone() for y in net.Y]
    net.per_layer()
    torch.cuda.synchronize()
    for a, b in zip(got, net.Y):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    net.close()


@pytest.mark.parametrize("batch", [1, 4])
def test_grouped_fp32_per_layer_shapes(env, oracle, batch):
    """Layers of one group with their own split, layout and K tile: still
    bit-identical to the oracle's chain over each layer's own packing and to
    one launch per layer."""
    srt, torch, layers = env
    so, cb = oracle
    _log("cnn fp32 per-layer shapes batch=%d" % batch)
    X = srt.cnn.activations(layers, batch, seed=6)
    cands = srt.cnn.layer_candidates(8)
    layer_shapes = [cands[(3 * i) % len(cands)] for i in range(len(layers))]
    net = srt.cnn.CnnFp32(layers, batch, X, mapping=(4, "orig"),
                          shape=dict(vec_width=4, waves_per_block=8, nsplit=8), layer_shapes=layer_shapes)
    for y in net.Y:
        y.fill_(float("nan"))
    net()
    torch.cuda.synchronize()
    table = _oracle_table(so, 4, "orig")
    for l, p, x, y, ls in zip(layers, net.plans, X, net.Y, layer_shapes):
        real = p.get_config()
        assert real.layout == ls["layout"] and real.nsplit == ls["nsplit"]
        A = l.A.to_dense().numpy()
        B = x.cpu().numpy()
        pk = cb.pack_dense_ktiled(A, 4, table, real.K_c) if real.K_c < A.shape[1] else cb.pack_dense(A, 4, table)
        ref = cb.exec_packed(pk, B, nsplit=p.split_counts())
        assert np.array_equal(_bits(y.cpu().numpy()), _bits(ref)), l.name
    got = [y.clone() for y in net.Y]
    net.per_layer()
    torch.cuda.synchronize()
    for a, b in zip(got, net.Y):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    net.close()
    with pytest.raises(ValueError):
        srt.cnn.CnnFp32(layers[:2], batch, X[:2], layer_shapes=[dict(vec_width=2), dict(nsplit=1)])


def test_grouped_fp32_graph_replay_and_rebind(env):
    srt, torch, layers = env
    _log("cnn fp32 graph + rebind")
    sub = layers[::3]
    X = srt.cnn.activations(sub, 2, seed=7)
    net = srt.cnn.CnnFp32(sub, 2, X, mapping=(4, "orig"))
    net()
    torch.cuda.synchronize()
    eager = [y.clone() for y in net.Y]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        net.group(stream=s)
    for rep in range(3):
        for y in net.Y:
            y.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(eager, net.Y):
            assert torch.equal(a.view(torch.int32), b.view(torch.int32)), "replay %d" % rep
    # rebind to narrower operands: columns [0, n) of each layer, same bits
    n = [max(1, y.shape[1] // 3) for y in net.Y]
    Y2 = [torch.full((y.shape[0], ni), float("nan"), device="cuda") for y, ni in zip(net.Y, n)]
    X2 = [x[:, :ni].contiguous() for x, ni in zip(X, n)]
    net.group.rebind(X2, Y2)
    net.group()
    torch.cuda.synchronize()
    for a, b, ni in zip(eager, Y2, n):
        assert torch.equal(a[:, :ni].contiguous().view(torch.int32), b.view(torch.int32))
    del g
    net.close()


def test_grouped_fp32_rejects_mixed_plans(env):
    srt, torch, layers = env
    _log("cnn fp32 mixed plans")
    l0, l1 = layers[0], layers[5]
    p0 = srt.MatMulSpecialized.from_csr(l0.A, 64, mapping_id=srt.named_mapping(4, "orig"))
    p1 = srt.MatMulSpecialized.from_csr(l1.A, 64, mapping_id=srt.named_mapping(8, "orig"))
    p2 = srt.MatMulSpecialized.from_csr(l1.A, 64, mapping_id=srt.named_mapping(4, "orig"),
                                        tile_config=srt.TileConfig(waves_per_block=4, nsplit=4))
    B0 = torch.rand((l0.K, 64), device="cuda")
    B1 = torch.rand((l1.K, 64), device="cuda")
    C0 = torch.empty((l0.M, 64), device="cuda")
    C1 = torch.empty((l1.M, 64), device="cuda")
    with pytest.raises(srt.SrtError, match="executors"):
        srt.cnn.GroupedPlans([p0, p1], [B0, B1], [C0, C1])
    with pytest.raises(srt.SrtError, match="waves_per_block"):
        srt.cnn.GroupedPlans([p0, p2], [B0, B1], [C0, C1])
    with pytest.raises(ValueError):
        srt.cnn.GroupedPlans([p0, p0], [B0, B0[:32]], [C0, C0])  # B has fewer rows than A has columns
    # an empty problem (n = 0) is skipped
    g = srt.cnn.GroupedPlans([p0, p0], [B0, B0], [C0, C1[:l0.M]], n=[64, 0])
    g()
    torch.cuda.synchronize()
    ref = torch.empty_like(C0)
    p0(ref, B0)
    torch.cuda.synchronize()
    assert torch.equal(ref.view(torch.int32), C0.view(torch.int32))
    for p in (p0, p1, p2):
        p.close()


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
@pytest.mark.parametrize("shape", [0, 1, 4, 6, 8])
def test_grouped_panels_match_single(env, dtype, shape):
    srt, torch, layers = env
    batch = 1
    _log("cnn %s shape=%d" % (dtype, shape))
    sub = layers[1::2]
    X = srt.cnn.activations(sub, batch, seed=9)
    net = srt.cnn.CnnPanels(sub, batch, X, dtype=dtype, shape=shape)
    for y in net.Y:
        y.fill_(float("nan"))
    net()
    torch.cuda.synchronize()
    for l, p, xt, s, x, y in zip(sub, net.plans, net.Xt, net.scales, X, net.Y):
        single = torch.full_like(y, float("nan"))
        p(single, xt, s, shape=shape, ksplit=1)
        torch.cuda.synchronize()
        assert torch.equal(single.view(torch.int32), y.view(torch.int32)), l.name
        A = l.A.to_dense().numpy()
        B = x.cpu().numpy()
        C = y.cpu().numpy()
        ref, bound = po.reference_f64(A, B, dtype)
        assert np.all(np.abs(C - ref) <= bound + 1e-30), l.name
        if dtype == "bf16":
            assert po.rel_error_vs_fp64(C, A, B) <= 1e-2
    net.close()


def test_tuners_pick_valid_configs(env):
    srt, torch, layers = env
    _log("cnn tuners")
    sub = layers[:4]
    X = srt.cnn.activations(sub, 1, seed=2)
    recs = srt.cnn.tune_fp32(sub, 1, X, mappings=[(4, "orig"), (8, "orig")],
                             shapes=srt.cnn.FP32_SHAPES[:2], reps=3)
    assert len(recs) == 4 and recs[0]["seconds"] <= recs[-1]["seconds"]
    assert all(r["seconds"] > 0 and r["workgroups"] > 0 for r in recs)
    recs = srt.cnn.tune_panels(sub, 1, X, "bf16", shapes=(0, 2), reps=3)
    assert [r["seconds"] for r in recs] == sorted(r["seconds"] for r in recs)
