"""GPU parity: the HIP path against the oracle, through the C ABI.

Bar: packing (per-panel nkern_counts, col_indices, values, padding) is
identical to the oracle's; C = A.B is bit-identical to the oracle's
sequential fmaf chain over the same packed data (plain FP32 computes the
same fused multiply-adds in the same order), and within the FP32 summation
bound of an FP64 dense product:  |C - C64| <= 2 * L * 2^-24 * (|A| |B|),
L = longest row of A.  The CSR kernel is held to the same bar against the
restated CSR executor.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN_MAPPINGS

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    torch.cuda.init()
    return torch


def _oracle_table(so, M_r, name):
    """The oracle's own restatement of a named strategy: its table and the
    strategy's kernel list (the ids the product must reproduce)."""
    return so.named_table(M_r, name)


def _plan(srt, torch, A, n, M_r, name, index_dtype=np.int64, **cfg):
    rp, ci, v = srt.workloads.to_csr(A, index_dtype)
    return srt.MatMulSpecialized(A.shape[0], A.shape[1], n, torch.from_numpy(v), torch.from_numpy(rp),
                                 torch.from_numpy(ci), tile_config=srt.TileConfig(**cfg),
                                 mapping_id=srt.named_mapping(M_r, name))


def _check_packed(pk_gpu, pk_ref):
    assert pk_gpu["patterns"] == pk_ref["patterns"]
    assert np.array_equal(pk_gpu["counts"], pk_ref["counts"])
    assert np.array_equal(pk_gpu["col_ptr"], pk_ref["col_ptr"])
    assert np.array_equal(pk_gpu["val_ptr"], pk_ref["val_ptr"])
    assert np.array_equal(pk_gpu["cols"], pk_ref["cols"])
    assert np.array_equal(pk_gpu["vals"].view(np.uint32), pk_ref["vals"].view(np.uint32))


def _run(srt, torch, plan, B):
    Bt = torch.from_numpy(B).cuda()
    C = torch.full((plan.m, B.shape[1]), float("nan"), device="cuda")
    plan(C, Bt)
    torch.cuda.synchronize()
    return C.cpu().numpy()


def _bitexact(C, C_ref):
    assert np.array_equal(C.view(np.uint32), C_ref.view(np.uint32)), \
        "max |diff| %g" % np.abs(C - C_ref).max()


MAPPINGS = [(4, "identity"), (4, "orig"), (8, "orig"), (8, "alt"), (8, "identity"), (2, "identity"),
            (6, "identity")]


@pytest.mark.parametrize("M_r,name", MAPPINGS)
def test_cpu_config_full_size(srt, oracle, torch_cuda, M_r, name):
    """2048 x 512 at density 0.2, N(0,1) values, B 512 x 256 U[-1,1], seed 0."""
    so, cb = oracle
    torch = torch_cuda
    _log("case cpu-config", M_r, name)
    A, B = srt.workloads.cpu_case(0)
    plan = _plan(srt, torch, A, B.shape[1], M_r, name)
    table = _oracle_table(so, M_r, name)
    pk = cb.pack_dense(A, M_r, table)
    got = plan.packed()
    _check_packed(got, pk)
    st = plan.stats()
    assert st["padding"] == pk["padding"] and st["nnz"] == np.count_nonzero(A)
    C = _run(srt, torch, plan, B)
    _bitexact(C, cb.exec_packed(pk, B, nsplit=plan.split_counts()))
    C64 = cb.dense_f64(A, B)
    assert np.all(np.abs(C - C64) <= so.fp64_error_bound(A, B))


GPU_CASES = [(2048, 512, 0.90, 512), (512, 512, 0.95, 128), (512, 2048, 0.80, 256),
             (1024, 4096, 0.70, 512), (2048, 512, 0.70, 128)]


@pytest.mark.parametrize("case", GPU_CASES, ids=lambda c: "%dx%d_s%d_N%d" % (c[0], c[1], c[2] * 100, c[3]))
@pytest.mark.parametrize("M_r,name", [(8, "orig"), (8, "identity"), (4, "identity")])
def test_pruned_ffn_cases(srt, oracle, torch_cuda, case, M_r, name):
    so, cb = oracle
    torch = torch_cuda
    _log("case", case, M_r, name)
    A, B = srt.workloads.case_inputs(srt.workloads.SpMMCase(*case), seed=0)
    plan = _plan(srt, torch, A, case[3], M_r, name)
    pk = cb.pack_dense(A, M_r, _oracle_table(so, M_r, name))
    _check_packed(plan.packed(), pk)
    C = _run(srt, torch, plan, B)
    _bitexact(C, cb.exec_packed(pk, B, nsplit=plan.split_counts()))
    assert np.all(np.abs(C - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B))


@pytest.mark.parametrize("V", [1, 2, 4])
@pytest.mark.parametrize("U", [1, 2, 4, 8])
@pytest.mark.parametrize("waves,nsplit,chunk,xcd", [(1, 1, 0, 0), (4, 1, 0, 1), (8, 8, 16, 1),
                                                    (16, 16, 8, 1), (8, 4, 32, 0), (16, 2, 0, 1),
                                                    (2, 2, 0, 1)])
@pytest.mark.parametrize("staging", [1, 2])
@pytest.mark.parametrize("layout", [1, 2])
def test_launch_variants_bitexact(srt, oracle, torch_cuda, V, U, waves, nsplit, chunk, xcd, staging, layout):
    so, cb = oracle
    torch = torch_cuda
    _log("variant V=%d U=%d waves=%d nsplit=%d chunk=%d xcd=%d staging=%d layout=%d"
         % (V, U, waves, nsplit, chunk, xcd, staging, layout))
    A, B = srt.workloads.case_inputs(srt.workloads.SpMMCase(520, 300, 0.85, 264), seed=3)
    plan = _plan(srt, torch, A, 264, 8, "orig", vec_width=V, unroll=U, waves_per_block=waves,
                 xcd_remap=xcd, nsplit=nsplit, chunk_cols=chunk, b_staging=staging, layout=layout)
    cfg = plan.get_config()
    assert cfg.b_staging == staging and cfg.layout == layout
    assert cfg.unroll == U and cfg.nsplit <= cfg.waves_per_block
    assert cfg.N_c == 64 * cfg.vec_width and cfg.M_c == 8 * cfg.waves_per_block
    kp = plan.split_counts()
    assert kp.min() >= 1 and kp.max() <= cfg.nsplit
    pk = cb.pack_dense(A, 8, _oracle_table(so, 8, "orig"))
    _bitexact(_run(srt, torch, plan, B), cb.exec_packed(pk, B, nsplit=kp))


@pytest.mark.parametrize("K,V,waves,nsplit", [(640, 1, 16, 16), (320, 2, 16, 4), (160, 4, 8, 8),
                                               (641, 1, 8, 2), (64, 4, 16, 16)])
def test_lds_tile_limits(srt, oracle, torch_cuda, K, V, waves, nsplit):
    """B tiles up to the full 160 KiB of LDS a workgroup can hold (and one row
    past it, which falls back to global-memory B rows); the chunk reduction
    reuses the tile's LDS."""
    so, cb = oracle
    torch = torch_cuda
    _log("lds limit K=%d V=%d waves=%d nsplit=%d" % (K, V, waves, nsplit))
    A, B = srt.workloads.case_inputs(srt.workloads.SpMMCase(1030, K, 0.9, 64 * V * 3 - 4 * V), seed=K)
    N = B.shape[1]
    plan = _plan(srt, torch, A, N, 8, "orig", vec_width=V, waves_per_block=waves, nsplit=nsplit,
                 chunk_cols=8, b_staging=1)
    cfg = plan.get_config()
    _log("realised b_staging=%d V=%d waves=%d" % (cfg.b_staging, cfg.vec_width, cfg.waves_per_block))
    if K * 64 * 4 > 160 * 1024:
        assert cfg.b_staging == 2
    pk = cb.pack_dense(A, 8, _oracle_table(so, 8, "orig"))
    _bitexact(_run(srt, torch, plan, B), cb.exec_packed(pk, B, nsplit=plan.split_counts()))


@pytest.mark.parametrize("nsplit,chunk", [(1, 0), (2, 64), (4, 32), (8, 16), (16, 8)])
@pytest.mark.parametrize("M_r,name", [(8, "identity"), (4, "orig")])
@pytest.mark.parametrize("layout", [1, 2])
def test_split_chains_on_uneven_rows(srt, oracle, torch_cuda, nsplit, chunk, M_r, name, layout):
    """Heavy and empty panels (magnitude pruning with lognormal row scales),
    load-balanced splits, against the oracle's chunked chains."""
    so, cb = oracle
    torch = torch_cuda
    _log("split", nsplit, chunk, M_r, name, layout)
    A, B = srt.workloads.case_inputs(srt.workloads.SpMMCase(2048, 512, 0.9, 512), seed=0)
    plan = _plan(srt, torch, A, 512, M_r, name, nsplit=nsplit, waves_per_block=16, chunk_cols=chunk,
                 vec_width=1, layout=layout)
    kp = plan.split_counts()
    assert kp.max() <= nsplit
    if nsplit > 1 and layout == 1:
        assert kp.max() > kp.min()          # heavy panels get more waves
    pk = cb.pack_dense(A, M_r, _oracle_table(so, M_r, name))
    C = _run(srt, torch, plan, B)
    _bitexact(C, cb.exec_packed(pk, B, nsplit=kp))
    assert np.all(np.abs(C - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B))


@pytest.mark.parametrize("shape", [(37, 50, 24), (1, 1, 1), (9, 3, 7), (1000, 300, 100), (64, 64, 65),
                                   (130, 257, 3)])
@pytest.mark.parametrize("index_dtype", [np.int32, np.int64])
def test_ragged_shapes(srt, oracle, torch_cuda, shape, index_dtype):
    """M not a multiple of M_r, N not a multiple of the lane width, tiny K."""
    so, cb = oracle
    torch = torch_cuda
    M, K, N = shape
    _log("ragged", shape, index_dtype.__name__)
    rng = np.random.default_rng(M * 1000 + K)
    A = ((rng.random((M, K)) < 0.3) * rng.standard_normal((M, K))).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    for M_r, name in [(8, "orig"), (4, "orig")]:
        plan = _plan(srt, torch, A, N, M_r, name, index_dtype=index_dtype)
        pk = cb.pack_dense(A, M_r, _oracle_table(so, M_r, name))
        _check_packed(plan.packed(), pk)
        _bitexact(_run(srt, torch, plan, B), cb.exec_packed(pk, B, nsplit=plan.split_counts()))


def test_empty_and_sparse_rows(srt, oracle, torch_cuda):
    so, cb = oracle
    torch = torch_cuda
    _log("empty matrix / empty rows")
    A = np.zeros((40, 30), dtype=np.float32)
    B = np.ones((30, 16), dtype=np.float32)
    plan = _plan(srt, torch, A, 16, 8, "orig")
    assert plan.stats()["packed_cols"] == 0
    assert np.array_equal(_run(srt, torch, plan, B), np.zeros((40, 16), np.float32))
    A[17, 3] = 2.0
    A[39, 29] = -1.0
    for ns in (1, 8):
        plan = _plan(srt, torch, A, 16, 8, "orig", nsplit=ns, waves_per_block=8, chunk_cols=1)
        _bitexact(_run(srt, torch, plan, B),
                  cb.exec_packed(cb.pack_dense(A, 8, _oracle_table(so, 8, "orig")), B,
                                 nsplit=plan.split_counts()))


def test_balanced_layout_splits_panels_beside_empty_ones(srt, oracle, torch_cuda):
    """Rows emptied in whole 16-row blocks (a hybrid plan's dense panels, dead
    rows of a pruned matrix) under the balanced layout: the empty panels
    spread over the workgroups instead of filling one, so every non-empty
    panel still gets its nsplit waves (an unsplit panel is a chain of
    thousands of columns on one wave, several times the launch's time), and
    the result is the oracle's chain over that split."""
    so, cb = oracle
    torch = torch_cuda
    _log("balanced layout with empty panels")
    rng = np.random.default_rng(41)
    M, K, N = 1024, 1024, 512
    A = ((rng.random((M, K)) < 0.3) * rng.standard_normal((M, K))).astype(np.float32)
    for g in (0, 5, 6, 13, 21, 22, 30, 41, 50, 57, 63):
        A[16 * g:16 * (g + 1)] = 0.0
    B = srt.workloads.dense_b(K, N, seed=42)
    plan = _plan(srt, torch, A, N, 4, "identity", vec_width=4, waves_per_block=16, nsplit=8, chunk_cols=100,
                 layout=2)
    kp = plan.split_counts()
    ncols = np.diff(plan.packed()["col_ptr"])
    assert (ncols[ncols == 0].size == 44) and (kp[ncols > 0] == 8).all(), np.bincount(kp[ncols > 0])
    _bitexact(_run(srt, torch, plan, B),
              cb.exec_packed(cb.pack_dense(A, 4, _oracle_table(so, 4, "identity")), B, nsplit=kp))
    plan.close()


def test_strided_b_c_and_stream(srt, oracle, torch_cuda):
    """B and C as row-strided views; execution on a side stream; n < B width."""
    so, cb = oracle
    torch = torch_cuda
    _log("strided views + stream")
    A, Bfull = srt.workloads.case_inputs(srt.workloads.SpMMCase(256, 128, 0.8, 200), seed=5)
    plan = _plan(srt, torch, A, 200, 8, "orig")
    Bbig = torch.zeros((128, 260), device="cuda")
    Bbig[:, 4:204] = torch.from_numpy(Bfull).cuda()
    Bv = Bbig[:, 4:204]
    Cbig = torch.zeros((256, 300), device="cuda")
    Cv = Cbig[:, 8:208]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan(Cv, Bv, n=136)
    s.synchronize()
    C = Cv.cpu().numpy()
    ref = cb.exec_packed(cb.pack_dense(A, 8, _oracle_table(so, 8, "orig")), Bfull[:, :136],
                         nsplit=plan.split_counts())
    _bitexact(np.ascontiguousarray(C[:, :136]), ref)
    assert np.all(C[:, 136:] == 0)
    assert torch.all(Cbig[:, :8] == 0) and torch.all(Cbig[:, 208:] == 0)


def test_invalid_inputs_raise(srt, torch_cuda):
    torch = torch_cuda
    _log("invalid inputs")
    rp = torch.tensor([0, 1, 2], dtype=torch.int64)
    ci = torch.tensor([0, 7], dtype=torch.int64)         # 7 >= k = 4
    v = torch.tensor([1.0, 2.0])
    with pytest.raises(srt.SrtError):
        srt.MatMulSpecialized(2, 4, 8, v, rp, ci, mapping_id=srt.named_mapping(8, "orig"))
    rp_bad = torch.tensor([0, 2, 1], dtype=torch.int64)  # decreasing row pointer
    with pytest.raises(srt.SrtError):
        srt.MatMulSpecialized(2, 4, 8, v, rp_bad, torch.tensor([0, 1]), mapping_id=srt.named_mapping(8, "orig"))
    with pytest.raises(srt.SrtError):
        srt.MatMulSpecialized(2, 4, 8, v, rp, torch.tensor([0, 1]), executor_id="cdna4_M4_36ed2acc48",
                              mapping_id=srt.named_mapping(8, "orig"))


# --------------------------------------------------------------- SOPTile
def _random_tile(shape, sparsity, seed):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) <= (1 - sparsity)).astype(np.float32) * rng.standard_normal(shape).astype(np.float32)


@pytest.mark.parametrize("M_r,name", [(4, "identity"), (4, "orig"), (8, "orig"), (8, "alt")])
@pytest.mark.parametrize("sparsity", [0.7, 0.9, 0.95])
def test_soptile_matches_reference_packing(srt, oracle, torch_cuda, M_r, name, sparsity):
    """48 x 256 tiles as sbench/SOP/sop_bench.py:171-181 draws them; packed
    panels equal the literal restatement of pack_panel; C = tile . B."""
    so, cb = oracle
    torch = torch_cuda
    _log("soptile", M_r, name, sparsity)
    tile = _random_tile((48, 256), sparsity, int(sparsity * 100) + M_r)
    B = np.random.default_rng(1).uniform(-1, 1, (256, 128)).astype(np.float32)
    m = srt.named_mapping(M_r, name)
    module = srt.make_sop_module(srt.Acc(M_r, 2), m.patterns, m)
    st = module.make_sop_tile(torch.from_numpy(tile))
    table = _oracle_table(so, M_r, name)
    pats = so.kernel_pattern_table(table)
    panels = so.make_sop_tile(tile, M_r, so.table_mapping(table), pats)
    got = st.packed()
    for p, pd in enumerate(panels):
        assert np.array_equal(got["counts"][p], pd["nkern_counts"])
        assert np.array_equal(got["cols"][got["col_ptr"][p]:got["col_ptr"][p + 1]], pd["col_indices"])
        assert np.array_equal(got["vals"][got["val_ptr"][p]:got["val_ptr"][p + 1]], pd["values"])
    assert st.padding() == sum(pd["padding"] for pd in panels)
    # the reference's single buffer (template:219-261): offsets, required
    # bytes and every segment's content
    buf = st.buffer()
    offs, end = so.pack_tile_offsets(panels)
    assert buf["required_bytes"] == so.pack_tile_required_bytes(panels)
    assert buf["used_bytes"] == end and np.array_equal(buf["offsets"], np.array(offs, dtype=np.int64))
    data = buf["data"]
    for (v, c, n), pd in zip(offs, panels):
        assert v % 64 == 0 and c % 64 == 0 and n % 64 == 0
        assert np.array_equal(data[v:v + 4 * pd["num_nnz"]].view(np.float32), pd["values"])
        assert np.array_equal(data[c:c + 4 * pd["num_col_indices"]].view(np.int32), pd["col_indices"])
        assert np.array_equal(data[n:n + 4 * pd["num_nkern"]].view(np.int32), pd["nkern_counts"])
    pk = cb.pack_dense(tile, M_r, table)
    for nsplit in (1, 4):
        ns, C = module.execute_tile(st, torch.from_numpy(B).cuda(), num_runs=2, measured_iterations=3,
                                    tile_config=srt.TileConfig(nsplit=nsplit, waves_per_block=8,
                                                               chunk_cols=8))
        assert ns > 0
        C = C.cpu().numpy()
        _bitexact(C, cb.exec_packed(pk, B, nsplit=st.split_counts()))
    assert torch.allclose(torch.from_numpy(C), torch.from_numpy(tile) @ torch.from_numpy(B), atol=1e-4)


def test_soptile_col_offset_and_errors(srt, oracle, torch_cuda):
    so, cb = oracle
    torch = torch_cuda
    _log("soptile col_offset / errors")
    m = srt.named_mapping(8, "orig")
    t = srt.SOPTile(8, 20, 1, m)
    dense = torch.zeros((10, 8))
    dense[2, 0] = 1.5
    dense[2, 1] = 2.5
    dense[7, 5] = -1.0
    # column 2 has code 0b11 -> 195 under 400fa; column 7 code 0b100000 -> 32
    t.pack_panel(0, 10, [195, 32], [2, 7], dense)
    t.pack_tile()
    got = t.packed()
    assert list(got["cols"]) == [17, 12]             # id(32) < id(195)
    assert list(got["vals"]) == [-1.0, 1.5, 2.5, 0.0, 0.0]
    assert t.padding() == 2
    B = torch.randn(20, 64, device="cuda")
    ns, C = srt.execute_tile(64, t, B, num_runs=0)
    ref = torch.zeros(8, 64, device="cuda")
    ref[0] = 1.5 * B[12]
    ref[1] = 2.5 * B[12]
    ref[5] = -1.0 * B[17]
    assert torch.allclose(C, ref)
    t2 = srt.SOPTile(8, 20, 1, m)
    with pytest.raises(srt.SrtError):                # pattern 3 is not in the set
        t2.pack_panel(0, 0, [3], [2], dense)
    with pytest.raises(srt.SrtError):                # entries do not cover the panel
        t2.pack_panel(0, 0, [195], [2], dense)
    with pytest.raises(srt.SrtError):                # col_offset + K beyond the tile
        t2.pack_panel(0, 15, [195, 32], [2, 7], dense)
    with pytest.raises(srt.SrtError):                # pack_tile before every panel
        t2.pack_tile()


def test_sort_and_encode_patterns(srt, oracle, torch_cuda):
    so, cb = oracle
    torch = torch_cuda
    _log("sort_and_encode_patterns")
    m = srt.named_mapping(8, "orig")
    rng = np.random.default_rng(2)
    codes = rng.choice(m.patterns + [0, 0, 0], size=1000).astype(np.int32)
    enc, order = srt.sort_and_encode_patterns(torch.from_numpy(codes), m)
    e_ref, o_ref = so.sort_and_encode_patterns(codes.tolist(), m.patterns)
    assert enc.cpu().tolist() == e_ref
    assert order.cpu().tolist() == o_ref


# ------------------------------------------------------------------- CSR
@pytest.mark.parametrize("n_tile", [16, 64, 128, 256])
@pytest.mark.parametrize("index_dtype", ["int32", "int64"])
def test_csr_bitexact(srt, oracle, torch_cuda, n_tile, index_dtype):
    so, cb = oracle
    torch = torch_cuda
    _log("csr", n_tile, index_dtype)
    A, B = srt.workloads.case_inputs(srt.workloads.SpMMCase(2048, 512, 0.9, 512), seed=0)
    csr = torch.from_numpy(A).to_sparse_csr()
    if index_dtype == "int32":
        csr = torch.sparse_csr_tensor(csr.crow_indices().int(), csr.col_indices().int(), csr.values(), A.shape)
    ns, C = srt.csr.executor(csr, torch.from_numpy(B).cuda(), num_runs=2, n_tile=n_tile,
                             measured_iterations=3)
    rp, ci, v = srt.workloads.to_csr(A)
    C_ref = cb.csr_exec(rp, ci, v, A.shape[0], B)
    _bitexact(C.cpu().numpy(), C_ref)
    for N in (1, 3, 100):
        C2 = srt.csr.spmm(csr, torch.from_numpy(B[:, :N].copy()).cuda(), n_tile=n_tile)
        _bitexact(C2.cpu().numpy(), cb.csr_exec(rp, ci, v, A.shape[0], B[:, :N].copy()))


def test_csr_out_of_range_indices_are_skipped(srt, torch_cuda):
    torch = torch_cuda
    _log("csr malformed")
    rp = torch.tensor([0, 2, 10], dtype=torch.int32, device="cuda")  # row 1 runs past nnz = 3
    ci = torch.tensor([0, 99, 1], dtype=torch.int32, device="cuda")  # 99 >= k
    v = torch.tensor([1.0, 5.0, 2.0], device="cuda")
    B = torch.arange(8, dtype=torch.float32, device="cuda").reshape(2, 4)
    C = torch.empty((2, 4), device="cuda")
    L = srt._lib.lib()
    from ctypes import c_void_p
    rc = L.srt_csr_spmm(2, 2, 4, 3, c_void_p(rp.data_ptr()), c_void_p(ci.data_ptr()), 4,
                        c_void_p(v.data_ptr()), c_void_p(B.data_ptr()), 4, c_void_p(C.data_ptr()), 4, 64,
                        None)
    assert rc == 0
    torch.cuda.synchronize()
    assert torch.equal(C[0], B[0])                   # 1*B[0] + (skipped 99)
    assert torch.equal(C[1], 2.0 * B[1])             # entry 2 only (clamped to nnz)
