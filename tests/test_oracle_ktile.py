"""K-tiled packing in the oracle (count_nkerns.cpp:52-126: A cut into K_c
column tiles, each panel packed per tile, load_c chaining the tiles).

Pinned by what the reference's execution fixes: running the tiles one after
another with load_c (C of tile t is the starting accumulator of tile t + 1)
is bit-identical to one chain over the concatenated packing; every tile's
columns stay inside its K range; the packing is the untiled one when K_c >= K;
and the result stays within the FP64 bound."""
import numpy as np
import pytest


@pytest.mark.parametrize("k_tile", [64, 100, 128, 255])
@pytest.mark.parametrize("M_r,name", [(4, "orig"), (8, "orig")])
def test_ktiled_chain_equals_tile_by_tile_load_c(srt, oracle, k_tile, M_r, name):
    so, cb = oracle
    A, B = srt.workloads.case_inputs(srt.workloads.SpMMCase(96, 300, 0.8, 40), seed=4)
    table = srt.named_mapping(M_r, name).table
    pk = cb.pack_dense_ktiled(A, M_r, table, k_tile)
    KT = -(-300 // k_tile)
    assert pk["KT"] == KT and pk["counts"].shape == (pk["P"], KT, pk["NK"])
    C = cb.exec_packed(pk, B)
    # tile by tile with load_c: the reference's K-tiled execution
    Ct = np.zeros_like(C)
    for t, k0 in enumerate(range(0, 300, k_tile)):
        sl = cb.pack_dense(A[:, k0:k0 + k_tile], M_r, table, col_offset=k0)
        assert sl["cols"].size == 0 or (sl["cols"].min() >= k0 and sl["cols"].max() < k0 + k_tile)
        assert np.array_equal(pk["counts"][:, t, :], sl["counts"])
        Ct = cb.exec_packed(sl, B, C=Ct, load_c=t > 0)
    assert np.array_equal(C.view(np.uint32), Ct.view(np.uint32))
    assert np.all(np.abs(C - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B))
    assert pk["padding"] == sum(cb.pack_dense(A[:, k0:k0 + k_tile], M_r, table)["padding"]
                                for k0 in range(0, 300, k_tile))


def test_ktiled_degenerates_to_untiled(srt, oracle):
    _, cb = oracle
    A, B = srt.workloads.case_inputs(srt.workloads.SpMMCase(40, 64, 0.7, 16), seed=1)
    table = srt.named_mapping(4, "orig").table
    a, b = cb.pack_dense_ktiled(A, 4, table, 64), cb.pack_dense(A, 4, table)
    assert np.array_equal(a["cols"], b["cols"]) and np.array_equal(a["vals"].view(np.uint32), b["vals"].view(np.uint32))
    # a planted mistake: dropping the load_c of the later tiles loses tile 0
    pk = cb.pack_dense_ktiled(A, 4, table, 32)
    C = cb.exec_packed(pk, B)
    sl1 = cb.pack_dense(A[:, 32:], 4, table, col_offset=32)
    wrong = cb.exec_packed(sl1, B)
    assert not np.array_equal(C, wrong)
