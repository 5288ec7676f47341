"""GPU parity of K-tiled packing (TileConfig.K_c < K: every panel packed as
K_c-column tiles, the chain running tile after tile as the reference's
K-tiled execution with load_c, count_nkerns.cpp:52-126): packed arrays equal
the oracle's tile-by-tile packing and C is bit-identical to the oracle's
chain over it, with chunk splits that cross tile boundaries."""
import sys

import numpy as np
import pytest

from test_gpu_parity import _oracle_table, _plan, _run, _bitexact

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k_tile", [64, 100, 128, 299, 300, 1000])
@pytest.mark.parametrize("M_r,name", [(4, "orig"), (8, "orig")])
@pytest.mark.parametrize("nsplit", [1, 8])
def test_ktiled_plan(srt, oracle, k_tile, M_r, name, nsplit):
    import torch
    so, cb = oracle
    print("ktile", k_tile, M_r, name, nsplit, flush=True)
    sys.stdout.flush()
    A, B = srt.workloads.case_inputs(srt.workloads.SpMMCase(520, 300, 0.85, 264), seed=3)
    plan = _plan(srt, torch, A, 264, M_r, name, K_c=k_tile, nsplit=nsplit, waves_per_block=8, chunk_cols=16)
    st = plan.stats()
    KT = -(-300 // k_tile) if k_tile < 300 else 1
    assert st["k_tiles"] == KT and plan.get_config().K_c == min(k_tile, 300)
    pk = cb.pack_dense_ktiled(A, M_r, _oracle_table(so, M_r, name), k_tile)
    got = plan.packed()
    want_counts = pk["counts"] if KT > 1 else pk["counts"].reshape(pk["P"], -1)
    assert np.array_equal(got["counts"], want_counts)
    assert np.array_equal(got["col_ptr"], pk["col_ptr"]) and np.array_equal(got["cols"], pk["cols"])
    assert np.array_equal(got["vals"].view(np.uint32), pk["vals"].view(np.uint32))
    assert st["padding"] == pk["padding"]
    C = _run(srt, torch, plan, B)
    _bitexact(C, cb.exec_packed(pk, B, nsplit=plan.split_counts()))
    assert np.all(np.abs(C - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B))
    plan.close()


def test_reference_default_k_tile_on_the_2048x512_workload(srt, oracle):
    """K_c = 256, the reference's SOPConfig default, on 2048 x 512 (90 %) x 512."""
    import torch
    so, cb = oracle
    A, B = srt.workloads.case_inputs(srt.workloads.PRUNED_2048_512_S90_N512, seed=0)
    plan = _plan(srt, torch, A, 512, 4, "orig", K_c=256)
    assert plan.stats()["k_tiles"] == 2
    pk = cb.pack_dense_ktiled(A, 4, _oracle_table(so, 4, "orig"), 256)
    _bitexact(_run(srt, torch, plan, B), cb.exec_packed(pk, B, nsplit=plan.split_counts()))
    plan.close()


@pytest.mark.parametrize("strategy,n", [(1, 256), (2, 512), (2, 96)])
def test_tiling_strategy_k_tile(srt, oracle, strategy, n):
    """tiling_strategy 1 (CAKE) sizes K_c to a workgroup's LDS B tile;
    2 adds the TLB compensation K_c = entries * page / (n * 4) whenever the
    n x K_c B block needs more pages (SpMM_GECSB.h:339-348).  The packing is
    the K-tiled one, bit-exact."""
    import torch
    so, cb = oracle
    A = srt.workloads.magnitude_pruned(256, 2048, 0.9, seed=3)
    B = srt.workloads.dense_b(2048, n, seed=4)
    plan = _plan(srt, torch, A, n, 4, "orig", tiling_strategy=strategy, vec_width=4,
                 max_tlb_entries=64, tlb_page_size=4096)
    kc = plan.get_config().K_c
    cake = 160 * 1024 // (64 * 4 * 4)          # 160 KiB of LDS / (64 lanes x 4 floats x 4 bytes) = 160 rows
    tlb = 64 * 4096 // (n * 4) if n * cake * 4 // 4096 > 64 else cake
    assert kc == (cake if strategy == 1 else tlb), (strategy, n, kc)
    pk = cb.pack_dense_ktiled(A, 4, _oracle_table(so, 4, "orig"), kc)
    _bitexact(_run(srt, torch, plan, B), cb.exec_packed(pk, B, nsplit=plan.split_counts()))
    plan.close()
