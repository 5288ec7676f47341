"""The shipped tuning table: every pruned-FFN workload has a record, and the
plan built from it (autotune.tuned_plan) realises the recorded configuration
and stays bit-exact with the oracle (K-tiled packing included)."""
import numpy as np
import pytest

from test_gpu_parity import _run, _bitexact

pytestmark = pytest.mark.gpu


def test_table_covers_every_workload(srt):
    at = srt.autotune
    cache = at.TuningCache(at.SHIPPED)
    for case in srt.workloads.gpu_cases():
        recs = cache.get(case.M, case.K, case.sparsity, case.N)
        assert recs, case
        assert recs == sorted(recs, key=lambda r: r["seconds"])


@pytest.mark.parametrize("case", [(2048, 512, 0.9, 512), (512, 2048, 0.7, 128), (1024, 4096, 0.95, 256)])
def test_tuned_plan_exact(srt, oracle, case):
    import torch
    so, cb = oracle
    at = srt.autotune
    c = srt.workloads.SpMMCase(*case)
    A, B = srt.workloads.case_inputs(c, seed=0)
    csr = torch.from_numpy(A).to_sparse_csr()
    plan = at.tuned_plan(csr, c.N)
    rec = at.TuningCache(at.SHIPPED).get(c.M, c.K, c.sparsity, c.N)[0]
    real = plan.get_config()
    assert plan.mapping.M_r == rec["M_r"] and real.vec_width == rec["vec_width"]
    assert real.K_c == (rec.get("K_c") or c.K)
    pk = cb.pack_dense_ktiled(A, rec["M_r"], plan.mapping.mapping.table, real.K_c)
    _bitexact(_run(srt, torch, plan, B), cb.exec_packed(pk, B, nsplit=plan.split_counts()))
    plan.close()
