"""GPU parity of an executor generated at run time for a pattern set chosen
by the set-cover model for one matrix (pattern_select.specialized_mapping
-> codegen -> jit): bit-identical to the oracle's fmaf chain over the same
packed data and within the FP64 bound."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_specialized_mapping_executor(srt, oracle):
    import torch
    so, cb = oracle
    A = srt.workloads.magnitude_pruned(512, 256, 0.85, seed=11)
    m, cost = srt.pattern_select.specialized_mapping(A, 4, max_groups=6)
    print("specialized", m, "cost %.1f" % cost, flush=True)
    sys.stdout.flush()
    native = srt.NativeMapping.from_mapping(m)
    from importlib import import_module
    jit = import_module("sparse-register-tiling_amd.jit")
    eid = jit.ensure_executor(native, lean=True)
    B = srt.workloads.dense_b(256, 192, seed=2)
    plan = srt.MatMulSpecialized.from_dense(torch.from_numpy(A), 192, mapping_id=native,
                                            tile_config=srt.TileConfig(b_staging=1))
    C = torch.full((512, 192), float("nan"), device="cuda")
    plan(C, torch.from_numpy(B).cuda())
    torch.cuda.synchronize()
    C = C.cpu().numpy()
    pk = cb.pack_dense(A, 4, m.table, kernel_list=m.kernel_list)
    assert plan.stats()["padding"] == pk["padding"]
    ref = cb.exec_packed(pk, B, nsplit=plan.split_counts())
    assert np.array_equal(C.view(np.uint32), ref.view(np.uint32))
    assert np.all(np.abs(C - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B))
    plan.close()
