"""GPU parity on the reference's bundled random MatrixMarket matrices
(768 x 768 at 98 % and 99 % sparsity), read through matrix_io, packed and
executed through the C ABI: bit-identical to the oracle's fmaf chain over
the same packed data and within the FP64 summation bound."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN
from test_gpu_parity import _oracle_table

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["random_98_768_768.mtx", "random_99_768_768.mtx"])
@pytest.mark.parametrize("M_r,mapping", [(4, "orig"), (8, "orig")])
def test_mtx_fixture_spmm(srt, oracle, name, M_r, mapping):
    import torch
    so, cb = oracle
    print("mtx", name, M_r, mapping, flush=True)
    sys.stdout.flush()
    m = srt.matrix_io.read_mtx(os.path.join(GOLDEN, "mtx", name))
    A = m.to_dense()
    B = srt.workloads.dense_b(m.cols, 256, seed=4)
    plan = srt.MatMulSpecialized(m.rows, m.cols, 256, torch.from_numpy(m.Lx), torch.from_numpy(m.Lp),
                                 torch.from_numpy(m.Li), mapping_id=srt.named_mapping(M_r, mapping))
    C = torch.full((m.rows, 256), float("nan"), device="cuda")
    plan(C, torch.from_numpy(B).cuda())
    torch.cuda.synchronize()
    C = C.cpu().numpy()
    pk = cb.pack_dense(A, M_r, _oracle_table(so, M_r, mapping))
    ref = cb.exec_packed(pk, B, nsplit=plan.split_counts())
    assert np.array_equal(C.view(np.uint32), ref.view(np.uint32))
    assert np.all(np.abs(C - cb.dense_f64(A, B)) <= so.fp64_error_bound(A, B))
    plan.close()
