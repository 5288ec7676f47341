"""The device packer's memory follows the nonzeros, not the dense size.

A 4096 x 262144 matrix at 99.9 % sparsity (1.07 M nonzeros; its dense form
would be 4 GiB) is packed from CSR on the GPU, with the column indices of
half the rows shuffled (the packer sorts entries itself).  The packed arrays
and C = A . B are compared bit for bit with the oracle, which packs the same
rows as dense slices of 256 rows."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _big_csr(M, K, density, seed):
    rng = np.random.default_rng(seed)
    n = int(M * K * density)
    key = np.unique(rng.integers(0, M, n).astype(np.int64) * K + rng.integers(0, K, n))
    rows, cols = key // K, key % K
    # odd rows get their column indices in random order (the packer sorts)
    order = np.lexsort((np.where(rows % 2 == 1, rng.random(len(key)), cols / K), rows))
    rows, cols = rows[order], cols[order]
    rp = np.zeros(M + 1, dtype=np.int64)
    rp[1:] = np.cumsum(np.bincount(rows, minlength=M))
    v = rng.standard_normal(len(cols)).astype(np.float32)
    return rp, cols.astype(np.int64), v


def _dense_rows(rp, ci, v, r0, r1, K):
    A = np.zeros((r1 - r0, K), dtype=np.float32)
    for i in range(r0, r1):
        A[i - r0, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    return A


def test_pack_4096x262144_sparse(srt, oracle):
    import torch
    so, cb = oracle
    M, K, N = 4096, 262144, 64
    rp, ci, v = _big_csr(M, K, 0.001, seed=1)
    print("pack 4096x262144, nnz %d" % len(v), flush=True)
    plan = srt.MatMulSpecialized(M, K, N, torch.from_numpy(v), torch.from_numpy(rp), torch.from_numpy(ci),
                                 mapping_id=srt.named_mapping(4, "orig"))
    st = plan.stats()
    assert st["nnz"] == len(v)
    # packed storage ~ nonzeros: far below the dense size
    assert plan.require_storage < 64 * len(v) + 64 * st["num_panels"]
    got = plan.packed()
    B = np.random.default_rng(2).uniform(-1, 1, (K, N)).astype(np.float32)
    Bd = torch.from_numpy(B).cuda()
    C = torch.full((M, N), float("nan"), device="cuda")
    plan(C, Bd)
    torch.cuda.synchronize()
    C = C.cpu().numpy()
    kp = plan.split_counts()
    table = so.named_table(4, "orig")
    rows = 256
    for r0 in range(0, M, rows):
        A = _dense_rows(rp, ci, v, r0, r0 + rows, K)
        pk = cb.pack_dense(A, 4, table)
        p0, p1 = r0 // 4, (r0 + rows) // 4
        assert np.array_equal(got["counts"][p0:p1], pk["counts"]), r0
        cs, ce = got["col_ptr"][p0], got["col_ptr"][p1]
        vs, ve = got["val_ptr"][p0], got["val_ptr"][p1]
        assert np.array_equal(got["col_ptr"][p0:p1 + 1] - cs, pk["col_ptr"]), r0
        assert np.array_equal(got["val_ptr"][p0:p1 + 1] - vs, pk["val_ptr"]), r0
        assert np.array_equal(got["cols"][cs:ce], pk["cols"]), r0
        assert np.array_equal(got["vals"][vs:ve].view(np.uint32), pk["vals"].view(np.uint32)), r0
        ref = cb.exec_packed(pk, B, nsplit=kp[p0:p1])
        assert np.array_equal(C[r0:r0 + rows].view(np.uint32), ref.view(np.uint32)), r0
    plan.close()


def test_pack_rejects_duplicate_entries(srt):
    import torch
    rp = torch.tensor([0, 3, 4], dtype=torch.int64)
    ci = torch.tensor([5, 1, 5, 2], dtype=torch.int64)      # row 0 holds column 5 twice
    v = torch.tensor([1.0, 2.0, 3.0, 4.0])
    with pytest.raises(srt.SrtError) as e:
        srt.MatMulSpecialized(2, 8, 16, v, rp, ci, mapping_id=srt.named_mapping(4, "orig"))
    assert "twice" in str(e.value)
