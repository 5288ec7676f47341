"""Matrix file formats: SMTX, MatrixMarket, file lists (no GPU)."""
import os

import numpy as np
import pytest
import scipy.io

from conftest import GOLDEN

MTX = os.path.join(GOLDEN, "mtx")


@pytest.mark.parametrize("name", ["random_98_768_768.mtx", "random_99_768_768.mtx"])
def test_read_mtx_matches_scipy(srt, name):
    path = os.path.join(MTX, name)
    csr = srt.matrix_io.read_mtx(path)
    ref = scipy.io.mmread(path).tocsr()
    ref.sort_indices()
    assert (csr.rows, csr.cols) == ref.shape and csr.nnz == ref.nnz
    assert np.array_equal(csr.Lp, ref.indptr) and np.array_equal(csr.Li, ref.indices)
    assert np.array_equal(csr.Lx, ref.data.astype(np.float32))
    # density encoded in the file name: random_<sparsity %>_...
    sparsity = int(name.split("_")[1]) / 100
    assert abs(csr.nnz / (768 * 768) - (1 - sparsity)) < 0.005


def test_symmetric_and_pattern_mtx(srt, tmp_path):
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n% c\n3 3 3\n1 1 2.0\n3 1 -1.5\n2 2 4\n")
    A = srt.matrix_io.read_mtx(str(p)).to_dense()
    assert np.array_equal(A, np.array([[2, 0, -1.5], [0, 4, 0], [-1.5, 0, 0]], dtype=np.float32))
    q = tmp_path / "p.mtx"
    q.write_text("%%MatrixMarket matrix coordinate pattern general\n2 3 2\n2 3\n1 2\n")
    B = srt.matrix_io.read_mtx(str(q))
    assert np.array_equal(B.to_dense(), np.array([[0, 1, 0], [0, 0, 1]], dtype=np.float32))
    bad = tmp_path / "d.mtx"
    bad.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(ValueError, match="coordinate"):
        srt.matrix_io.read_mtx(str(bad))
    out = tmp_path / "w.mtx"
    csr = srt.matrix_io.read_mtx(str(p))
    srt.matrix_io.write_mtx(str(out), csr)
    assert np.array_equal(srt.matrix_io.read_mtx(str(out)).to_dense(), A)


def test_smtx_read_write(srt, tmp_path):
    # header with commas, row pointers, then column indices on the next line
    p = tmp_path / "a.smtx"
    p.write_text("3, 5, 4\n0 2 2 4\n4 1\n0 3\n")
    csr = srt.matrix_io.read_smtx(str(p))
    assert (csr.rows, csr.cols, csr.nnz) == (3, 5, 4)
    assert csr.Lp.tolist() == [0, 2, 2, 4] and csr.Li.tolist() == [1, 4, 0, 3]  # row 0 sorted
    assert csr.Lx.tolist() == [1.0] * 4
    # the pattern of an MTX fixture survives an SMTX round trip
    m = srt.matrix_io.read_mtx(os.path.join(MTX, "random_99_768_768.mtx"))
    q = tmp_path / "b.smtx"
    srt.matrix_io.write_smtx(str(q), m.rows, m.cols, m.Lp, m.Li)
    back = srt.matrix_io.read_matrix(str(q))
    assert np.array_equal(back.Lp, m.Lp) and np.array_equal(back.Li, m.Li)
    assert np.array_equal(back.to_dense(), (m.to_dense() != 0).astype(np.float32))
    bad = tmp_path / "c.smtx"
    bad.write_text("2, 2, 2\n0 1 3\n0 1\n")
    with pytest.raises(ValueError):
        srt.matrix_io.read_smtx(str(bad))


def test_filelist(srt, tmp_path):
    fl = tmp_path / "list.txt"
    fl.write_text("mobilenet/70/conv1x1_0_3136.smtx,3136\n\ndlmc/rn50/a.smtx\n/abs/b.mtx,49\n")
    got = srt.matrix_io.read_filelist(str(fl), root="/data")
    assert got == [("/data/mobilenet/70/conv1x1_0_3136.smtx", 3136), ("/data/dlmc/rn50/a.smtx", None),
                   ("/abs/b.mtx", 49)]
