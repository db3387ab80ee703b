"""Matrix Market reader / writer (spaqr_mm_*; the reference's include/mmio.hpp:174-377), on the
CPU, against scipy.io's independent implementation: files written by scipy are read here and
files written here are read by scipy, for every type / property the reference reads
(real, integer, pattern; general, symmetric, skew-symmetric), with comments, blank lines and
duplicate entries (summed, as Eigen's setFromTriplets does); array files for coordinates and
vectors; and every error code.
"""
import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

from pspaqr_public_amd import mmio


def _rand(rng, m, n, density=0.2):
    return sp.random(m, n, density=density, random_state=rng, format="csc", data_rvs=rng.standard_normal)


def test_general_real_round_trips(tmp_path):
    rng = np.random.default_rng(1)
    A = _rand(rng, 37, 23)
    f = tmp_path / "a.mtx"
    sio.mmwrite(str(f), A, precision=17)
    B = mmio.read_sparse(f)
    assert B.shape == A.shape and (B != A).nnz == 0
    assert np.all(np.diff(B.indptr) >= 0)
    for j in range(B.shape[1]):  # rows sorted inside each column
        assert np.all(np.diff(B.indices[B.indptr[j]:B.indptr[j + 1]]) > 0)
    g = tmp_path / "b.mtx"
    mmio.write_sparse(g, A)
    C = sp.csc_matrix(sio.mmread(str(g)))
    assert (C != A).nnz == 0  # 17 digits: the same doubles


def test_symmetric_skew_pattern_integer(tmp_path):
    rng = np.random.default_rng(2)
    L = sp.tril(_rand(rng, 30, 30, 0.3), k=-1)
    D = sp.diags(rng.standard_normal(30))
    S = (L + L.T + D).tocsc()
    f = tmp_path / "s.mtx"
    sio.mmwrite(str(f), S, symmetry="symmetric", precision=17)
    assert "symmetric" in f.read_text().splitlines()[0]
    assert (mmio.read_sparse(f) != S).nnz == 0
    K = (L - L.T).tocsc()
    g = tmp_path / "k.mtx"
    sio.mmwrite(str(g), K, symmetry="skew-symmetric", precision=17)
    assert (mmio.read_sparse(g) != K).nnz == 0
    P = (abs(L) > 0).astype(np.float64).tocsc()
    h = tmp_path / "p.mtx"
    sio.mmwrite(str(h), P, field="pattern")
    assert "pattern" in h.read_text().splitlines()[0]
    assert (mmio.read_sparse(h) != P).nnz == 0
    I = sp.csc_matrix(np.array([[1, 0, 3], [0, -2, 0]], dtype=np.int64))
    i = tmp_path / "i.mtx"
    sio.mmwrite(str(i), I, field="integer")
    assert (mmio.read_sparse(i) != I.astype(np.float64)).nnz == 0
    # the writer's symmetric form holds the lower triangle and reads back whole
    w = tmp_path / "w.mtx"
    mmio.write_sparse(w, S, symmetric=True)
    assert (sp.csc_matrix(sio.mmread(str(w))) != S).nnz == 0 and (mmio.read_sparse(w) != S).nnz == 0


def test_comments_blank_lines_and_duplicates(tmp_path):
    f = tmp_path / "d.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n% a comment\n\n3 2 4\n1 1 1.5\n"
                 "% inside\n3 2 -2\n1 1 0.25\n\n2 1 4e-3\n")
    A = mmio.read_sparse(f)
    assert A.shape == (3, 2) and A.nnz == 3
    assert A[0, 0] == 1.75 and A[2, 1] == -2.0 and A[1, 0] == 4e-3


def test_dense_and_vector(tmp_path):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((3, 50))
    f = tmp_path / "x.mtx"
    sio.mmwrite(str(f), X, precision=17)
    assert np.array_equal(mmio.read_dense(f), X)
    g = tmp_path / "y.mtx"
    mmio.write_dense(g, X)
    assert np.array_equal(sio.mmread(str(g)), X)
    b = rng.standard_normal(17)
    h = tmp_path / "b.mtx"
    mmio.write_dense(h, b)
    assert np.array_equal(mmio.read_vector(h), b)
    with pytest.raises(ValueError):
        mmio.read_vector(g)


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", -2),
    ("%%MatrixMarket vector coordinate real general\n1 1 1\n1 1 1\n", -2),
    ("garbage\n", -2),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", -3),
    ("%%MatrixMarket matrix coordinate real symmetric\n2 2 1\n1 2 1.0\n", -3),
    ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n1 1 1.0\n", -3),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n", -4),
    ("%%MatrixMarket matrix coordinate real general\n2 2\n", -3),
])
def test_error_codes(tmp_path, text, code):
    from pspaqr_public_amd._lib import lib

    f = tmp_path / "e.mtx"
    f.write_text(text)
    nr, nc = np.zeros(1, np.int32), np.zeros(1, np.int32)
    nnz = np.zeros(1, np.int64)
    assert lib().spaqr_mm_read_sparse(str(f).encode(), nr.ctypes.data, nc.ctypes.data, nnz.ctypes.data, None, None,
                                      None) == code
    assert lib().spaqr_last_error(None)
    with pytest.raises(OSError):
        mmio.read_sparse(f)


def test_missing_file_and_array_errors(tmp_path):
    from pspaqr_public_amd._lib import lib

    assert lib().spaqr_mm_read_sparse(str(tmp_path / "none.mtx").encode(), None, None, None, None, None, None) == -1
    f = tmp_path / "a.mtx"
    f.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n")
    assert lib().spaqr_mm_read_dense(str(f).encode(), None, None, None) == -4
    f.write_text("%%MatrixMarket matrix array real symmetric\n2 2\n1\n2\n3\n")
    assert lib().spaqr_mm_read_dense(str(f).encode(), None, None, None) == -2
    f.write_text("%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n")
    assert lib().spaqr_mm_read_dense(str(f).encode(), None, None, None) == -2  # not an array file
