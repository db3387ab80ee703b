"""Pins the oracle's whole-path restatement (ordering, factorization, solve, GMRES) by what
the mathematics fixes:

* tol = 0 (skip = nlevels - 1, pspaQR.cpp:150) => no sparsification => the tree solve is the
  exact inverse of the permuted matrix (checked against scipy's sparse direct solver), and
  GMRES converges at its first step (returned k == 0).
* the nested-dissection ordering: separators really separate (no A^T A edge between the
  leaves of different parts), the permutation is a permutation, the cluster hierarchy sums.
* with sparsification (tol > 0) the preconditioned GMRES reaches the 1e-12 true residual.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

from oracle.oracle import OracleTree
from pspaqr_public_amd.problems import make_problem, mt_normal, mt_uniform


@pytest.mark.parametrize("cfg,side", [(1, 32), (2, 10), (3, 12), (5, 12)])
def test_exact_inverse_when_tol_zero(cfg, side):
    P = make_problem(cfg, side)
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, 0.0, P.nlevels - 1)
    O.factorize()
    cp = O.cperm()
    b = P.b[cp]
    x = O.solve(b)
    App = P.A[cp][:, cp].tocsc()
    xs = spl.spsolve(App, b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-12
    xg, it, hist = O.gmres(b, iters=50, tol=1e-12)
    assert it == 0
    assert np.linalg.norm(b - App @ xg) / np.linalg.norm(b) < 1e-13


@pytest.mark.parametrize("cfg,side", [(1, None), (2, 12), (3, 14)])
def test_ordering_invariants(cfg, side):
    P = make_problem(cfg, side)
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, P.skip)
    cp = O.cperm()
    assert np.array_equal(np.sort(cp), np.arange(P.A.shape[0]))
    cl = O.clusters()
    # leaf clusters (merge level 0) tile the permuted index range
    leaf = cl[cl[:, 0] == 0]
    starts = leaf[:, 9]
    sizes = leaf[:, 10]
    assert starts[0] == 0 and np.array_equal(starts[1:], np.cumsum(sizes)[:-1])
    assert sizes.sum() == P.A.shape[0]
    # parents' sizes are the sum of their children's
    for o in range(cl.shape[0]):
        ch = np.nonzero(cl[:, 11] == o)[0]
        if len(ch):
            assert cl[o, 10] == cl[ch, 10].sum()
    # separators separate: columns of two different level-0 interiors share no row
    owner = np.empty(P.A.shape[0], dtype=np.int64)
    for o in np.nonzero(cl[:, 0] == 0)[0]:
        owner[cl[o, 9]: cl[o, 9] + cl[o, 10]] = o
    is_leaf = (cl[owner, 1] == 0)
    App = P.A[cp][:, cp].tocsr()
    for r in range(App.shape[0]):
        cols = App.indices[App.indptr[r]:App.indptr[r + 1]]
        leaves = set(owner[c] for c in cols if is_leaf[c])
        assert len(leaves) <= 1


@pytest.mark.parametrize("cfg,side", [(1, None), (2, 12), (3, 16)])
def test_sparsified_gmres_reaches_1e12(cfg, side):
    P = make_problem(cfg, side)
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, P.skip)
    O.factorize()
    cp = O.cperm()
    b = P.b[cp]
    x, it, hist = O.gmres(b, iters=200, tol=1e-12)
    App = P.A[cp][:, cp].tocsc()
    assert np.linalg.norm(b - App @ x) / np.linalg.norm(b) <= 1e-12
    assert 0 < it < 60
    # the estimate the loop stops on is monotone (GMRES minimises the residual)
    assert np.all(np.diff(hist) <= 1e-15)


def test_random_streams_match_std_mt19937():
    """First values of std::mt19937(0) + uniform_real_distribution<double>(-1,1) (the
    reference's random(), src/util.cpp:992-1001), computed with libstdc++."""
    u = mt_uniform(0, 3)
    assert np.array_equal(u, np.array([0.18568923303336526, 0.68853148851319657, 0.71589123997965975]))
    z = mt_normal(2, 100000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
