"""The MI355X path on matrices without coordinates (graph-bisection or caller-given leaf parts,
the reference's METIS path) and read from Matrix Market files (the driver's input path,
pspaQR.cpp:85-167).

* Given parts: the GPU factorization equals the oracle's on the same parts -- cluster table
  (ranks included), preconditioner apply within 1e-10, GMRES iteration count (FP64).
* Tiny tolerance on an irregular graph: the sparsified path is an exact inverse, solve(b)
  equal to scipy's direct solve within 1e-10.
* From files: config 1 written to a .mtx file, read back without coordinates (graph ordering)
  and with a coordinate file (geometric ordering, which equals the generator's tensor grid):
  FP32 fronts + flexible GMRES to true relres 1e-12.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle.oracle import OracleTree
from pspaqr_public_amd import mmio
from pspaqr_public_amd.problems import Problem, column_scale, load_problem, make_problem, mt_uniform
from pspaqr_public_amd.ptree import build_tree, linspace_nd

pytestmark = pytest.mark.gpu


def _graph_problem(P):
    return Problem(P.name + "_graph", P.A, P.b, 0, 0, P.nlevels, P.tol, P.skip)


@pytest.mark.parametrize("cfg,side", [(1, None), (3, 12), (2, 12)])
def test_given_parts_factorization_equals_oracle(cfg, side):
    P = _graph_problem(make_problem(cfg, side))
    T = build_tree(P, 64)  # graph bisection
    parts = T.parts()
    O = OracleTree(P.A, 0, 0, P.nlevels, P.tol, P.skip, parts=parts)
    O.factorize()
    Q = Problem(P.name, P.A, P.b, 0, 0, P.nlevels, P.tol, P.skip, parts=parts)
    G = build_tree(Q, 64)  # the same parts given back
    G.factorize()
    assert np.array_equal(G.cperm(), O.cperm())
    assert np.array_equal(G.clusters(), O.clusters())
    b = P.b[O.cperm()]
    xo = O.solve(b)
    assert np.linalg.norm(G.solve(b) - xo) / np.linalg.norm(xo) < 1e-10
    _, ito, _ = O.gmres(b, iters=300, tol=1e-12)
    xg, res, _ = G.gmres(b, iters=300, tol=1e-12)
    print("[graph] %s: levels %d, gmres %d (oracle %d)" % (P.name, P.nlevels, res.iters, ito))
    assert res.iters == ito
    assert np.linalg.norm(b - O.spmv(xg)) / np.linalg.norm(b) <= 1e-12


def test_tiny_tol_graph_ordering_is_exact():
    rng = np.random.default_rng(4)
    n = 3000
    A = sp.random(n, n, density=4.0 / n, random_state=rng, format="csc") + sp.eye(n) * 3.0
    A = A + sp.diags(-np.ones(n - 1), 1) + sp.diags(-np.ones(n - 1), -1)
    A = column_scale(sp.csc_matrix(A))
    P = Problem("irregular_3000", A, mt_uniform(7, n), 0, 0, 6, 1e-13, 0)
    T = build_tree(P, 64)
    T.factorize()
    b = P.b[T.cperm()]
    Ap = A[T.cperm(), :][:, T.cperm()].tocsc()
    xd = spl.spsolve(Ap, b)
    x = T.solve(b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-10


@pytest.mark.parametrize("coords", [False, True])
def test_matrix_market_input_reaches_1e12(tmp_path, coords):
    P0 = make_problem(1)  # 64 x 64 Poisson, unscaled matrix written to disk
    from pspaqr_public_amd.problems import poisson
    A = poisson(64, 2)
    f = tmp_path / "poisson64.mtx"
    mmio.write_sparse(f, A)
    c = None
    if coords:
        c = tmp_path / "coords.mtx"
        mmio.write_dense(c, linspace_nd(64, 2))
    P = load_problem(f, nlevels=P0.nlevels, tol=P0.tol, coordinates=c)
    assert (P.A != P0.A).nnz == 0 and np.array_equal(P.b, P0.b)
    T = build_tree(P, 32)
    if coords:
        R = build_tree(P0, 32)
        assert np.array_equal(T.cperm(), R.cperm())
    T.factorize()
    b = P.b[T.cperm()]
    x, res, _ = T.gmres(b, iters=300, tol=1e-12, restart=50, true_tol=1e-12, flexible=True)
    print("[mtx] coords=%s: gmres %d, true relres %.2e" % (coords, res.iters, res.true_res))
    assert res.true_res <= 1e-12
