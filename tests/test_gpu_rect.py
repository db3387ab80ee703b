"""Rectangular (least-squares) systems on the GPU against the oracle.

A = [K; lam G] (m > n, problems.least_squares) ordered by the grid's geometric leaf parts (the
METIS path with given parts).  FP64: the same row filing, reassignment, cluster table and ranks
(column and extra-row), ParTree::solve(b, x) and solve_nrml within 1e-10, CGLS with the same
iteration count.  A tolerance that drops nothing makes solve(b) the least-squares solution.
FP32 fronts: CGLS (FP64) still reaches 1e-12.
"""
import numpy as np
import pytest

from oracle.oracle import OracleTree
from pspaqr_public_amd.problems import least_squares
from pspaqr_public_amd.ptree import build_tree

import nd_geometry

pytestmark = pytest.mark.gpu


def _problem(n, d, nl, tol):
    P = least_squares(n, d, tol=tol, nlevels=nl)
    P.parts = nd_geometry.leaf_parts(n, d, nl).ravel().astype(np.int32)
    return P


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("n,d,nl,tol", [(16, 2, 5, 1e-2), (32, 2, 6, 1e-2), (8, 3, 5, 1e-2), (16, 3, 7, 1e-2),
                                        (32, 2, 6, 1e-14)])
def test_fp64_matches_the_oracle(n, d, nl, tol):
    P = _problem(n, d, nl, tol)
    print("case", n, d, nl, tol, flush=True)
    T = build_tree(P, 64)
    T.factorize()
    O = OracleTree(P.A, 0, 0, nl, tol, 0, parts=P.parts)
    O.factorize()
    assert np.array_equal(T.cperm(), O.cperm())
    assert np.array_equal(T.rperm(), O.rperm())
    assert np.array_equal(T.clusters(), O.clusters())
    assert np.array_equal(T.rows(), O.rows())
    rp, cp = O.rperm(), O.cperm()
    b = P.b[rp]
    assert _rel(T.solve_ls(b), O.solve_ls(b)) < 1e-10
    g = (P.A.T @ P.b)[cp]
    assert _rel(T.solve_nrml(g), O.solve_nrml(g)) < 1e-10
    assert _rel(T.spmv_t(b), O.spmv_t(b)) < 1e-13
    x, it, rel = T.cgls(b, 300, 1e-12)
    xo, ito, relo = O.cgls(b, 300, 1e-12)
    assert it == ito and rel < 1e-12
    assert _rel(x, xo) < 1e-9
    if tol < 1e-12:
        xd = np.linalg.lstsq(P.A.toarray(), P.b, rcond=None)[0][cp]
        assert _rel(T.solve_ls(b), xd) < 1e-10


@pytest.mark.parametrize("n,d,nl", [(32, 2, 6), (16, 3, 7)])
def test_fp32_fronts_reach_the_least_squares_solution(n, d, nl):
    P = _problem(n, d, nl, 1e-2)
    print("case", n, d, nl, flush=True)
    T = build_tree(P, 32)
    T.factorize()
    rp, cp = T.rperm(), T.cperm()
    x, it, rel = T.cgls(P.b[rp], 300, 1e-12)
    assert rel < 1e-12 and it < 40
    xd = np.linalg.lstsq(P.A.toarray(), P.b, rcond=None)[0][cp]
    assert _rel(x, xd) < 1e-9


def test_graph_bisection_ordering_is_exact_at_tiny_tolerance():
    """No parts: the library bisects the graph of A^T A itself; a tolerance that drops nothing
    still makes the factorization an exact QR."""
    P = least_squares(24, 2, tol=1e-14, nlevels=5)
    P.parts = None
    T = build_tree(P, 64)
    T.factorize()
    rp, cp = T.rperm(), T.cperm()
    xd = np.linalg.lstsq(P.A.toarray(), P.b, rcond=None)[0][cp]
    assert _rel(T.solve_ls(P.b[rp]), xd) < 1e-10
    x, it, rel = T.cgls(P.b[rp], 50, 1e-12)
    assert it <= 1 and _rel(x, xd) < 1e-10


def test_square_only_paths_refuse_a_rectangular_tree():
    P = _problem(16, 2, 5, 1e-2)
    T = build_tree(P, 64)
    T.factorize()
    with pytest.raises(RuntimeError):
        T.gmres(np.zeros(P.A.shape[0]), 10)


def test_square_tree_solve_nrml_matches_the_oracle():
    """solve_nrml on a square tree (the transposed pass on the square structure)."""
    from pspaqr_public_amd.problems import make_problem

    P = make_problem(2, 12)
    T = build_tree(P, 64)
    T.factorize()
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, P.skip)
    O.factorize()
    y = np.random.default_rng(5).standard_normal(P.A.shape[0])
    assert _rel(T.solve_nrml(y), O.solve_nrml(y)) < 1e-10


def test_matrix_market_least_squares(tmp_path):
    """The driver's input path for an overdetermined A: a Matrix Market file, no coordinates
    (graph bisection), b from std::mt19937(seed); cgls to 1e-12 in FP64."""
    from pspaqr_public_amd import mmio
    from pspaqr_public_amd.problems import load_problem

    P0 = least_squares(20, 2, tol=1e-2, nlevels=5)
    f = str(tmp_path / "ls.mtx")
    mmio.write_sparse(f, P0.A)
    P = load_problem(f, nlevels=5, tol=1e-2, seed=4)
    assert P.A.shape == P0.A.shape and P.b.size == P0.A.shape[0]
    T = build_tree(P, 64)
    T.factorize()
    rp, cp = T.rperm(), T.cperm()
    x, it, rel = T.cgls(P.b[rp], 200, 1e-12)
    assert rel < 1e-12
    xd = np.linalg.lstsq(P.A.toarray(), P.b, rcond=None)[0][cp]
    assert _rel(x, xd) < 1e-9


def test_fp32_sketch_ranks_on_a_rectangular_tree():
    """FP32 fronts with the sketched rank path (rank mode 2) for the column RRQR; the extra-row
    RRQR stays geqp3.  cgls reaches 1e-12."""
    P = _problem(32, 2, 6, 1e-2)
    T = build_tree(P, 32)
    T.set_rank_mode(2)
    T.factorize()
    rp, cp = T.rperm(), T.cperm()
    x, it, rel = T.cgls(P.b[rp], 300, 1e-12)
    assert rel < 1e-12
    xd = np.linalg.lstsq(P.A.toarray(), P.b, rcond=None)[0][cp]
    assert _rel(x, xd) < 1e-9
