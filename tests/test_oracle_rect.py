"""The oracle's rectangular (least-squares) path, pinned without the reference.

ParTree on an overdetermined A (m > n) files every row under a cluster (form_rmap,
src/tree.cpp:102-165), reassigns the extra rows of each eliminated interior to a neighbour
(src/ptree.cpp:1149-1356), compresses them with a second RRQR (sparsify_extra :441-493) and
keeps them through the solve (:2267-2598).  CGLS (src/is.cpp:274-345) takes solve_nrml
(src/tree.cpp:1524-1570) as its preconditioner.

Pins:
* a tolerance that drops nothing makes the factorization an exact QR of A: solve(b) is the
  least-squares solution and solve_nrml(A^T b) too (numpy's lstsq on the dense matrix);
* solve_nrml is symmetric (it is M^{-1} M^{-T}), so the transposed pass is the transpose;
* CGLS with tol 1e-2 factors reaches the least-squares solution;
* row filing and row reassignment follow their definitions, checked row by row with numpy
  (the largest sum of squares over the column clusters a row touches; the largest row norm of
  Q_c^T A(c, m) over the candidate neighbours, Q_c from LAPACK's Householder QR);
* planted mistakes (one extra row too few kept; the solve skipping Q_ex^T) fail the pins.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle.oracle import OracleTree, planted_fault
from pspaqr_public_amd.problems import least_squares

import nd_geometry


def _tree(n, d, nlevels, tol):
    P = least_squares(n, d, tol=tol, nlevels=nlevels)
    parts = nd_geometry.leaf_parts(n, d, nlevels).ravel().astype(np.int32)
    O = OracleTree(P.A, 0, 0, nlevels, tol, 0, parts=parts)
    O.factorize()
    return P, O


def _lstsq(P):
    return np.linalg.lstsq(P.A.toarray(), P.b, rcond=None)[0]


def _unperm(xp, cp):
    x = np.zeros_like(xp)
    x[cp] = xp
    return x


@pytest.mark.parametrize("n,d,nl", [(16, 2, 5), (32, 2, 6), (8, 3, 5)])
def test_exact_at_tiny_tolerance(n, d, nl):
    P, O = _tree(n, d, nl, 1e-14)
    rows = O.rows()
    cl = O.clusters()
    assert (cl[:, 12] >= 0).sum() > 0, "no interface was sparsified"
    assert rows[:, 2].sum() > 0, "no extra row was reassigned"
    assert rows[:, 0].max() < P.A.shape[0]
    xd = _lstsq(P)
    rp, cp = O.rperm(), O.cperm()
    assert np.array_equal(np.sort(rp), np.arange(P.A.shape[0]))
    x = _unperm(O.solve_ls(P.b[rp]), cp)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-10
    g = P.A.T @ P.b
    y = _unperm(O.solve_nrml(g[cp]), cp)
    assert np.linalg.norm(y - xd) / np.linalg.norm(xd) < 1e-10


@pytest.mark.parametrize("n,d,nl", [(32, 2, 6), (16, 3, 7)])
def test_cgls_reaches_the_least_squares_solution(n, d, nl):
    P, O = _tree(n, d, nl, 1e-2)
    cl = O.clusters()
    sp = cl[:, 12] >= 0
    assert cl[sp, 12].sum() < cl[sp, 10].sum(), "tol 1e-2 drops no column"
    xd = _lstsq(P)
    rp, cp = O.rperm(), O.cperm()
    # the preconditioner alone is an approximation
    x1 = _unperm(O.solve_ls(P.b[rp]), cp)
    assert np.linalg.norm(x1 - xd) / np.linalg.norm(xd) > 1e-6
    xc, it, rel = O.cgls(P.b[rp], 200, 1e-12)
    assert rel < 1e-12 and it < 30
    x = _unperm(xc, cp)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    # cgls's own accounting: ||A^T (b - A x)|| / ||A^T b||
    r = P.b - P.A @ x
    assert np.linalg.norm(P.A.T @ r) / np.linalg.norm(P.A.T @ P.b) < 1e-11


def test_solve_nrml_is_symmetric_positive_definite():
    P, O = _tree(32, 2, 6, 1e-2)
    rng = np.random.default_rng(3)
    n = P.A.shape[1]
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    Mu, Mv = O.solve_nrml(u), O.solve_nrml(v)
    assert abs(u @ Mv - v @ Mu) <= 1e-12 * abs(u @ Mv)
    assert u @ Mu > 0 and v @ Mv > 0


def test_row_filing_follows_its_definition():
    """Every row is filed under a column of the cluster holding its largest sum of squares
    (first such cluster in column order), at that cluster's first column the row touches; rows
    are ordered by that column, then by row index."""
    n, d, nl = 16, 2, 5
    P, O = _tree(n, d, nl, 1e-2)
    A = P.A.tocsr()
    cp = O.cperm()
    cinv = np.empty_like(cp)
    cinv[cp] = np.arange(cp.size)
    cl = O.clusters()
    lvl0 = cl[cl[:, 0] == 0]
    cstart = lvl0[:, 9]
    owner = np.searchsorted(cstart, np.arange(cp.size), side="right") - 1  # permuted column -> leaf cluster
    filed = np.empty(A.shape[0], dtype=np.int64)
    for i in range(A.shape[0]):
        cols = cinv[A.indices[A.indptr[i]:A.indptr[i + 1]]]
        vals = A.data[A.indptr[i]:A.indptr[i + 1]]
        o = np.argsort(cols)
        cols, vals = cols[o], vals[o]
        best, bcol = -1.0, -1
        for c in np.unique(owner[cols]):
            sel = owner[cols] == c
            w = float(np.sum(vals[sel] ** 2))
            if w > best:
                best, bcol = w, int(cols[sel][0])
        filed[i] = bcol
    expect = np.lexsort((np.arange(A.shape[0]), filed))
    assert np.array_equal(O.rperm(), expect)
    rows = O.rows()[: len(lvl0)]
    counts = np.bincount(owner[filed], minlength=len(lvl0))
    assert np.array_equal(rows[:, 1], counts)


def test_row_reassignment_follows_its_definition():
    """Each extra row of a level-0 interior goes to the candidate neighbour m (an out- and
    in-neighbour of the interior) with the largest norm of that row of Q^T A(c, m)."""
    n, d, nl = 16, 2, 5
    P, O = _tree(n, d, nl, 1e-2)
    rp, cp = O.rperm(), O.cperm()
    App = P.A.tocsc()[rp][:, cp].toarray()
    cl = O.clusters()
    rows = O.rows()
    lvl0 = np.nonzero(cl[:, 0] == 0)[0]
    csl = {o: slice(cl[o, 9], cl[o, 9] + cl[o, 10]) for o in lvl0}
    rsl = {o: slice(rows[o, 0], rows[o, 0] + rows[o, 1]) for o in lvl0}
    checked = 0
    for c in lvl0:
        if cl[c, 1] != 0:  # interiors only
            continue
        k, mr = cl[c, 10], rows[c, 1]
        if mr <= k:
            continue
        Q, _ = sla.qr(App[rsl[c], csl[c]], mode="full")
        cand = []
        for m in lvl0:
            if m == c:
                continue
            into_c = np.any(App[rsl[c], csl[m]] != 0)    # edge m -> c
            from_c = np.any(App[rsl[m], csl[c]] != 0)    # edge c -> m
            if into_c and from_c:
                cand.append(m)
        norms = np.array([np.linalg.norm((Q.T @ App[rsl[c], csl[m]])[k:], axis=1) for m in cand])
        got = O.r2c(c)
        assert got.size == mr - k
        for i in range(mr - k):
            top = norms[:, i].max()
            if top == 0:
                assert got[i] == c
                continue
            # clear winners only (a near tie may go either way with rounding)
            srt = np.sort(norms[:, i])[::-1]
            if len(srt) > 1 and srt[1] > (1 - 1e-9) * srt[0]:
                continue
            assert got[i] == cand[int(np.argmax(norms[:, i]))]
            checked += 1
    assert checked > 50


@pytest.mark.parametrize("fault", [7, 8])
def test_planted_faults_fail_exactness(fault):
    with planted_fault(fault):
        P, O = _tree(16, 2, 5, 1e-14)
        xd = _lstsq(P)
        x = _unperm(O.solve_ls(P.b[O.rperm()]), O.cperm())
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) > 1e-6
