"""Pins of the oracle's tree composition and ordering by facts independent of its code.

* Ordering: the oracle's permutation and level-0 cluster table equal an independent grid
  restatement of the reference's nested dissection (tests/nd_geometry.py), and the separator
  sizes equal the closed-form geometry of the median-split rule (two grid planes per cut).
* Sparsified composition: with a tiny tolerance the sparsification drops only columns whose
  couplings are below tol, so scaling, RRQR gather, truncation, sparsify apply and merge
  together must still give the exact inverse: the preconditioner applied to b equals
  scipy's sparse direct solve.  The cases are chosen so that columns really are dropped.
* want_sparsify's level range: the top two levels and the levels below skip are never
  sparsified (src/ptree.cpp:29-41).
Each check is shown to catch a planted mistake in the oracle (oracle.planted_fault).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

from oracle.oracle import OracleTree, planted_fault
from pspaqr_public_amd.problems import make_problem

import nd_geometry as NDG

ORDER_CASES = [(1, None), (2, 12), (3, 16), (5, 16), (2, None)]


def _leaf_table(cl):
    c0 = cl[cl[:, 0] == 0]
    ids = [((r[1], r[2]), (r[3], r[4]), (r[5], r[6]), int(r[7]), int(r[8])) for r in c0]
    return ids, [int(r[10]) for r in c0]


def _ordering_matches(P, O):
    perm, leaves = NDG.ordering(P.A, P.n_side, P.dim, P.nlevels)
    ids, sizes = _leaf_table(O.clusters())
    return (np.array_equal(O.cperm(), perm) and ids == [l[0] for l in leaves] and sizes == [l[2] for l in leaves])


@pytest.mark.parametrize("cfg,side", ORDER_CASES)
def test_ordering_equals_grid_restatement(cfg, side):
    P = make_problem(cfg, side)
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, P.skip)
    assert _ordering_matches(P, O)


def test_ordering_check_catches_planted_side_error():
    P = make_problem(1, None)
    with planted_fault(1):
        O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, P.skip)
    assert not _ordering_matches(P, O)


@pytest.mark.parametrize("n,d,nlevels,expect", [
    # largest separator per level, top to bottom: two grid planes of the box being cut, less
    # the planes already taken by the cuts above (the raw sizes of the survey's table)
    (64, 2, 6, [128, 62, 62, 30, 30]),
    (32, 3, 9, [2048, 960, 450, 450, 210, 98, 98, 42]),
    (64, 3, 12, [8192, 3968, 1922, 1922, 930, 450, 450, 210, 98, 98, 42]),
])
def test_separator_sizes_closed_form(n, d, nlevels, expect):
    sizes = NDG.separator_sizes(n, d, nlevels)
    got = [max(sizes[l]) for l in range(nlevels - 1, 0, -1)]
    assert got == expect
    # the top cut is two full planes of the grid
    assert sizes[nlevels - 1] == [2 * n ** (d - 1)]


def test_oracle_separator_columns_match_geometry():
    """Per level, the oracle's separator clusters hold exactly the geometric separators."""
    P = make_problem(2, None)
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, P.skip)
    cl = O.clusters()
    c0 = cl[cl[:, 0] == 0]
    geo = NDG.separator_sizes(P.n_side, P.dim, P.nlevels)
    for lvl in range(1, P.nlevels):
        rows = c0[c0[:, 1] == lvl]
        per_sep = np.bincount(rows[:, 2], weights=rows[:, 10]).astype(int)
        assert sorted(per_sep[per_sep > 0].tolist()) == geo[lvl]


# ---------------------------------------------------------------- sparsified composition
EXACT_CASES = [
    # (config, side, tol, columns must be dropped)
    (1, None, 1e-13, True),
    (3, 20, 1e-12, True),
    (2, 16, 1e-13, False),
    (3, 16, 1e-13, False),
]


def _exact_inverse_error(P, tol, skip=0):
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, tol, skip)
    O.factorize()
    cp = O.cperm()
    App = P.A[cp][:, cp].tocsc()
    b = P.b[cp]
    x = O.solve(b)
    xe = spl.spsolve(App, b)
    tr = O.trace()
    rr = tr[tr[:, 1] == 6]  # RRQR records: (level, 6, k, w, rank)
    return np.linalg.norm(x - xe) / np.linalg.norm(xe), int(rr[:, 2].sum()), int(rr[:, 4].sum()), len(rr)


@pytest.mark.parametrize("cfg,side,tol,drops", EXACT_CASES)
def test_sparsified_path_is_exact_at_tiny_tol(cfg, side, tol, drops):
    P = make_problem(cfg, side)
    err, ksum, rsum, nsp = _exact_inverse_error(P, tol)
    assert nsp > 0  # the scaling / RRQR / sparsify / merge path ran
    if drops:
        assert rsum < ksum  # columns were really discarded
    assert err < 1e-10, err


@pytest.mark.parametrize("fault", [2, 3, 4])
@pytest.mark.parametrize("cfg,side", [(1, None), (3, 16)])
def test_exactness_check_catches_planted_composition_errors(fault, cfg, side):
    P = make_problem(cfg, side)
    with planted_fault(fault):
        err, *_ = _exact_inverse_error(P, 1e-13)
    assert err > 1e-9, (fault, err)  # at least ten times the exactness bar above


def _sparsified_levels(P, skip):
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, skip)
    O.factorize()
    cl = O.clusters()
    return sorted(set(cl[cl[:, 12] >= 0, 0].tolist()))


@pytest.mark.parametrize("skip", [0, 1, 2])
def test_top_two_levels_and_skip_never_sparsified(skip):
    P = make_problem(1, None)
    lv = _sparsified_levels(P, skip)
    assert lv == list(range(skip, P.nlevels - 2))


def test_level_range_check_catches_planted_error():
    """Sparsifying the second level from the top either changes the sparsified level set or
    leaves the root with nothing to eliminate (the oracle then refuses to factorize)."""
    P = make_problem(1, None)
    with planted_fault(5):
        try:
            lv = _sparsified_levels(P, 0)
        except RuntimeError:
            return
    assert lv != list(range(0, P.nlevels - 2))
