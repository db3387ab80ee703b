"""Orderings without coordinates (spaqr_tree_set_parts / the library's graph bisection; the
reference's METIS path, src/ptree.cpp:503-518 -> src/partition.cpp:354-424,637-682), on the CPU.

* Given leaf parts: the planner's permutation and cluster table equal the oracle's (its
  restatement of MetisPartition with the same parts) -- with the geometric path's parts of the
  grid problems, with the library's own graph parts, and with arbitrary parts.  Given the
  geometric parts, the separators are the geometric path's (the same rule on the same parts).
* The library's graph bisection, on grids, an irregular graph and a disconnected one: a valid
  nested dissection -- every A^T A edge joins clusters on one root-to-leaf path of the separator
  tree (what lets each interior be eliminated without touching unrelated clusters) -- with parts
  split as evenly as the counts allow at every bisection.
* Bad parts are refused before anything runs; coordinates take precedence over parts.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.oracle import OracleTree
from pspaqr_public_amd.problems import make_problem
from pspaqr_public_amd.ptree import ParTree, linspace_nd


def _tree(nlevels, tol=1e-2, X=None, parts=None):
    t = ParTree(nlevels, 0)
    t.set_tol(tol)
    if X is not None:
        t.set_Xcoo(X)
    if parts is not None:
        t.set_parts(parts)
    return t


def _selves(t):
    """Per natural index: (self lvl, self sep, l lvl, l sep, r lvl, r sep) from the leaf clusters."""
    cl = t.clusters()
    cp = t.cperm()
    out = np.zeros((len(cp), 6), dtype=np.int64)
    covered = np.zeros(len(cp), dtype=bool)
    for r in cl[cl[:, 0] == 0]:  # columns 9, 10: cstart and size at creation
        out[cp[r[9]: r[9] + r[10]]] = r[1:7]
        covered[cp[r[9]: r[9] + r[10]]] = True
    assert covered.all()
    return out


def _ancestor(a, b):
    """True when separator a = (lvl, sep) is b or an ancestor of b."""
    la, sa = a
    lb, sb = b
    return la >= lb and (sb >> (la - lb)) == sa


def _irregular(n, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=3.0 / n, random_state=rng, format="csc")
    path = sp.diags([-np.ones(n - 1), -np.ones(n - 1)], [1, -1])  # keeps it connected
    return sp.csc_matrix(A + A.T + sp.eye(n) * 4.0 + path)


def _grids():
    a = make_problem(1, 16).A
    b = make_problem(3, 8).A
    return sp.block_diag([a, b], format="csc")


@pytest.mark.parametrize("case", ["poisson2d_64", "convdiff3d_12", "irregular", "disconnected"])
def test_graph_bisection_is_a_valid_nested_dissection(case):
    if case == "poisson2d_64":
        A, nl = make_problem(1).A, 6
    elif case == "convdiff3d_12":
        A, nl = make_problem(3, 12).A, 6
    elif case == "irregular":
        A, nl = _irregular(1500, 5), 5
    else:
        A, nl = _grids(), 5
    t = _tree(nl)
    t.partition(A)
    n = A.shape[0]
    cp = t.cperm()
    assert np.array_equal(np.sort(cp), np.arange(n))
    parts = t.parts()
    assert parts.min() >= 0 and parts.max() < 2 ** (nl - 1)
    # balance at every bisection: part k of depth d splits into halves of sizes differing by <= 1
    for depth in range(nl - 1):
        sh = nl - 2 - depth
        cnt = np.bincount(parts >> sh, minlength=2 ** (depth + 1))
        for k in range(2 ** depth):
            assert 0 <= cnt[2 * k] - cnt[2 * k + 1] <= 1
    # every A^T A edge joins clusters on one root-to-leaf path
    S = _selves(t)
    G = (abs(A).T @ abs(A)).tocoo()
    for i, j in zip(G.row, G.col):
        a, b = (S[i, 0], S[i, 1]), (S[j, 0], S[j, 1])
        assert _ancestor(a, b) or _ancestor(b, a), (i, j, a, b)
    # interfaces: l and r are the vertex's own cluster or its ancestors' -- on its path
    for i in range(n):
        s, l, r = (S[i, 0], S[i, 1]), (S[i, 2], S[i, 3]), (S[i, 4], S[i, 5])
        if s[0] == 0:
            assert l == s and r == s


@pytest.mark.parametrize("source", ["geometric", "graph", "random"])
def test_given_parts_match_the_oracle(source):
    P = make_problem(1, 32) if source != "graph" else make_problem(3, 12)
    nl = P.nlevels
    if source == "geometric":
        tg = _tree(nl, X=linspace_nd(P.n_side, P.dim))
        tg.partition(P.A)
        parts = tg.parts()
    elif source == "graph":
        tq = _tree(nl)
        tq.partition(P.A)
        parts = tq.parts()
    else:
        parts = np.random.default_rng(9).integers(0, 2 ** (nl - 1), P.A.shape[0]).astype(np.int32)
    t = _tree(nl, parts=parts)
    t.partition(P.A)
    assert np.array_equal(t.parts(), parts)
    O = OracleTree(P.A, 0, 0, nl, P.tol, 0, parts=parts)
    assert np.array_equal(t.cperm(), O.cperm())
    assert np.array_equal(t.clusters()[:, :12], O.clusters()[:, :12])
    if source == "geometric":  # the same separator rule on the same parts
        Sg, Sp = _selves(tg), _selves(t)
        assert np.array_equal(Sg[:, :2], Sp[:, :2])


def test_bad_parts_and_precedence():
    P = make_problem(1, 16)
    t = _tree(P.nlevels, parts=np.zeros(10, np.int32))
    with pytest.raises(RuntimeError, match="leaf part"):
        t.partition(P.A)
    t = _tree(P.nlevels, parts=np.full(P.A.shape[0], 2 ** (P.nlevels - 1), np.int32))
    with pytest.raises(RuntimeError, match="outside"):
        t.partition(P.A)
    # coordinates present: the geometric path runs, the parts are ignored
    t = _tree(P.nlevels, X=linspace_nd(P.n_side, P.dim), parts=np.zeros(P.A.shape[0], np.int32))
    t.partition(P.A)
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, 0)
    assert np.array_equal(t.cperm(), O.cperm())


def test_metis_path_closed_form_on_a_path_graph():
    """The METIS path's rules (src/partition.cpp:380-424, 637-682) by hand on a path graph:
    A tridiagonal (32 vertices), so A^T A joins vertices up to 2 apart; 3 levels, leaf parts
    0..3 of 8 consecutive vertices each.  Level-2 separator: the vertices of parts {0, 1}
    within distance 2 of parts {2, 3}: 14, 15, with left (0, 1) (their own leaf) and right
    (0, 2).  Level-1 separators: 6, 7 of part 0 next to part 1 (left (0, 0), right (0, 1)) and
    22, 23 of part 2 next to part 3 (left (0, 2), right (0, 3)).  Every other vertex is its
    leaf, with l = r = self.  Both the library and the oracle must give exactly this."""
    n = 32
    A = sp.diags([-np.ones(n - 1), 4 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csc")
    parts = np.repeat(np.arange(4), 8).astype(np.int32)
    want = np.zeros((n, 6), dtype=np.int64)
    for i in range(n):
        want[i] = (0, parts[i], 0, parts[i], 0, parts[i])
    for i in (14, 15):
        want[i] = (2, 0, 0, 1, 0, 2)
    for i in (6, 7):
        want[i] = (1, 0, 0, 0, 0, 1)
    for i in (22, 23):
        want[i] = (1, 1, 0, 2, 0, 3)
    t = _tree(3, parts=parts)
    t.partition(A)
    assert np.array_equal(_selves(t), want)
    O = OracleTree(A, 0, 0, 3, 1e-2, 0, parts=parts)
    assert np.array_equal(O.cperm(), t.cperm())
    assert np.array_equal(O.clusters()[:, :12], t.clusters()[:, :12])
    # and the pin catches a planted mistake in the oracle's interface rule
    from oracle.oracle import planted_fault
    with planted_fault(6):
        Of = OracleTree(A, 0, 0, 3, 1e-2, 0, parts=parts)
        assert not np.array_equal(Of.clusters()[:, :12], t.clusters()[:, :12])
