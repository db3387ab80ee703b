"""The product's host planner without a GPU: a tree touches the GPU only from assemble() on, so
ordering, cluster hierarchy and input checks run here through the C ABI.

* The planner's permutation and cluster table equal the oracle's and the grid restatement of the
  reference's nested dissection (tests/nd_geometry.py).
* A batch orders every system as it would be ordered alone.
* Failure injection: non-finite entries, a zero column, a non-square matrix and bad coordinates
  are refused by partition(); tunables and thresholds out of range are refused.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.oracle import OracleTree
from pspaqr_public_amd import _lib as G
from pspaqr_public_amd.problems import make_batch, make_problem
from pspaqr_public_amd.ptree import ParTree, linspace_nd

import nd_geometry as NDG


def _tree(P, nlevels=None):
    t = ParTree(P.nlevels if nlevels is None else nlevels, P.skip)
    t.set_tol(P.tol)
    t.set_Xcoo(linspace_nd(P.n_side, P.dim))
    return t


@pytest.mark.parametrize("cfg,side", [(1, None), (2, 12), (3, 16), (5, 16), (2, None)])
def test_planner_ordering_on_the_host(cfg, side):
    P = make_problem(cfg, side)
    t = _tree(P)
    t.partition(P.A)
    assert G.lib().spaqr_tree_stream(t.handle) is None  # no GPU used so far
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, P.skip)
    assert np.array_equal(t.cperm(), O.cperm())
    cg, co = t.clusters(), O.clusters()
    assert np.array_equal(cg[:, :12], co[:, :12])
    perm, leaves = NDG.ordering(P.A, P.n_side, P.dim, P.nlevels)
    assert np.array_equal(t.cperm(), perm)


def test_batch_orders_each_system_as_alone():
    B = make_batch(count=4, sides=(8, 12, 16))
    t = ParTree(max(B.levels), 0)
    t.set_tol(B.tol)
    t.set_Xcoo(B.X)
    t.partition_batch(B.A, B.sizes, B.levels)
    cp = t.cperm()
    cs = t.cluster_systems()
    for s, (P, off) in enumerate(zip(B.systems, B.offsets)):
        ts = _tree(P)
        ts.partition(P.A)
        n = P.A.shape[0]
        assert np.array_equal(cp[off:off + n] - off, ts.cperm())
        assert (cs == s).sum() == ts.clusters().shape[0]


def _poisson(n=8):
    P = make_problem(1, n)
    return P


def test_partition_refuses_bad_input():
    P = _poisson()
    A = P.A.tocsc(copy=True)
    t = _tree(P)
    A.data[5] = np.nan
    with pytest.raises(RuntimeError, match="not finite"):
        t.partition(A)
    A = P.A.tocsc(copy=True)
    A.data[A.indptr[3]:A.indptr[4]] = 0.0
    with pytest.raises(RuntimeError, match="singular"):
        t.partition(A)
    with pytest.raises(RuntimeError, match="square"):
        t.partition(sp.csc_matrix(np.ones((6, 5))))
    t2 = ParTree(P.nlevels, 0)
    t2.set_Xcoo(np.zeros((2, 10)))
    with pytest.raises(RuntimeError, match="coordinate"):
        t2.partition(P.A)
    X = linspace_nd(P.n_side, P.dim)
    X[0, 3] = np.inf
    t3 = ParTree(P.nlevels, 0)
    t3.set_Xcoo(X)
    with pytest.raises(RuntimeError, match="finite"):
        t3.partition(P.A)
    t.partition(P.A)  # the same tree accepts a good matrix afterwards


def test_settings_out_of_range_are_refused():
    P = _poisson()
    t = _tree(P)
    L = G.lib()
    assert L.spaqr_tree_set_tuning(t.handle, 1, 48) == -3     # tile not in {16, 32, 64}
    assert L.spaqr_tree_set_tuning(t.handle, 99, 64) == -2
    assert L.spaqr_tree_set_tuning(t.handle, 1, 32) == 0 and L.spaqr_tree_get_tuning(t.handle, 1) == 32
    assert L.spaqr_tree_set_pivot_check(t.handle, 1.5) == -2
    assert L.spaqr_tree_set_rank_retries(t.handle, 3) == -2
    assert L.spaqr_tree_set_precision(t.handle, 16) == -2
    with pytest.raises(RuntimeError, match="assemble"):
        t.factorize()  # nothing assembled: refused before any GPU work
