"""The per-matrix symbolic phase on the device (row f1) against the host phase.

ParTree::partition and ParTree::assemble (src/ptree.cpp:495-864) for square systems with
coordinates run as kernels and device sorts (csrc/symbolic.hip).  The host phase is the one
pinned against the oracle (the same permutation and cluster table, test_gpu_tree.py); the
device phase must give exactly its result: the same permutation, leaf parts and cluster table,
blocks bit for bit (so the factorization, the preconditioner and SpMV agree bit for bit), and
the same distance-two connections (so the fill-in, ranks and cluster table after factorization
agree too).
"""
import numpy as np
import pytest

from pspaqr_public_amd.problems import make_problem
from pspaqr_public_amd.ptree import build_tree

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,side,prec", [(1, None, 64), (2, None, 64), (3, 24, 64), (5, 16, 32), (3, 32, 32)])
def test_device_symbolic_equals_the_host_phase(cfg, side, prec):
    P = make_problem(cfg, side)
    print("case", cfg, side, prec, P.A.shape, flush=True)
    H = build_tree(P, prec)
    D = build_tree(P, prec, device_symbolic=True)
    assert np.array_equal(H.cperm(), D.cperm())
    assert np.array_equal(H.parts(), D.parts())
    assert np.array_equal(H.clusters(), D.clusters())
    rng = np.random.default_rng(1)
    x = rng.standard_normal(P.A.shape[0])
    assert np.array_equal(H.spmv(x), D.spmv(x))
    H.factorize()
    D.factorize()
    assert np.array_equal(H.clusters(), D.clusters())
    assert np.array_equal(H.level_stats(), D.level_stats())
    root = H.clusters().shape[0] - 1
    assert np.array_equal(H.diag_R(root), D.diag_R(root))
    assert np.array_equal(H.solve(x), D.solve(x))
    b = P.b[D.cperm()]
    xh, rh, _ = H.gmres(b, 200, 1e-12, restart=50, flexible=prec == 32)
    xd, rd, _ = D.gmres(b, 200, 1e-12, restart=50, flexible=prec == 32)
    assert rh.iters == rd.iters and np.array_equal(xh, xd)
    print("partition/assemble ms host %.1f/%.1f device %.1f/%.1f" % (H.symbolic_ms() + D.symbolic_ms()))


def test_device_symbolic_falls_back_without_coordinates():
    """Leaf parts (the METIS path) keep the host phase; the result is unchanged."""
    from pspaqr_public_amd.problems import Problem

    P = make_problem(1, 32)
    parts = build_tree(P, 64).parts()
    Q = Problem(P.name, P.A, P.b, 0, 0, P.nlevels, P.tol, parts=parts)
    H = build_tree(Q, 64)
    D = build_tree(Q, 64, device_symbolic=True)
    assert np.array_equal(H.cperm(), D.cperm()) and np.array_equal(H.clusters(), D.clusters())
    H.factorize()
    D.factorize()
    x = P.b[H.cperm()]
    assert np.array_equal(H.solve(x), D.solve(x))


def test_device_symbolic_refactorizes_and_reassembles():
    """Two factorizations after a device assembly, and a second assembly in the other precision."""
    P = make_problem(2, 16)
    D = build_tree(P, 64, device_symbolic=True)
    D.factorize()
    x1 = D.solve(P.b[D.cperm()])
    D.factorize()
    assert np.array_equal(x1, D.solve(P.b[D.cperm()]))
    D.set_precision(32)
    D.assemble()
    D.factorize()
    H = build_tree(P, 32)
    H.factorize()
    assert np.array_equal(H.solve(P.b[H.cperm()]), D.solve(P.b[D.cperm()]))
