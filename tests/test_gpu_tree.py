"""GPU parity of the whole path (partition, assemble, factorize, solve, GMRES) against the
oracle on the same inputs.

Bars (FP64): identical permutation and cluster table, identical sparsification ranks,
diagonal R factors of every interior within 1e-12 relative after row-sign normalisation,
preconditioner apply within 1e-10, the same GMRES iteration count and a true relative
residual <= 1e-12.  FP32 fronts: restarted FP64 GMRES still reaches 1e-12.
"""
import numpy as np
import pytest

from oracle.oracle import OracleTree
from pspaqr_public_amd.problems import make_problem
from pspaqr_public_amd.ptree import build_tree

pytestmark = pytest.mark.gpu

CASES = [(1, None), (2, 12), (3, 12), (2, 16), (3, 16), (5, 16)]


def _pair(cfg, side, tol=None, skip=None):
    P = make_problem(cfg, side)
    t = tol if tol is not None else P.tol
    s = skip if skip is not None else P.skip
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, t, s)
    O.factorize()
    T = build_tree(P, 64, tol=t, skip=s)
    T.factorize()
    return P, O, T


def _sign_norm_rel(Rg, Ro):
    d = np.sign(np.diag(Rg)) * np.sign(np.diag(Ro))
    d[d == 0] = 1
    return np.linalg.norm(d[:, None] * Rg - Ro) / max(np.linalg.norm(Ro), 1e-300)


@pytest.mark.parametrize("cfg,side", CASES)
def test_structure_and_ranks(cfg, side):
    P, O, T = _pair(cfg, side)
    assert np.array_equal(T.cperm(), O.cperm())
    cg, co = T.clusters(), O.clusters()
    assert cg.shape == co.shape
    assert np.array_equal(cg[:, :12], co[:, :12])
    # sparsification ranks, interface by interface
    assert np.array_equal(cg[:, 12], co[:, 12]), np.nonzero(cg[:, 12] != co[:, 12])


@pytest.mark.parametrize("cfg,side", CASES)
def test_diag_R_parity(cfg, side):
    P, O, T = _pair(cfg, side)
    cl = O.clusters()
    worst = 0.0
    for order in range(cl.shape[0]):
        if cl[order, 0] != cl[order, 1]:  # interiors only (merge level == separator level)
            continue
        Ro = O.diag_R(order)
        Rg = T.diag_R(order)
        assert Rg is not None and Ro is not None
        worst = max(worst, _sign_norm_rel(Rg, Ro))
    assert worst < 1e-12, worst


@pytest.mark.parametrize("cfg,side", CASES)
def test_solve_and_gmres_parity(cfg, side):
    P, O, T = _pair(cfg, side)
    cp = O.cperm()
    b = P.b[cp]
    xo = O.solve(b)
    xg = T.solve(b)
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10
    assert np.linalg.norm(T.spmv(b) - O.spmv(b)) / np.linalg.norm(O.spmv(b)) < 1e-14
    xo, ito, _ = O.gmres(b, iters=300, tol=1e-12)
    xg, res, _ = T.gmres(b, iters=300, tol=1e-12)
    assert res.iters == ito
    r = b - O.spmv(xg)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-12
    assert res.true_res <= 1e-12


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("cfg,side", [(2, 12), (3, 16)])
def test_solve_on_opt_in_chain_kernels(cfg, side, prec, monkeypatch):
    """The opt-in one-wave and pipelined chain kernels of the solve (SPAQR_WAVE_CHAINS,
    SPAQR_PIPE_CHAINS) give the solve of the default workgroup chain kernel: FP64 against the
    oracle at the 1e-10 bar, FP32 against the default FP32 solve."""
    P = make_problem(cfg, side)
    Td = build_tree(P, prec)
    Td.factorize()
    b = P.b[Td.cperm()]
    xd = Td.solve(b)
    monkeypatch.setenv("SPAQR_WAVE_CHAINS", "64")
    monkeypatch.setenv("SPAQR_PIPE_CHAINS", "512")
    Tp = build_tree(P, prec)
    Tp.factorize()
    xp = Tp.solve(b)
    if prec == 64:
        O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, P.skip)
        O.factorize()
        xo = O.solve(b)
        assert np.linalg.norm(xp - xo) / np.linalg.norm(xo) < 1e-10
    else:
        assert np.linalg.norm(xp - xd) / np.linalg.norm(xd) < 1e-3


@pytest.mark.parametrize("cfg,side", [(1, 32), (2, 12), (3, 12)])
def test_exact_factorization_tol0(cfg, side):
    """tol = 0 (skip = nlevels - 1, pspaQR.cpp:150): no sparsification, the solve is the
    exact inverse."""
    P = make_problem(cfg, side)
    T = build_tree(P, 64, tol=0.0, skip=P.nlevels - 1)
    T.factorize()
    cp = T.cperm()
    b = P.b[cp]
    x = T.solve(b)
    r = b - T.spmv(x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-13


@pytest.mark.parametrize("cfg,side", [(1, None), (3, 16)])
def test_fp32_fronts_reach_fp64_residual(cfg, side):
    P = make_problem(cfg, side)
    T = build_tree(P, 32)
    T.factorize()
    b = P.b[T.cperm()]
    x, res, _ = T.gmres(b, iters=300, tol=1e-12, restart=50, true_tol=1e-12)
    r = b - T.spmv(x)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-12
    assert res.true_res <= 1e-12


@pytest.mark.parametrize("cfg,side", [(1, None), (3, 16)])
def test_flexible_gmres(cfg, side):
    """Right-preconditioned flexible GMRES (no reference counterpart): its recurrence estimates
    the true residual, so with FP32 factors it stops at true relres <= 1e-12 within a couple of
    steps of what the FP64 left-preconditioned path needs; with an exact factorization
    (tol = 0) it converges in one step, as the left-preconditioned form does."""
    P = make_problem(cfg, side)
    T64 = build_tree(P, 64)
    T64.factorize()
    b = P.b[T64.cperm()]
    _, r64, _ = T64.gmres(b, iters=300, tol=1e-12)
    T32 = build_tree(P, 32)
    T32.factorize()
    x, res, _ = T32.gmres(b, iters=300, tol=1e-12, restart=50, true_tol=1e-12, flexible=True)
    rel = np.linalg.norm(b - T32.spmv(x)) / np.linalg.norm(b)
    assert rel <= 1e-12 and res.true_res <= 1e-12
    assert res.iters <= r64.iters + 3, (res.iters, r64.iters)
    Te = build_tree(P, 64, tol=0.0, skip=P.nlevels - 1)
    Te.factorize()
    x, res, _ = Te.gmres(b, iters=10, tol=1e-12, flexible=True)
    assert res.iters == 0
    assert np.linalg.norm(b - Te.spmv(x)) / np.linalg.norm(b) < 1e-13


@pytest.mark.parametrize("cfg,side,prec,mode", [(2, 12, 64, 0), (3, 16, 32, 1), (5, 16, 32, 1)])
def test_refactorize_is_repeatable(cfg, side, prec, mode):
    """Factorizing again, and again after a new assemble(), gives the first factorization's
    results bit for bit (host planning, level-0 blocks and arenas all start afresh)."""
    P = make_problem(cfg, side)
    T = build_tree(P, prec)
    if mode:
        from pspaqr_public_amd import rank_model
        T.set_rank_mode(mode)
        T.set_rank_model(rank_model.load_weights())
    T.factorize()
    b = P.b[T.cperm()]
    x1 = T.solve(b)
    c1 = T.clusters()
    for rep in range(2):
        T.factorize()
        assert np.array_equal(T.solve(b), x1) and np.array_equal(T.clusters(), c1)
    T.assemble()
    T.factorize()
    assert np.array_equal(T.solve(b), x1)


def test_headline_workload_at_full_size_matches_oracle():
    """Config 3 at its stated size (64^3 upwind convection-diffusion, n = 262144, 12 levels,
    tol 1e-2), FP64: the oracle's permutation and cluster table (ranks included), its
    preconditioner apply within 1e-10 and its GMRES iteration count; then the FP32 headline
    path (learned ranks, bf16 sketches, flexible GMRES) to a true relative residual 1e-12."""
    P = make_problem(3)
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, P.skip)
    print("[full] oracle factorization %.1f s" % O.factorize(), flush=True)
    T = build_tree(P, 64)
    T.factorize()
    assert np.array_equal(T.cperm(), O.cperm())
    assert np.array_equal(T.clusters(), O.clusters())
    b = P.b[O.cperm()]
    xo = O.solve(b)
    assert np.linalg.norm(T.solve(b) - xo) / np.linalg.norm(xo) < 1e-10
    _, ito, _ = O.gmres(b, iters=300, tol=1e-12)
    xg, res, _ = T.gmres(b, iters=300, tol=1e-12)
    print("[full] FP64 gmres %d (oracle %d), true relres %.2e" % (res.iters, ito, res.true_res), flush=True)
    assert res.iters == ito and res.true_res <= 1e-12
    T.close()
    from pspaqr_public_amd import rank_model
    F = build_tree(P, 32)
    F.set_rank_mode(1)
    F.set_rank_model(rank_model.load_weights())
    F.factorize()
    _, r32, _ = F.gmres(b, iters=300, tol=1e-12, restart=50, true_tol=1e-12, flexible=True)
    print("[full] FP32 learned: gmres %d, true relres %.2e" % (r32.iters, r32.true_res), flush=True)
    assert r32.true_res <= 1e-12 and r32.iters <= ito


def test_largest_workload_fp32_reaches_1e12():
    """Config 5 at its stated size (96^3 anisotropic diffusion, n = 884736, 14 levels), FP32
    fronts with learned ranks: flexible GMRES to a true relative residual 1e-12."""
    from pspaqr_public_amd import rank_model
    P = make_problem(5)
    F = build_tree(P, 32)
    F.set_rank_mode(1)
    F.set_rank_model(rank_model.load_weights())
    F.factorize()
    b = P.b[F.cperm()]
    _, res, _ = F.gmres(b, iters=300, tol=1e-12, restart=50, true_tol=1e-12, flexible=True)
    st = F.rank_stats()
    print("[full] 96^3 FP32: gmres %d, true relres %.2e, %d interfaces, %d fell back" % (
        res.iters, res.true_res, st["interfaces"], st["fallback"]), flush=True)
    assert res.true_res <= 1e-12
