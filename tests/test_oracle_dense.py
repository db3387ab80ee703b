"""Pins the oracle's dense restatements (oracle/dense.hpp) against LAPACK and closed forms.

* tests/golden/lapack_fixtures.npz holds scipy-LAPACK outputs of dgeqrt, dtpqrt, dtpmqrt and
  dgeqp3 (made by tests/golden/make_lapack_fixtures.py); the oracle must reproduce them.
* Closed forms the arithmetic cannot fake: I - V T V^T equals the explicit product of the
  Householder reflectors, Q R reproduces the input, pivoted QR picks the brute-force argmax
  of projected column norms, choose_rank's special cases (src/util.cpp:830-851).
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla

from oracle.oracle import F, lib, unF

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "lapack_fixtures.npz"))


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("i", range(5))
def test_geqrt_matches_lapack(i):
    A = GOLD["geqrt%d_in" % i]
    nb = int(GOLD["geqrt%d_nb" % i][0])
    m, n = A.shape
    a = F(A)
    T = np.zeros(nb * min(m, n))
    lib().orc_geqrt(m, n, nb, a, T)
    assert rel(unF(a, m, n), GOLD["geqrt%d_a" % i]) < 1e-13
    Tg = GOLD["geqrt%d_t" % i]
    To = unF(T, nb, min(m, n))
    for p0 in range(0, min(m, n), nb):
        w = min(nb, min(m, n) - p0)
        assert rel(np.triu(To[:w, p0:p0 + w]), np.triu(Tg[:w, p0:p0 + w])) < 1e-13


@pytest.mark.parametrize("i", range(3))
def test_tpqrt_tpmqrt_match_lapack(i):
    R, B = GOLD["tpqrt%d_R" % i], GOLD["tpqrt%d_B" % i]
    m, n = B.shape
    nb = min(32, n)
    a, b, T = F(R), F(B), np.zeros(nb * n)
    lib().orc_tpqrt(m, n, nb, a, b, T)
    assert rel(np.triu(unF(a, n, n)), np.triu(GOLD["tpqrt%d_a" % i])) < 1e-13
    assert rel(unF(b, m, n), GOLD["tpqrt%d_b" % i]) < 1e-13
    Tg = GOLD["tpqrt%d_t" % i]
    To = unF(T, nb, n)
    for p0 in range(0, n, nb):
        w = min(nb, n - p0)
        assert rel(np.triu(To[:w, p0:p0 + w]), np.triu(Tg[:w, p0:p0 + w])) < 1e-13
    A2, C2 = F(GOLD["tpmqrt%d_A" % i]), F(GOLD["tpmqrt%d_C" % i])
    lib().orc_tpmqrt_lt(m, 11, n, nb, b, T, A2, C2)
    assert rel(unF(A2, n, 11), GOLD["tpmqrt%d_a" % i]) < 1e-13
    assert rel(unF(C2, m, 11), GOLD["tpmqrt%d_b" % i]) < 1e-13


@pytest.mark.parametrize("i", range(4))
def test_geqp3_matches_lapack(i):
    A = GOLD["geqp3%d_in" % i]
    m, n = A.shape
    a = F(A)
    jp = np.zeros(n, dtype=np.int32)
    tau = np.zeros(min(m, n))
    lib().orc_geqp3(m, n, a, jp, tau)
    k = min(m, n)
    # pivots are decided by norms; compare the pivots and factors where the spectrum is
    # resolved (|R_ii| well above rounding), which covers every rank choose_rank can return
    d = np.abs(np.diag(GOLD["geqp3%d_qr" % i][:k, :k]))
    r = int(np.sum(d / d[0] > 1e-10))
    assert np.array_equal(jp[:r], GOLD["geqp3%d_jpvt" % i][:r])
    assert rel(unF(a, m, n)[:, :r], GOLD["geqp3%d_qr" % i][:, :r]) < 1e-12
    assert rel(tau[:r], GOLD["geqp3%d_tau" % i][:r]) < 1e-12


def test_pivoted_qr_brute_force_pivots():
    rng = np.random.default_rng(7)
    for trial in range(20):
        m, n = rng.integers(2, 7), rng.integers(2, 9)
        A = rng.standard_normal((m, n))
        a = F(A)
        jp = np.zeros(n, dtype=np.int32)
        tau = np.zeros(min(m, n))
        lib().orc_geqp3(m, n, a, jp, tau)
        chosen = []
        for step in range(min(m, n)):
            if chosen:
                Qc, _ = np.linalg.qr(A[:, chosen])
                P = np.eye(m) - Qc @ Qc.T
            else:
                P = np.eye(m)
            norms = np.array([np.linalg.norm(P @ A[:, j]) if j not in chosen else -1 for j in range(n)])
            chosen.append(int(np.argmax(norms)))
        assert list(jp[: min(m, n)]) == chosen
        # Q R == A P
        Rf = np.triu(unF(a, m, n))
        V = np.tril(unF(a, m, n), -1)[:, : min(m, n)] + np.eye(m, min(m, n))
        Q = np.eye(m)
        for j in range(min(m, n)):
            v = V[:, j]
            Q = Q @ (np.eye(m) - tau[j] * np.outer(v, v))
        assert rel(Q @ Rf, A[:, jp]) < 1e-13


def test_larft_is_product_of_reflectors():
    rng = np.random.default_rng(8)
    m, k = 20, 7
    V = np.tril(rng.standard_normal((m, k)), -1) + np.eye(m, k)
    tau = 2.0 / np.sum(V * V, axis=0)  # exact reflectors
    T = np.zeros(k * k)
    lib().orc_larft(m, k, F(V), tau, T)
    H = np.eye(m)
    for j in range(k):
        H = H @ (np.eye(m) - tau[j] * np.outer(V[:, j], V[:, j]))
    Tm = unF(T, k, k)
    assert rel(np.eye(m) - V @ Tm @ V.T, H) < 1e-14
    # larfb left/transposed and right/no-trans against the explicit H
    C = rng.standard_normal((m, 9))
    c = F(C)
    lib().orc_larfb_lt(m, 9, k, F(V), T, c)
    assert rel(unF(c, m, 9), H.T @ C) < 1e-14
    C2 = rng.standard_normal((6, m))
    c2 = F(C2)
    lib().orc_larfb_rn(6, m, k, F(V), T, c2)
    assert rel(unF(c2, 6, m), C2 @ H) < 1e-14
    # ormqr (tau form) in both directions
    x = rng.standard_normal(m)
    y = x.copy()
    lib().orc_ormqr_lt(m, k, F(V), tau, y)
    assert rel(y, H.T @ x) < 1e-14
    z = x.copy()
    lib().orc_ormqr_ln(m, k, F(V), tau, z)
    assert rel(z, H @ x) < 1e-14


def test_triangular_solves():
    rng = np.random.default_rng(9)
    n = 17
    R = np.triu(rng.standard_normal((n, n))) + 5 * np.eye(n)
    B = rng.standard_normal((8, n))
    b = F(B)
    lib().orc_trsm_right_upper(8, n, F(R), b)
    assert rel(unF(b, 8, n), B @ np.linalg.inv(R)) < 1e-13
    x = rng.standard_normal(n)
    y = x.copy()
    lib().orc_trsv_upper(n, F(R), y)
    assert rel(y, sla.solve_triangular(R, x)) < 1e-14


def test_choose_rank_special_cases():
    cr = lambda s, tol: lib().orc_choose_rank(np.ascontiguousarray(s, dtype=np.float64), len(s), tol)
    s = np.array([4.0, -2.0, 0.5, 0.01, 0.3])
    assert cr(s, 0.0) == 5            # tol == 0: full
    assert cr(s, 1.0) == 0            # tol >= 1: nothing
    assert cr(s, 0.1) == 3            # stops at the first |s_i|/|s_0| < tol (0.0025)
    assert cr(s, 0.5) == 2            # |-2|/4 = 0.5 >= 0.5 counts
    assert cr(np.array([3.0]), 0.5) == 1
    assert cr(np.array([]), 0.5) == 0
    assert cr(np.array([0.0, 0.0]), 0.5) == 1  # 0/0 fails the test: rank 1


def test_pivot_tie_band_prefers_lowest_index():
    """With exactly tied columns the tie band makes the pivot independent of rounding: the
    lowest-index member of the band wins; with no ties the band changes nothing."""
    rng = np.random.default_rng(11)
    base = rng.standard_normal((6, 3))
    A = np.concatenate([base[:, [1]], base, base[:, [2, 0]]], axis=1)  # duplicated columns
    m, n = A.shape
    a = F(A)
    jp = np.zeros(n, dtype=np.int32)
    tau = np.zeros(min(m, n))
    lib().orc_geqp3_tie(m, n, a, jp, tau, 1e-10)
    norms = np.linalg.norm(A, axis=0)
    top = norms.max()
    assert jp[0] == np.nonzero(norms >= top * (1 - 1e-10))[0][0]
    B = rng.standard_normal((9, 14))
    b1, b2 = F(B), F(B)
    j1, j2 = np.zeros(14, dtype=np.int32), np.zeros(14, dtype=np.int32)
    t1, t2 = np.zeros(9), np.zeros(9)
    lib().orc_geqp3(9, 14, b1, j1, t1)
    lib().orc_geqp3_tie(9, 14, b2, j2, t2, 1e-10)
    assert np.array_equal(j1, j2) and np.array_equal(b1, b2)
