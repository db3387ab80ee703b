"""The oracle of the randomized rank path (oracle/rank_path.hpp), pinned by mathematics.

* Omega: the Box-Muller mapping gives N(0, 1) entries (moments and a Kolmogorov-Smirnov test),
  entries depend on (i, j, seed) only.
* Column norms and the sketch: against numpy.
* Tail selection: against a brute-force search on small inputs, and its special cases.
* The whole sketch pipeline on matrices with a prescribed spectrum: with singular values 1
  (r0 of them) and 1e-12 (the rest), the rank at tol = 1e-6 is exactly r0 whenever the sketch
  has at least r0 columns, for pivoted and unpivoted sketch QRs; narrower sketches fail.
* Features: closed forms for a known R diagonal and column norms.
"""
import numpy as np
import pytest
import scipy.stats

from oracle.oracle import F, lib as olib

SEED = 0x5EED5CA1E


def gauss(rows, cols, seed=SEED):
    om = np.zeros(rows * cols)
    olib().orc_gauss(rows, cols, seed, om)
    return om.reshape((rows, cols), order="F")


def test_gauss_is_standard_normal_and_position_keyed():
    om = gauss(300, 400)
    v = om.ravel()
    assert abs(v.mean()) < 0.01 and abs(v.std() - 1.0) < 0.01
    assert scipy.stats.kstest(v, "norm").pvalue > 1e-3
    assert np.array_equal(gauss(120, 70), om[:120, :70])  # a sub-block regenerates identically
    assert not np.array_equal(gauss(50, 50, seed=1), om[:50, :50])
    # columns are uncorrelated
    c = np.corrcoef(om[:, :20].T)
    assert np.max(np.abs(c - np.eye(20))) < 0.25


def test_colnorm_and_sketch_match_numpy():
    rng = np.random.default_rng(0)
    k, w, s = 37, 150, 21
    Cm = rng.standard_normal((k, w))
    cn = np.zeros(w)
    mx = olib().orc_colnorm(k, w, F(Cm), cn)
    assert np.allclose(cn, (Cm * Cm).sum(0), rtol=1e-14)
    assert mx == cn.max()
    om = gauss(w + 5, s)
    Y = np.zeros(k * s)
    olib().orc_sketch(k, w, s, F(Cm), F(om), w + 5, Y)
    assert np.allclose(Y.reshape((k, s), order="F"), Cm @ om[:w, :], rtol=1e-12, atol=1e-12)


def _tail_brute(B, ry, cmax2, tol):
    k, w = B.shape
    for r in range(0, min(ry, k) + 1):
        if all(np.sum(B[r:, j] ** 2) <= tol * tol * cmax2 for j in range(w)):
            return r
    return -1


def test_tail_select_brute_force_and_special_cases():
    rng = np.random.default_rng(1)
    for trial in range(40):
        k, w = rng.integers(1, 30), rng.integers(1, 20)
        B = rng.standard_normal((k, w)) * (0.5 ** np.arange(k))[:, None]
        cmax2 = float(rng.uniform(0.5, 5))
        tol = float(10 ** rng.uniform(-4, -0.5))
        ry = int(rng.integers(0, k + 1))
        got = olib().orc_tail_select(k, w, F(B), ry, cmax2, tol)
        assert got == _tail_brute(B, ry, cmax2, tol), (k, w, ry, tol)
    B = np.zeros((10, 4))
    assert olib().orc_tail_select(10, 4, F(B), 10, 1.0, 1e-3) == 0  # nothing above tol: rank 0
    B[9, 2] = 1.0
    assert olib().orc_tail_select(10, 4, F(B), 10, 1.0, 1e-3) == 10
    assert olib().orc_tail_select(10, 4, F(B), 9, 1.0, 1e-3) == -1  # rY too small


def _spectrum(rng, k, w, r0, floor=1e-12):
    U, _ = np.linalg.qr(rng.standard_normal((k, min(k, w))))
    V, _ = np.linalg.qr(rng.standard_normal((w, min(k, w))))
    s = np.full(min(k, w), floor)
    s[:r0] = 1.0 + rng.uniform(0, 1, r0)
    return (U * s) @ V.T


@pytest.mark.parametrize("pivoted", [1, 0])
@pytest.mark.parametrize("k,w,r0", [(60, 200, 5), (60, 200, 17), (90, 300, 40), (40, 35, 12)])
def test_sketch_rank_prescribed_spectrum(pivoted, k, w, r0):
    rng = np.random.default_rng(k * 1000 + w + r0)
    Cm = _spectrum(rng, k, w, r0)
    om = gauss(w, min(k, w))
    for s in (r0, r0 + 3, min(k, w)):
        got = olib().orc_sketch_rank(k, w, s, F(Cm), F(om), w, 1e-6, pivoted)
        assert got == r0, (s, got)
    if r0 > 3:
        assert olib().orc_sketch_rank(k, w, r0 - 3, F(Cm), F(om), w, 1e-6, pivoted) == -1
    # tol > 1 keeps nothing: every column, hence every tail, is below tol * max column norm
    assert olib().orc_sketch_rank(k, w, r0, F(Cm), F(om), w, 2.0, pivoted) == 0


def test_features_closed_forms():
    k, w, level, tol = 40, 100, 3, 1e-2
    rdiag = 10.0 ** (-np.arange(16) / 2.0)  # |R_jj| / |R_00| = 10^(-j/2)
    cn = np.full(w, 1e-8)
    cn[:10] = 4.0  # ten columns above tol^2 * cmax2
    x = np.zeros(32, dtype=np.float32)
    olib().orc_rank_features(k, w, rdiag, cn, 4.0, 7, 5, level, tol, x)
    expect_decay = np.clip(1 + (-np.arange(16) / 2.0) / 8.0, 0, 1)
    assert np.allclose(x[:16], expect_decay, atol=1e-6)
    tot = cn.sum()
    expect = [np.log2(k + 1) / 11, np.log2(w + 1) / 14, k / w, level / 16, 2 / 8, 7 / 32, 5 / 32, 10 / w, 10 / k,
              4.0 / tot, min(1.0, tot / 4.0 / k)]
    assert np.allclose(x[16:27], expect, rtol=1e-6)
    assert np.all(x[27:] == 0)
