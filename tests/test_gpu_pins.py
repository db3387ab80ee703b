"""GPU checks pinned by facts independent of the oracle's code, and the rank path's kernels
against the rank-path oracle (oracle/rank_path.hpp), all through the C ABI.

* The GPU planner's ordering equals the grid restatement of the reference's nested dissection.
* At a tiny tolerance the whole sparsified FP64 path (scaling, RRQR gather, truncation,
  sparsify apply, merge) is an exact inverse: solve(b) equals scipy's direct solve, with the
  oracle's ranks.
* Rank-path kernels on fixed inputs: Omega, column norms, sketch, tail selection, features
  (FP64 within 1e-12, FP32 within the stated tolerances); the FP32 geqp3.
* A predictor that asks for 8-column sketches with no retries forces the geqp3 fallback; the
  fallen-back interfaces get geqp3's ranks.
* A precision change after assemble() is refused before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse.linalg as spl

from oracle.oracle import F, OracleTree, lib as olib, unF
from pspaqr_public_amd import _lib as G
from pspaqr_public_amd import rank_model
from pspaqr_public_amd.problems import make_problem
from pspaqr_public_amd.ptree import build_tree

import nd_geometry as NDG

pytestmark = pytest.mark.gpu

NB = 32
SEED = 0x5EED5CA1E


def dbuf(a, dtype=np.float64):
    return G.DeviceBuffer.from_array(np.ascontiguousarray(a, dtype=dtype))


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64)) / max(np.linalg.norm(b), 1e-300)


# ---------------------------------------------------------------- ordering and exactness
@pytest.mark.parametrize("cfg,side", [(1, None), (2, 12), (3, 16), (2, None)])
def test_gpu_ordering_equals_grid_restatement(cfg, side):
    P = make_problem(cfg, side)
    T = build_tree(P, 64)
    perm, leaves = NDG.ordering(P.A, P.n_side, P.dim, P.nlevels)
    assert np.array_equal(T.cperm(), perm)
    cl = T.clusters()
    c0 = cl[cl[:, 0] == 0]
    ids = [((r[1], r[2]), (r[3], r[4]), (r[5], r[6]), int(r[7]), int(r[8])) for r in c0]
    assert ids == [l[0] for l in leaves]
    assert [int(r[10]) for r in c0] == [l[2] for l in leaves]


@pytest.mark.parametrize("cfg,side,tol,drops", [(1, None, 1e-13, True), (3, 20, 1e-12, True), (2, 16, 1e-13, False),
                                                (3, 16, 1e-13, False)])
def test_gpu_sparsified_path_is_exact_at_tiny_tol(cfg, side, tol, drops):
    P = make_problem(cfg, side)
    T = build_tree(P, 64, tol=tol, skip=0)
    T.factorize()
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, tol, 0)
    O.factorize()
    cg, co = T.clusters(), O.clusters()
    assert np.array_equal(cg[:, 12], co[:, 12])  # the oracle's ranks
    sp = cg[cg[:, 12] >= 0]
    assert len(sp) > 0
    if drops:
        tr = O.trace()
        rr = tr[tr[:, 1] == 6]
        assert rr[:, 4].sum() < rr[:, 2].sum()
    cp = T.cperm()
    b = P.b[cp]
    xe = spl.spsolve(P.A[cp][:, cp].tocsc(), b)
    x = T.solve(b)
    assert rel(x, xe) < 1e-10


# ---------------------------------------------------------------- rank-path kernels
def gauss_oracle(rows, cols):
    om = np.zeros(rows * cols)
    olib().orc_gauss(rows, cols, SEED, om)
    return unF(om, rows, cols)


@pytest.mark.parametrize("prec", [64, 32])
def test_gauss_fill_matches_oracle(prec):
    dt = np.float64 if prec == 64 else np.float32
    rows, cols, ld = 300, 70, 310
    buf = G.DeviceBuffer(ld * cols * np.dtype(dt).itemsize)
    assert G.lib().spaqr_gauss_fill(prec, buf.ptr, rows, cols, ld, SEED, None) == 0
    got = unF(buf.download(dt, ld * cols), ld, cols)[:rows]
    assert rel(got, gauss_oracle(rows, cols)) < (1e-13 if prec == 64 else 1e-6)


def _spectrum(rng, k, w, r0, floor):
    U, _ = np.linalg.qr(rng.standard_normal((k, min(k, w))))
    V, _ = np.linalg.qr(rng.standard_normal((w, min(k, w))))
    s = np.full(min(k, w), floor)
    s[:r0] = 1.0 + rng.uniform(0, 1, r0)
    return (U * s) @ V.T


@pytest.mark.parametrize("prec", [64, 32])
def test_colnorm_and_sketch_match_oracle(prec):
    dt = np.float64 if prec == 64 else np.float32
    es = np.dtype(dt).itemsize
    rng = np.random.default_rng(7)
    shapes = [(37, 150, 21), (130, 400, 70), (8, 9, 8), (300, 1000, 130)]
    wmax = max(w for _, w, _ in shapes)
    om = gauss_oracle(wmax, 140)
    omb = dbuf(F(om), dt)
    keep, cds, sds, refs = [], [], [], []
    for (k, w, s) in shapes:
        Cm = rng.standard_normal((k, w))
        cb, cnb, mb, yb = dbuf(F(Cm), dt), G.DeviceBuffer(w * es), dbuf(np.zeros(1), dt), G.DeviceBuffer(k * s * es)
        keep += [cb, cnb, mb, yb]
        cds.append(G.ColNormDesc(cb.ptr, cnb.ptr, mb.ptr, k, w))
        sds.append(G.SketchDesc(cb.ptr, yb.ptr, k, w, s, 0))
        cn = np.zeros(w)
        mx = olib().orc_colnorm(k, w, F(Cm), cn)
        Y = np.zeros(k * s)
        olib().orc_sketch(k, w, s, F(Cm), F(om), wmax, Y)
        refs.append((cn, mx, unF(Y, k, s)))
    assert G.lib().spaqr_colnorm_batched(prec, len(cds), (G.ColNormDesc * len(cds))(*cds), None) == 0
    assert G.lib().spaqr_sketch_batched(prec, len(sds), (G.SketchDesc * len(sds))(*sds), omb.ptr, wmax, None) == 0
    tol = 1e-13 if prec == 64 else 2e-6
    for i, ((k, w, s), (cn, mx, Y)) in enumerate(zip(shapes, refs)):
        cb, cnb, mb, yb = keep[4 * i: 4 * i + 4]
        assert rel(cnb.download(dt, w), cn) < tol
        assert abs(float(mb.download(dt, 1)[0]) - mx) <= tol * mx
        assert rel(unF(yb.download(dt, k * s), k, s), Y) < (1e-12 if prec == 64 else 1e-5)


def _tail_case(rng, k, w, r_need):
    """B whose columns have tails above the threshold up to row r_need - 1 exactly, with a
    clear margin on both sides (so precision cannot move the answer)."""
    B = rng.standard_normal((k, w)) * 1e-6
    j = rng.integers(0, w)
    B[: r_need, :] = rng.standard_normal((r_need, w))
    if r_need > 0:
        B[r_need - 1, j] = 3.0
    return B


@pytest.mark.parametrize("prec", [64, 32])
def test_tail_select_matches_oracle(prec):
    dt = np.float64 if prec == 64 else np.float32
    rng = np.random.default_rng(8)
    cases = [(40, 300, 13, 40), (64, 64, 0, 64), (200, 128, 77, 120), (50, 20, 30, 29), (17, 5, 17, 17),
             (300, 513, 150, 300)]
    for (k, w, need, ry) in cases:
        B = _tail_case(rng, k, w, need)
        cmax2 = max(1.0, float(np.max((B * B).sum(0))))  # >= 1: the 1e-6 noise rows pass for need = 0
        tol = 1e-3
        expect = olib().orc_tail_select(k, w, F(B), ry, cmax2, tol)
        assert expect == (need if need <= ry else -1)
        bb, cb = dbuf(F(B), dt), dbuf(np.array([cmax2]), dt)
        ryb, rb = dbuf(np.array([ry], np.int32), np.int32), dbuf(np.zeros(1, np.int32), np.int32)
        # the planner's split: 128-column chunks of one interface share its rank
        ds = [G.TailDesc(bb.ptr + c0 * k * np.dtype(dt).itemsize, cb.ptr, ryb.ptr, rb.ptr, k, min(128, w - c0))
              for c0 in range(0, w, 128)]
        assert G.lib().spaqr_tail_select_batched(prec, len(ds), (G.TailDesc * len(ds))(*ds), tol, None) == 0
        got = int(rb.download(np.int32, 1)[0])
        assert got == (expect if expect >= 0 else 2 ** 31 - 1), (k, w, need, ry, got)


@pytest.mark.parametrize("prec", [64, 32])
def test_tail_lead_matches_oracle(prec):
    """The product form of the tail test: B's leading ry rows and the column norms of the full B
    (= those of C, Q orthogonal) select the oracle's rank on the full B."""
    dt = np.float64 if prec == 64 else np.float32
    rng = np.random.default_rng(18)
    for (k, w, need, ry) in [(40, 300, 13, 40), (64, 64, 0, 64), (200, 128, 77, 120), (50, 20, 30, 29),
                             (1200, 260, 500, 640)]:
        B = _tail_case(rng, k, w, need)
        cmax2 = max(1.0, float(np.max((B * B).sum(0))))
        tol = 1e-3
        expect = olib().orc_tail_select(k, w, F(B), ry, cmax2, tol)
        lead = np.ascontiguousarray(B[:ry, :])
        bb = dbuf(F(lead), dt)
        cnb = dbuf((B.astype(np.float64) ** 2).sum(0), dt)
        cb = dbuf(np.array([cmax2]), dt)
        ryb, rb = dbuf(np.array([ry], np.int32), np.int32), dbuf(np.zeros(1, np.int32), np.int32)
        isz = np.dtype(dt).itemsize
        ds = [G.TailLeadDesc(bb.ptr + c0 * ry * isz, cnb.ptr + c0 * isz, cb.ptr, ryb.ptr, rb.ptr, ry, min(128, w - c0))
              for c0 in range(0, w, 128)]
        assert G.lib().spaqr_tail_lead_batched(prec, len(ds), (G.TailLeadDesc * len(ds))(*ds), tol, None) == 0
        got = int(rb.download(np.int32, 1)[0])
        assert got == (expect if expect >= 0 else 2 ** 31 - 1), (k, w, need, ry, got)


def test_tail_gemm_matches_numpy():
    """B = M(0:r, 0:k) C on the matrix cores (FP32), uneven shapes in one launch."""
    rng = np.random.default_rng(19)
    shapes = [(70, 130, 200), (1, 33, 5), (300, 1100, 257), (64, 64, 64)]
    keep, ds, ref = [], [], []
    for (r, k, w) in shapes:
        M = rng.standard_normal((k, k)).astype(np.float32)
        Cm = rng.standard_normal((k, w)).astype(np.float32)
        mb, cb = dbuf(F(M), np.float32), dbuf(F(Cm), np.float32)
        bb = G.DeviceBuffer(4 * r * w)
        keep += [mb, cb, bb]
        ds.append(G.TailGemmDesc(mb.ptr, cb.ptr, bb.ptr, r, k, w, k))
        ref.append(M[:r].astype(np.float64) @ Cm.astype(np.float64))
    assert G.lib().spaqr_tail_gemm_batched(len(ds), (G.TailGemmDesc * len(ds))(*ds), None) == 0
    for (r, k, w), d, e, i in zip(shapes, ds, ref, range(len(ds))):
        got = unF(keep[3 * i + 2].download(np.float32, r * w), r, w)
        assert rel(got, e) < 1e-5, (r, k, w)


@pytest.mark.parametrize("prec", [64, 32])
def test_rank_features_match_oracle(prec):
    """Features from the GPU's own QR'd 16-column sketch: equal to the oracle's features of the
    same R diagonal, and (through the uniqueness of |R_jj|) to those of the oracle's QR."""
    dt = np.float64 if prec == 64 else np.float32
    es = np.dtype(dt).itemsize
    rng = np.random.default_rng(9)
    k, w, tol, level = 90, 260, 1e-2, 3
    Cm = _spectrum(rng, k, w, 30, 1e-4)
    om = gauss_oracle(w, 16)
    cb, cnb, mb = dbuf(F(Cm), dt), G.DeviceBuffer(w * es), dbuf(np.zeros(1), dt)
    yb, tb, xb = G.DeviceBuffer(k * 16 * es), G.DeviceBuffer(NB * 16 * es), G.DeviceBuffer(32 * 4)
    omb = dbuf(F(om), dt)
    assert G.lib().spaqr_colnorm_batched(prec, 1, (G.ColNormDesc * 1)(G.ColNormDesc(cb.ptr, cnb.ptr, mb.ptr, k, w)),
                                         None) == 0
    assert G.lib().spaqr_sketch_batched(prec, 1, (G.SketchDesc * 1)(G.SketchDesc(cb.ptr, yb.ptr, k, w, 16, 0)),
                                        omb.ptr, w, None) == 0
    assert G.lib().spaqr_geqrt_batched(prec, 1, (G.GeqrtDesc * 1)(G.GeqrtDesc(yb.ptr, tb.ptr, k, 16, k, 0)), None) == 0
    fd = G.FeatDesc(yb.ptr, cnb.ptr, mb.ptr, xb.ptr, k, w, 6, 4, level, 0)
    assert G.lib().spaqr_rank_features_batched(prec, 1, (G.FeatDesc * 1)(fd), tol, None) == 0
    x = xb.download(np.float32, 32)
    Yg = unF(yb.download(dt, k * 16), k, 16)
    cn = cnb.download(dt, w).astype(np.float64)
    cmax2 = float(mb.download(dt, 1)[0])
    xo = np.zeros(32, dtype=np.float32)
    olib().orc_rank_features(k, w, np.ascontiguousarray(np.abs(np.diag(Yg)), dtype=np.float64), cn, cmax2, 6, 4,
                             level, tol, xo)
    assert np.max(np.abs(x - xo)) < (1e-6 if prec == 64 else 2e-5)
    # against the oracle's own sketch and QR
    Y = np.zeros(k * 16)
    olib().orc_sketch(k, w, 16, F(Cm), F(om), w, Y)
    rd = np.abs(np.diag(np.linalg.qr(unF(Y, k, 16), mode="r")))
    cn2 = np.zeros(w)
    mx2 = olib().orc_colnorm(k, w, F(Cm), cn2)
    xo2 = np.zeros(32, dtype=np.float32)
    olib().orc_rank_features(k, w, np.ascontiguousarray(rd), cn2, mx2, 6, 4, level, tol, xo2)
    assert np.max(np.abs(x - xo2)) < (1e-6 if prec == 64 else 1e-4)


@pytest.mark.parametrize("k,w,r0", [(60, 200, 17), (90, 300, 40), (200, 700, 64)])
def test_fp32_geqp3_rank_and_factorization(k, w, r0):
    """FP32 geqp3 (the planner's sketch QR below 96 columns) on a prescribed spectrum: the
    rank of the oracle, the first pivot (the largest column), and C(:, jpvt) = Q R to FP32
    accuracy on the kept columns."""
    rng = np.random.default_rng(k + w)
    Cm = _spectrum(rng, k, w, r0, 1e-6)
    tol = 1e-3
    c = F(Cm)
    jp = np.zeros(w, dtype=np.int32)
    tau = np.zeros(min(k, w))
    olib().orc_geqp3(k, w, c, jp, tau)
    Cf = unF(c, k, w)
    r = olib().orc_choose_rank(np.ascontiguousarray(np.diag(Cf[:, : min(k, w)])), min(k, w), tol)
    assert r == r0
    kmax = min(k, w)
    b = dict(C=dbuf(F(Cm), np.float32), V=G.DeviceBuffer(4 * k * kmax), tau=G.DeviceBuffer(4 * kmax),
             T=G.DeviceBuffer(4 * NB * kmax), vn=G.DeviceBuffer(8 * w), jp=G.DeviceBuffer(4 * w), rank=G.DeviceBuffer(4))
    d = G.Geqp3Desc(b["C"].ptr, b["V"].ptr, b["tau"].ptr, b["T"].ptr, b["vn"].ptr, b["vn"].ptr + 4 * w, b["jp"].ptr,
                    b["rank"].ptr, k, w, 0, 0)
    assert G.lib().spaqr_geqp3_batched(32, 1, (G.Geqp3Desc * 1)(d), tol, 1e-5, None) == 0
    rg = int(b["rank"].download(np.int32, 1)[0])
    assert rg == r
    jg = b["jp"].download(np.int32, w)[:rg]
    assert jg[0] == jp[0]
    V = unF(b["V"].download(np.float32, k * rg), k, rg).astype(np.float64)
    tg = b["tau"].download(np.float32, rg).astype(np.float64)
    R = np.triu(V[:rg, :rg])
    Q = np.eye(k)[:, :rg].copy()
    for j in range(rg - 1, -1, -1):  # Q = H_0 ... H_{r-1} applied to the first r unit vectors
        v = np.concatenate([np.zeros(j), [1.0], V[j + 1:, j]])
        Q -= tg[j] * np.outer(v, v @ Q)
    assert rel(Q @ R, Cm[:, jg]) < 1e-5
    assert np.max(np.abs(Q.T @ Q - np.eye(rg))) < 1e-5
    assert abs(abs(R[0, 0]) - abs(Cf[0, 0])) <= 1e-5 * abs(Cf[0, 0])


# ---------------------------------------------------------------- forced geqp3 fallback
def test_forced_fallback_takes_geqp3_ranks():
    P = make_problem(3, 16)
    W = np.zeros(rank_model.load_weights().size, dtype=np.float32)
    W[-1] = -30.0  # output bias: predicted rank fraction ~ 1e-13, so every sketch has 8 columns
    Tm = build_tree(P, 32)
    Tm.set_rank_mode(1)
    Tm.set_rank_model(W)
    Tm.set_rank_retries(0)
    Tm.set_rank_log(True)
    Tm.rank_stats_reset()
    Tm.factorize()
    st = Tm.rank_stats()
    assert st["fallback"] > 0 and st["retried"] == 0
    T0 = build_tree(P, 32)  # the reference method (geqp3 on every interface)
    T0.factorize()
    _, meta = Tm.rank_log()
    orders = Tm.rank_log_clusters()
    r0 = T0.clusters()[:, 12]
    first = meta[:, 3] == meta[:, 3].min()  # the first sparsified level: identical inputs in both trees
    fell = first & (meta[:, 4] == 0)
    assert fell.sum() > 0
    assert np.array_equal(meta[fell, 0], r0[orders[fell]])
    # and the sketched ones passed the tail test with at most 8 reflectors
    ok = first & (meta[:, 4] == 1)
    assert np.all(meta[ok, 0] <= 8)
    # GMRES still converges through the fallback ranks
    b = P.b[Tm.cperm()]
    _, res, _ = Tm.gmres(b, iters=200, tol=1e-12, restart=50, true_tol=1e-12, flexible=True)
    assert res.true_res <= 1e-12


# ---------------------------------------------------------------- precision sequence
def test_precision_change_after_assemble_is_refused():
    P = make_problem(1, 16)
    T = build_tree(P, 32)
    T.factorize()
    T.set_precision(64)  # allowed, but the assembled blocks are FP32
    rc = G.lib().spaqr_tree_factorize(T.handle)
    assert rc != 0
    assert b"assemble" in G.lib().spaqr_last_error(T.handle)
    # the FP32 factorization is still the one solve() uses
    b = P.b[T.cperm()]
    x = T.solve(b)
    assert np.all(np.isfinite(x))
    T.assemble()
    T.factorize()  # FP64 after a new assemble
    x64 = T.solve(b)
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, P.skip)
    O.factorize()
    assert rel(x64, O.solve(b)) < 1e-10
