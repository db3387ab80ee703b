"""GPU parity of each batched kernel against the oracle's LAPACK-semantics routine.

Every call goes through the C ABI (include/spaqr_gpu.h) with several tasks of uneven
sizes in one grouped launch.  FP64 tolerance: 1e-12 relative (different summation order
only); FP32 tolerance: 1e-4 relative.
"""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import F, lib as olib, unF
from pspaqr_public_amd import _lib as G

pytestmark = pytest.mark.gpu

NB = 32


def dbuf(a, dtype=np.float64):
    return G.DeviceBuffer.from_array(np.asarray(a, dtype=dtype))


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def oracle_geqrt(A):
    m, n = A.shape
    a = F(A)
    T = np.zeros(NB * n)
    olib().orc_geqrt(m, n, NB, a, T)
    return unF(a, m, n), unF(T, NB, n)


@pytest.mark.parametrize("prec", [64, 32])
def test_geqrt_grouped(prec):
    rng = np.random.default_rng(1)
    # (1500, 64) and (2600, 40): FP32 panels taller than the register panel (streamed panel)
    shapes = [(5, 3), (40, 40), (100, 70), (33, 33), (64, 64), (129, 97), (1, 1), (300, 31), (1500, 64), (2600, 40)]
    dt = np.float64 if prec == 64 else np.float32
    bufs, descs, ref = [], [], []
    for (m, n) in shapes:
        A = rng.standard_normal((m, n))
        Ab = dbuf(F(A), dt)
        Tb = G.DeviceBuffer(NB * n * np.dtype(dt).itemsize)
        bufs += [Ab, Tb]
        descs.append(G.GeqrtDesc(Ab.ptr, Tb.ptr, m, n, m, 0))
        ref.append(oracle_geqrt(A))
    arr = (G.GeqrtDesc * len(descs))(*descs)
    assert G.lib().spaqr_geqrt_batched(prec, len(descs), arr, None) == 0
    tol = 1e-12 if prec == 64 else 2e-4
    for (m, n), (Ar, Tr), i in zip(shapes, ref, range(len(shapes))):
        A = unF(bufs[2 * i].download(dt, m * n), m, n)
        T = unF(bufs[2 * i + 1].download(dt, NB * n), NB, n)
        assert rel(A, Ar) < tol, (m, n)
        # T: only the upper triangle of each 32x32 panel block is defined
        for p0 in range(0, min(m, n), NB):
            nb = min(NB, min(m, n) - p0)
            assert rel(np.triu(T[:nb, p0:p0 + nb]), np.triu(Tr[:nb, p0:p0 + nb])) < tol


@pytest.mark.parametrize("prec", [64, 32])
def test_larfb_and_trsm(prec):
    rng = np.random.default_rng(2)
    dt = np.float64 if prec == 64 else np.float32
    # FP32 blocks of m <= 512 and n <= 128 run on the C-resident kernel (m <= 128 / 256 / 512 buckets)
    cases = [(50, 20, 50), (70, 45, 33), (8, 3, 8), (128, 65, 128), (200, 128, 150), (257, 33, 100), (500, 70, 481),
             (700, 64, 300), (1000, 40, 96)]  # the last two: 512 < m <= 1024, n <= 64 (two-half kernel)
    keep, descs, refs = [], [], []
    for (m, n, k) in cases:
        A = rng.standard_normal((m, k))
        V, T = oracle_geqrt(A)
        Cm = rng.standard_normal((m, n))
        c = F(Cm)
        olib().orc_gemqrt_lt(m, n, k, NB, F(V), F(T), c)
        Vb, Tb, Cb = dbuf(F(V), dt), dbuf(F(T), dt), dbuf(F(Cm), dt)
        keep += [Vb, Tb, Cb]
        descs.append(G.LarfbDesc(Vb.ptr, Tb.ptr, Cb.ptr, None, m, n, k, m, m, 0, 0, 0))
        refs.append(unF(c, m, n))
    # trsm only tasks: C (m x k2) <- C R^{-1}
    # FP32 chunks of m <= 128 rows with k2 <= 128 / 256 / 512 run on the C-resident trsm
    for (m, k2) in [(40, 40), (17, 70), (100, 33), (128, 200), (77, 481), (130, 40)]:
        # well conditioned at every size (condition number about 2), like a scaling QR's R
        R = np.triu(rng.standard_normal((k2, k2)), 1) / np.sqrt(k2) + np.diag(3 + np.abs(rng.standard_normal(k2)))
        Cm = rng.standard_normal((m, k2))
        c = F(Cm)
        olib().orc_trsm_right_upper(m, k2, F(R), c)
        Rb, Cb = dbuf(F(R), dt), dbuf(F(Cm), dt)
        keep += [Rb, Cb]
        descs.append(G.LarfbDesc(None, None, Cb.ptr, Rb.ptr, m, k2, 0, 1, m, k2, k2, 0))
        refs.append(unF(c, m, k2))
    arr = (G.LarfbDesc * len(descs))(*descs)
    assert G.lib().spaqr_larfb_batched(prec, len(descs), arr, None) == 0
    tol = 1e-12 if prec == 64 else 2e-4
    cbufs = [keep[3 * i + 2] for i in range(len(cases))] + [keep[3 * len(cases) + 2 * j + 1] for j in range(3)]
    for d, Cb, ref in zip(descs, cbufs, refs):
        got = unF(Cb.download(dt, d.m * d.n), d.m, d.n)
        assert rel(got, ref) < tol


@pytest.mark.parametrize("prec", [64, 32])
def test_tpqrt_tpmqrt_chain(prec):
    rng = np.random.default_rng(3)
    dt = np.float64 if prec == 64 else np.float32
    tol = 1e-12 if prec == 64 else 5e-4
    # FP32 chains with links of <= 128 / 256 / 512 rows run on the C-resident kernel
    # links of 1100 and 2500 rows: FP32 pentagonal panels taller than the register panel
    fronts = [(45, [30, 7, 64]), (8, [5, 9]), (70, [100]), (40, [300, 20]), (66, [200, 150]), (50, [700, 30, 1000]),
              (40, [1100, 2500])]
    keep = []
    tq_desc, tq_links, tm_desc, tm_links = [], [], [], []
    expect_R, expect_B, expect_top, expect_C = [], [], [], []
    for (k, rows) in fronts:
        R0 = np.triu(rng.standard_normal((k, k))) + 3 * np.eye(k)
        Bs = [rng.standard_normal((r, k)) for r in rows]
        ncols = 37
        top = rng.standard_normal((k, ncols))
        Cs = [rng.standard_normal((r, ncols)) for r in rows]
        # oracle: chain of tpqrt, then the chain of tpmqrt on (top, C_j)
        Ro = F(R0)
        Vs, Ts = [], []
        for B in Bs:
            b = F(B)
            T = np.zeros(NB * k)
            olib().orc_tpqrt(B.shape[0], k, NB, Ro, b, T)
            Vs.append(b)
            Ts.append(T)
        to = F(top)
        cos = []
        for B, V, T, Cm in zip(Bs, Vs, Ts, Cs):
            c = F(Cm)
            olib().orc_tpmqrt_lt(B.shape[0], ncols, k, NB, V, T, to, c)
            cos.append(c)
        expect_R.append(unF(Ro, k, k))
        expect_B.append([unF(v, b.shape[0], k) for v, b in zip(Vs, Bs)])
        expect_top.append(unF(to, k, ncols))
        expect_C.append([unF(c, cc.shape[0], ncols) for c, cc in zip(cos, Cs)])
        # device
        Rb = dbuf(F(R0), dt)
        keep.append(Rb)
        l0 = len(tq_links)
        Bbufs, Tbufs = [], []
        for B in Bs:
            Bb = dbuf(F(B), dt)
            Tb = G.DeviceBuffer(NB * k * np.dtype(dt).itemsize)
            Bbufs.append(Bb)
            Tbufs.append(Tb)
            tq_links.append(G.TpqrtLink(Bb.ptr, Tb.ptr, B.shape[0], B.shape[0]))
        keep += Bbufs + Tbufs
        tq_desc.append(G.TpqrtDesc(Rb.ptr, k, k, l0, len(Bs)))
        topb = dbuf(F(top), dt)
        Cbufs = [dbuf(F(Cm), dt) for Cm in Cs]
        keep += [topb] + Cbufs
        m0 = len(tm_links)
        for Bb, Tb, Cb, B in zip(Bbufs, Tbufs, Cbufs, Bs):
            tm_links.append(G.TpmqrtLink(Bb.ptr, Tb.ptr, Cb.ptr, B.shape[0], B.shape[0], B.shape[0], 0))
        # two column tiles: [0, 20) and [20, 37)
        tm_desc.append(G.TpmqrtDesc(topb.ptr, k, k, 20, m0, len(Bs), 0))
        tm_desc.append(G.TpmqrtDesc(topb.ptr + 20 * k * np.dtype(dt).itemsize, k, k, 17, m0, len(Bs), 20))
        fronts_dev = (Rb, Bbufs, topb, Cbufs)
        keep.append(fronts_dev)
    ta = (G.TpqrtDesc * len(tq_desc))(*tq_desc)
    tl = (G.TpqrtLink * len(tq_links))(*tq_links)
    assert G.lib().spaqr_tpqrt_batched(prec, len(tq_desc), ta, len(tq_links), tl, None) == 0
    ma = (G.TpmqrtDesc * len(tm_desc))(*tm_desc)
    ml = (G.TpmqrtLink * len(tm_links))(*tm_links)
    assert G.lib().spaqr_tpmqrt_batched(prec, len(tm_desc), ma, len(tm_links), ml, None) == 0
    devs = [x for x in keep if isinstance(x, tuple)]
    for (k, rows), (Rb, Bbufs, topb, Cbufs), eR, eB, eT, eC in zip(fronts, devs, expect_R, expect_B, expect_top,
                                                                   expect_C):
        assert rel(np.triu(unF(Rb.download(dt, k * k), k, k)), np.triu(eR)) < tol
        for Bb, r, e in zip(Bbufs, rows, eB):
            assert rel(unF(Bb.download(dt, r * k), r, k), e) < tol
        assert rel(unF(topb.download(dt, k * 37), k, 37), eT) < tol
        for Cb, r, e in zip(Cbufs, rows, eC):
            assert rel(unF(Cb.download(dt, r * 37), r, 37), e) < tol


def _spectrum_matrix(rng, k, w, decay):
    U, _ = np.linalg.qr(rng.standard_normal((k, k)))
    Vt, _ = np.linalg.qr(rng.standard_normal((w, min(k, w))))
    s = decay ** np.arange(min(k, w))
    return (U[:, : len(s)] * s) @ Vt.T


@pytest.mark.parametrize("prec", [64])
def test_geqp3_rank_and_reflectors(prec):
    rng = np.random.default_rng(4)
    dt = np.float64
    cases = [(30, 200, 0.7, 1e-2), (64, 64, 0.5, 1e-3), (10, 3, 0.9, 1e-2), (100, 400, 0.8, 1e-2), (7, 50, 1.0, 1e-2),
             (40, 300, 0.9, 0.0)]
    keep, descs, refs = [], [], []
    for (k, w, decay, tol) in cases:
        Cm = _spectrum_matrix(rng, k, w, decay)
        c = F(Cm)
        jp = np.zeros(w, dtype=np.int32)
        tau = np.zeros(min(k, w))
        olib().orc_geqp3(k, w, c, jp, tau)
        Cf = unF(c, k, w)
        r = olib().orc_choose_rank(np.ascontiguousarray(np.diag(Cf[:, : min(k, w)])), min(k, w), tol)
        refs.append((r, Cf[:, :r].copy(), tau[:r].copy(), jp[:r].copy()))
        kmax = max(1, min(k, w))
        b = dict(C=dbuf(F(Cm)), V=G.DeviceBuffer(8 * k * kmax), tau=G.DeviceBuffer(8 * kmax),
                 T=G.DeviceBuffer(8 * NB * kmax), vn=G.DeviceBuffer(16 * w), jp=G.DeviceBuffer(4 * w),
                 rank=G.DeviceBuffer(4))
        keep.append((b, tol))
        descs.append(G.Geqp3Desc(b["C"].ptr, b["V"].ptr, b["tau"].ptr, b["T"].ptr, b["vn"].ptr, b["vn"].ptr + 8 * w,
                                 b["jp"].ptr, b["rank"].ptr, k, w, 0, 0))
    # one call per tolerance (the tolerance is a launch argument)
    for i, (d, (b, tol)) in enumerate(zip(descs, keep)):
        arr = (G.Geqp3Desc * 1)(d)
        assert G.lib().spaqr_geqp3_batched(prec, 1, arr, tol, 0.0, None) == 0
    for (k, w, _, _), (b, tol), (r, Vr, taur, jpr) in zip(cases, keep, refs):
        rg = int(b["rank"].download(np.int32, 1)[0])
        assert rg == r, (k, w, rg, r)
        if r == 0:
            continue
        V = unF(b["V"].download(dt, k * r), k, r)
        assert rel(V, Vr) < 1e-11
        assert rel(b["tau"].download(dt, r), taur) < 1e-11
        assert np.array_equal(b["jp"].download(np.int32, w)[:r], jpr)
        T = unF(b["T"].download(dt, NB * r), NB, r)
        for p0 in range(0, r, NB):
            nb = min(NB, r - p0)
            Tr = np.zeros(nb * nb)
            Vp = np.ascontiguousarray(F(Vr[p0:, p0:p0 + nb]))
            olib().orc_larft(k - p0, nb, Vp, np.ascontiguousarray(taur[p0:p0 + nb]), Tr)
            assert rel(np.triu(T[:nb, p0:p0 + nb]), np.triu(unF(Tr, nb, nb))) < 1e-11


@pytest.mark.parametrize("prec,shape", [(64, (40, 50, 17, 33)), (32, (40, 50, 17, 33)), (64, (300, 200, 150, 100)),
                                        (32, (300, 200, 150, 100)), (32, (120, 480, 60, 128))])
def test_spars_apply_matches_oracle(prec, shape):
    """Right transform A Q_n, left transform Q_c^T on A[:, :rn], scatter of the first rc rows.
    FP64: one task with both phases; FP32: the right phase in 128-row chunks, then the left
    phase, as the level planner issues them (the C-resident kernels' shapes)."""
    rng = np.random.default_rng(5)
    dt = np.float64 if prec == 64 else np.float32
    kc, kn, rc, rn = shape
    A = rng.standard_normal((kc, kn))
    Vc, Tc = oracle_geqrt(rng.standard_normal((kc, kc)))
    Vn, Tn = oracle_geqrt(rng.standard_normal((kn, kn)))
    # oracle: A Q_n (first rn reflectors, full T via larft), then Q_c^T on A[:, :rn]
    Vn_r = np.tril(Vn[:, :rn], -1) + np.eye(kn, rn)
    taun = np.array([Tn[i % NB, i] for i in range(rn)])
    Tfull_n = np.zeros(rn * rn)
    olib().orc_larft(kn, rn, F(Vn_r), taun, Tfull_n)
    a = F(A)
    olib().orc_larfb_rn(kc, kn, rn, F(Vn_r), Tfull_n, a)
    Ar = unF(a, kc, kn)
    Vc_r = np.tril(Vc[:, :rc], -1) + np.eye(kc, rc)
    tauc = np.array([Tc[i % NB, i] for i in range(rc)])
    Tfull_c = np.zeros(rc * rc)
    olib().orc_larft(kc, rc, F(Vc_r), tauc, Tfull_c)
    a11 = F(Ar[:, :rn])
    olib().orc_larfb_lt(kc, rn, rc, F(Vc_r), Tfull_c, a11)
    expect = unF(a11, kc, rn)[:rc, :]
    es = np.dtype(dt).itemsize
    Ab, Vnb, Tnb, Vcb, Tcb = dbuf(F(A), dt), dbuf(F(Vn), dt), dbuf(F(Tn), dt), dbuf(F(Vc), dt), dbuf(F(Tc), dt)
    ldd = rc + 5
    dst = dbuf(np.zeros(ldd * (rn + 3)), dt)
    dptr = dst.ptr + es * (2 + ldd * 1)
    if prec == 64:
        ds = [G.SparsDesc(Ab.ptr, Vnb.ptr, Tnb.ptr, Vcb.ptr, Tcb.ptr, dptr, kc, kn, kc, kn, rn, kc, rc, kn, kc, ldd, 0, 0)]
        assert G.lib().spaqr_spars_apply_batched(prec, len(ds), (G.SparsDesc * len(ds))(*ds), None) == 0
    else:
        right = [G.SparsDesc(Ab.ptr + es * r0, Vnb.ptr, Tnb.ptr, None, None, None, min(128, kc - r0), kn, kc, kn, rn, kc,
                             rc, kn, kc, ldd, 0, 1) for r0 in range(0, kc, 128)]
        assert G.lib().spaqr_spars_apply_batched(prec, len(right), (G.SparsDesc * len(right))(*right), None) == 0
        left = [G.SparsDesc(Ab.ptr + es * kc * c0, None, None, Vcb.ptr, Tcb.ptr, dptr + es * ldd * c0, kc, kn, kc, kn,
                            min(128, rn - c0), kc, rc, kn, kc, ldd, 0, 2) for c0 in range(0, rn, 128)]
        assert G.lib().spaqr_spars_apply_batched(prec, len(left), (G.SparsDesc * len(left))(*left), None) == 0
    out = unF(dst.download(dt, ldd * (rn + 3)), ldd, rn + 3)
    assert rel(out[2:2 + rc, 1:1 + rn], expect) < (1e-12 if prec == 64 else 2e-4)
    # untouched border stays zero
    assert np.all(out[:2, :] == 0) and np.all(out[:, 0] == 0)


def test_geqp3_tie_band_matches_oracle():
    """Exactly tied columns (duplicates): with the tie band the GPU and the oracle choose
    the same pivots even though their partial norms differ in the last bits."""
    rng = np.random.default_rng(12)
    base = _spectrum_matrix(rng, 24, 40, 0.8)
    A = np.concatenate([base, base[:, ::-1], base[:, 5:15]], axis=1)
    k, w = A.shape
    a = F(A)
    jp = np.zeros(w, dtype=np.int32)
    tau = np.zeros(k)
    olib().orc_geqp3_tie(k, w, a, jp, tau, 1e-10)
    Cf = unF(a, k, w)
    r = olib().orc_choose_rank(np.ascontiguousarray(np.diag(Cf[:, :k])), k, 1e-2)
    b = dict(C=dbuf(F(A)), V=G.DeviceBuffer(8 * k * k), tau=G.DeviceBuffer(8 * k), T=G.DeviceBuffer(8 * NB * k),
             vn=G.DeviceBuffer(16 * w), jp=G.DeviceBuffer(4 * w), rank=G.DeviceBuffer(4))
    d = G.Geqp3Desc(b["C"].ptr, b["V"].ptr, b["tau"].ptr, b["T"].ptr, b["vn"].ptr, b["vn"].ptr + 8 * w, b["jp"].ptr,
                    b["rank"].ptr, k, w, 0, 0)
    assert G.lib().spaqr_geqp3_batched(64, 1, (G.Geqp3Desc * 1)(d), 1e-2, 1e-10, None) == 0
    assert int(b["rank"].download(np.int32, 1)[0]) == r
    assert np.array_equal(b["jp"].download(np.int32, w)[:r], jp[:r])
    assert rel(unF(b["V"].download(np.float64, k * r), k, r), Cf[:, :r]) < 1e-12


def test_geqp3_hbm_slices_match_oracle():
    """An interface too tall for its column slices to sit in LDS (k = 6000 rows in FP64: one
    column is 47 KB) runs the cooperative geqp3 with the slices left in HBM; same rank, pivots,
    reflectors and T as the oracle."""
    rng = np.random.default_rng(21)
    k, w, tol = 6000, 300, 1e-2
    U, _ = np.linalg.qr(rng.standard_normal((k, w)))
    Vt, _ = np.linalg.qr(rng.standard_normal((w, w)))
    Cm = (U * 0.93 ** np.arange(w)) @ Vt.T
    c = F(Cm)
    jp = np.zeros(w, dtype=np.int32)
    tau = np.zeros(min(k, w))
    olib().orc_geqp3(k, w, c, jp, tau)
    Cf = unF(c, k, w)
    r = olib().orc_choose_rank(np.ascontiguousarray(np.diag(Cf[:, :w])), w, tol)
    assert 20 < r < w
    b = dict(C=dbuf(F(Cm)), V=G.DeviceBuffer(8 * k * w), tau=G.DeviceBuffer(8 * w), T=G.DeviceBuffer(8 * NB * w),
             vn=G.DeviceBuffer(16 * w), jp=G.DeviceBuffer(4 * w), rank=G.DeviceBuffer(4))
    d = G.Geqp3Desc(b["C"].ptr, b["V"].ptr, b["tau"].ptr, b["T"].ptr, b["vn"].ptr, b["vn"].ptr + 8 * w, b["jp"].ptr,
                    b["rank"].ptr, k, w, 0, 0)
    assert G.lib().spaqr_geqp3_batched(64, 1, (G.Geqp3Desc * 1)(d), tol, 0.0, None) == 0
    assert int(b["rank"].download(np.int32, 1)[0]) == r
    assert np.array_equal(b["jp"].download(np.int32, w)[:r], jp[:r])
    assert rel(unF(b["V"].download(np.float64, k * r), k, r), Cf[:, :r]) < 1e-11
    assert rel(b["tau"].download(np.float64, r), tau[:r]) < 1e-11


def test_full_T_from_the_blocked_geqrt():
    """The reference's geqrt_cluster asks dgeqrt for nb = k, one k x k T (src/ptree.cpp:187);
    spaqr_geqrt_batched returns the NB = 32 blocked T.  A caller that needs the k x k T builds it
    with larft's recurrence over all k reflectors (tau_j = T's diagonal), and that T satisfies
    I - V T V^T = H_0 H_1 ... H_{k-1}, the product of the reflectors -- checked here on the GPU's
    V and blocked T for a square block and a tall one."""
    rng = np.random.default_rng(5)
    for (m, n) in [(96, 96), (150, 70)]:
        A = rng.standard_normal((m, n))
        Ab = dbuf(F(A))
        Tb = G.DeviceBuffer(NB * n * 8)
        d = (G.GeqrtDesc * 1)(G.GeqrtDesc(Ab.ptr, Tb.ptr, m, n, m, 0))
        assert G.lib().spaqr_geqrt_batched(64, 1, d, None) == 0
        VR = unF(Ab.download(np.float64, m * n), m, n)
        Tblk = unF(Tb.download(np.float64, NB * n), NB, n)
        k = min(m, n)
        V = np.tril(VR[:, :k], -1) + np.eye(m, k)
        tau = np.array([Tblk[j % NB, j] for j in range(k)])
        # larft (forward, columnwise): T(0:j, j) = -tau_j T(0:j, 0:j) V(:, 0:j)^T v_j
        Tf = np.zeros((k, k))
        for j in range(k):
            Tf[j, j] = tau[j]
            if j:
                Tf[:j, j] = -tau[j] * (Tf[:j, :j] @ (V[:, :j].T @ V[:, j]))
        Q = np.eye(m)
        for j in range(k):
            v = V[:, j:j + 1]
            Q = Q @ (np.eye(m) - tau[j] * (v @ v.T))
        assert np.linalg.norm(np.eye(m) - V @ Tf @ V.T - Q) < 1e-12 * m
        # and each 32 x 32 diagonal block of the full T is the blocked T's panel block
        for p0 in range(0, k, NB):
            nb = min(NB, k - p0)
            assert rel(np.triu(Tf[p0:p0 + nb, p0:p0 + nb]), np.triu(Tblk[:nb, p0:p0 + nb])) < 1e-12
        # Q^T A = R
        R = np.triu(VR[:k, :n])
        assert np.linalg.norm(Q.T @ A - np.vstack([R, np.zeros((m - k, n))])) < 1e-11 * np.linalg.norm(A)
