"""spaqr_geqp3_batched_noalloc against spaqr_geqp3_batched: identical ranks, pivots,
reflectors, tau and T (the same kernels run on the same inputs, so bit for bit), on two calls in
a row, the second smaller; and a call whose scratch would exceed the fixed array is refused
before anything runs."""
import numpy as np
import pytest

from pspaqr_public_amd import _lib as G

pytestmark = pytest.mark.gpu

NB = 32


def _spectrum(rng, k, w, decay):
    U, _ = np.linalg.qr(rng.standard_normal((k, min(k, w))))
    V, _ = np.linalg.qr(rng.standard_normal((w, min(k, w))))
    return (U * decay ** np.arange(min(k, w))) @ V.T


class Task:
    def __init__(self, C, dt):
        k, w = C.shape
        es = np.dtype(dt).itemsize
        kmax = max(1, min(k, w))
        self.k, self.w, self.dt = k, w, dt
        self.C = G.DeviceBuffer.from_array(np.asfortranarray(C).ravel(order="F").astype(dt))
        self.V = G.DeviceBuffer.from_array(np.zeros(k * kmax, dt))
        self.tau = G.DeviceBuffer.from_array(np.zeros(kmax, dt))
        self.T = G.DeviceBuffer.from_array(np.zeros(NB * kmax, dt))
        self.vn = G.DeviceBuffer(2 * w * es)
        self.jp = G.DeviceBuffer.from_array(np.zeros(w, np.int32))
        self.rank = G.DeviceBuffer.from_array(np.full(1, -7, np.int32))
        self.desc = G.Geqp3Desc(self.C.ptr, self.V.ptr, self.tau.ptr, self.T.ptr, self.vn.ptr, self.vn.ptr + w * es,
                                self.jp.ptr, self.rank.ptr, k, w, 0, 0)

    def result(self):
        k, w, dt = self.k, self.w, self.dt
        r = int(self.rank.download(np.int32, 1)[0])
        kmax = max(1, min(k, w))
        return (r, self.V.download(dt, k * kmax)[: k * max(r, 0)], self.tau.download(dt, kmax)[: max(r, 0)],
                self.T.download(dt, NB * kmax)[: NB * max(r, 0)], self.jp.download(np.int32, w)[: max(r, 0)])


def _run(fn, prec, mats, dt, tol):
    tasks = [Task(C, dt) for C in mats]
    arr = (G.Geqp3Desc * len(tasks))(*[t.desc for t in tasks])
    rc = fn(prec, len(tasks), arr, tol, 1e-10 if prec == 64 else 1e-5, None)
    return rc, [t.result() for t in tasks]


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_geqp3_equals_allocating_entry_point(prec):
    dt = np.float64 if prec == 64 else np.float32
    rng = np.random.default_rng(31 + prec)
    tol = 1e-2
    first = [_spectrum(rng, k, w, d) for (k, w, d) in
             [(30, 200, 0.7), (64, 64, 0.5), (100, 400, 0.8), (7, 50, 1.0), (600, 900, 0.97), (300, 2500, 0.98)]]
    second = [_spectrum(rng, k, w, d) for (k, w, d) in [(20, 90, 0.7), (150, 300, 0.9)]]
    L = G.lib()
    for call, mats in (("first", first), ("second, smaller", second)):
        print("[no-alloc geqp3] prec %d, %s call: %d tasks" % (prec, call, len(mats)), flush=True)
        rc_a, ref = _run(L.spaqr_geqp3_batched, prec, mats, dt, tol)
        rc_b, got = _run(L.spaqr_geqp3_batched_noalloc, prec, mats, dt, tol)
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b, L.spaqr_last_error(None))
        for (ra, Va, ta, Ta, ja), (rb, Vb, tb, Tb, jb) in zip(ref, got):
            assert ra == rb and ra > 0
            assert np.array_equal(ja, jb)
            assert np.array_equal(Va, Vb) and np.array_equal(ta, tb) and np.array_equal(Ta, Tb)


def test_noalloc_geqp3_refuses_calls_beyond_its_scratch():
    """Nine 2000 x 2000 FP64 interfaces need about 2 MB of exchange slots each (cooperative
    geqp3 with the column slices in HBM), more than the 16 MiB array: the call returns -6 and
    leaves every output untouched."""
    L = G.lib()
    assert L.spaqr_geqp3_noalloc_scratch_bytes() == 16 << 20
    rng = np.random.default_rng(5)
    C = rng.standard_normal((2000, 2000))
    tasks = [Task(C, np.float64) for _ in range(9)]
    arr = (G.Geqp3Desc * len(tasks))(*[t.desc for t in tasks])
    print("[no-alloc geqp3] oversized call", flush=True)
    rc = L.spaqr_geqp3_batched_noalloc(64, len(tasks), arr, 1e-2, 1e-10, None)
    assert rc == -6
    assert b"MiB" in L.spaqr_last_error(None)
    for t in tasks:
        assert int(t.rank.download(np.int32, 1)[0]) == -7
    # a call that fits still works afterwards
    rc, res = _run(L.spaqr_geqp3_batched_noalloc, 64, [_spectrum(rng, 50, 120, 0.8)], np.float64, 1e-2)
    assert rc == 0 and res[0][0] > 0
