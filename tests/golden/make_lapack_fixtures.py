"""Writes tests/golden/lapack_fixtures.npz: inputs and outputs of the LAPACK routines the
reference calls on its hot path (dgeqrt, dtpqrt, dtpmqrt, dgeqp3), computed with the
LAPACK bundled in scipy.  These pin the oracle's restatements of the same routines.

Run from the repository root: python tests/golden/make_lapack_fixtures.py
"""
import os

import numpy as np
import scipy.linalg.lapack as la

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lapack_fixtures.npz")


def main():
    rng = np.random.default_rng(20240601)
    data = {}
    # dgeqrt(nb, A): the diagonal-block QR (src/ptree.cpp:187 uses nb = k; T is nb x k)
    for i, (m, n, nb) in enumerate([(12, 12, 12), (40, 25, 25), (33, 33, 32), (70, 70, 32), (9, 5, 5)]):
        A = rng.standard_normal((m, n))
        a, t, info = la.dgeqrt(nb, A)
        assert info == 0
        data["geqrt%d_in" % i] = A
        data["geqrt%d_nb" % i] = np.array([nb])
        data["geqrt%d_a" % i] = a
        data["geqrt%d_t" % i] = t
    # dtpqrt(l=0, nb, R, B) (src/ptree.cpp:206, nb = min(32, k))
    for i, (m, n) in enumerate([(20, 10), (7, 33), (50, 45)]):
        nb = min(32, n)
        R = np.triu(rng.standard_normal((n, n))) + 2 * np.eye(n)
        B = rng.standard_normal((m, n))
        a, b, t, info = la.dtpqrt(0, nb, R, B)
        assert info == 0
        data["tpqrt%d_R" % i] = R
        data["tpqrt%d_B" % i] = B
        data["tpqrt%d_a" % i] = a
        data["tpqrt%d_b" % i] = b
        data["tpqrt%d_t" % i] = t
        # dtpmqrt('L','T', l=0) on [A; C] with the V, T just produced (src/ptree.cpp:248)
        Atop = rng.standard_normal((n, 11))
        Cb = rng.standard_normal((m, 11))
        a2, b2, info = la.dtpmqrt(0, b, t, Atop, Cb, side="L", trans="T")
        assert info == 0
        data["tpmqrt%d_A" % i] = Atop
        data["tpmqrt%d_C" % i] = Cb
        data["tpmqrt%d_a" % i] = a2
        data["tpmqrt%d_b" % i] = b2
    # dgeqp3 on matrices with a decaying spectrum (src/util.cpp:501-512)
    for i, (m, n, decay) in enumerate([(15, 60, 0.6), (30, 30, 0.8), (8, 3, 0.5), (25, 120, 0.7)]):
        U, _ = np.linalg.qr(rng.standard_normal((m, m)))
        V, _ = np.linalg.qr(rng.standard_normal((n, min(m, n))))
        s = decay ** np.arange(min(m, n))
        A = (U[:, : len(s)] * s) @ V.T
        qr, jpvt, tau, work, info = la.dgeqp3(A)
        assert info == 0
        data["geqp3%d_in" % i] = A
        data["geqp3%d_qr" % i] = qr
        data["geqp3%d_jpvt" % i] = jpvt - 1
        data["geqp3%d_tau" % i] = tau
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, len(data), "arrays")


if __name__ == "__main__":
    main()
