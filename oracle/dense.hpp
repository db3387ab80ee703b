// Oracle dense kernels: plain FP64 restatements of the LAPACK/BLAS routines that the
// pspaQR reference calls on its hot path.  TEST INFRASTRUCTURE ONLY: these routines are
// the checker that the HIP kernels are compared against; nothing in the product links them.
//
// Every routine follows the published LAPACK definition of the routine named in its
// comment (column-major storage, leading dimensions as in LAPACK).  Parity of each
// routine is pinned in tests/test_oracle_dense.py against scipy's bundled LAPACK
// (dgeqrt, dtpqrt, dtpmqrt, dgeqp3) and against closed forms (Q orthogonality,
// Q*R == A, brute-force pivoting on tiny inputs).
#pragma once
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <vector>

namespace orc {

// Dense column-major matrix with lda == rows (the layout of Eigen::MatrixXd that the
// reference keeps in every block, include/edge.h:18).
struct Mat {
    int r = 0, c = 0;
    std::vector<double> a;
    Mat() {}
    Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, 0.0) {}
    double* data() { return a.data(); }
    const double* data() const { return a.data(); }
    double& operator()(int i, int j) { return a[i + (size_t)j * r]; }
    double operator()(int i, int j) const { return a[i + (size_t)j * r]; }
    void zero(int r_, int c_) { r = r_; c = c_; a.assign((size_t)r_ * c_, 0.0); }
    // Eigen conservativeResize: keep the top-left overlap, zero the rest
    // (the reference relies on it at src/ptree.cpp:1878).
    void conservative_resize(int nr, int nc) {
        Mat b(nr, nc);
        for (int j = 0; j < std::min(c, nc); ++j)
            for (int i = 0; i < std::min(r, nr); ++i) b(i, j) = (*this)(i, j);
        *this = std::move(b);
    }
    Mat block(int i0, int j0, int nr, int nc) const {
        Mat b(nr, nc);
        for (int j = 0; j < nc; ++j)
            for (int i = 0; i < nr; ++i) b(i, j) = (*this)(i0 + i, j0 + j);
        return b;
    }
    void set_block(int i0, int j0, const Mat& b) {
        for (int j = 0; j < b.c; ++j)
            for (int i = 0; i < b.r; ++i) (*this)(i0 + i, j0 + j) = b(i, j);
    }
};

// dlarfg: generate an elementary reflector H with H^T [alpha; x] = [beta; 0].
// alpha is overwritten by beta, x by the reflector tail v(2:n); returns tau.
// (LAPACK dlarfg; the safe-minimum rescaling loop is not needed for the matrices on this
// path and is omitted, both here and in the HIP kernels.)
constexpr double kAlphaSignBand = 1e-10;

inline std::atomic<long>& larfg_tiny_alpha() {  // diagnostic counter (updated from several threads)
    static std::atomic<long> count{0};
    return count;
}

inline double larfg(int n, double& alpha, double* x, int incx) {
    if (n <= 1) return 0.0;
    double ss = 0.0;
    for (int i = 0; i < n - 1; ++i) ss += x[(size_t)i * incx] * x[(size_t)i * incx];
    if (ss == 0.0) return 0.0;
    double xnorm = std::sqrt(ss);
    // An alpha at rounding-noise level carries no sign information: treat it as >= 0 so the
    // reflector does not flip with the summation order (the HIP kernels use the same rule).
    const bool tiny = std::fabs(alpha) <= kAlphaSignBand * xnorm;
    if (tiny) larfg_tiny_alpha()++;
    double beta = tiny ? -std::hypot(alpha, xnorm) : -std::copysign(std::hypot(alpha, xnorm), alpha);
    double tau = (beta - alpha) / beta;
    double scal = 1.0 / (alpha - beta);
    for (int i = 0; i < n - 1; ++i) x[(size_t)i * incx] *= scal;
    alpha = beta;
    return tau;
}

// dlarft('F','C'): T (k x k upper) of H = H(1)...H(k) = I - V T V^T, V unit lower trapezoidal.
inline void larft(int m, int k, const double* V, int ldv, const double* tau, double* T, int ldt) {
    for (int i = 0; i < k; ++i) {
        T[i + (size_t)i * ldt] = tau[i];
        if (tau[i] == 0.0) {
            for (int j = 0; j < i; ++j) T[j + (size_t)i * ldt] = 0.0;
            continue;
        }
        // w(j) = V(:,j)^T V(:,i), j < i   (V(i,i) = 1, V(r,i) = 0 for r < i)
        std::vector<double> w(i, 0.0);
        for (int j = 0; j < i; ++j) {
            double s = V[i + (size_t)j * ldv];  // row i: V(i,j) * 1
            for (int r = i + 1; r < m; ++r) s += V[r + (size_t)j * ldv] * V[r + (size_t)i * ldv];
            w[j] = s;
        }
        // T(0:i, i) = -tau_i * T(0:i, 0:i) * w
        for (int j = 0; j < i; ++j) {
            double s = 0.0;
            for (int l = j; l < i; ++l) s += T[j + (size_t)l * ldt] * w[l];
            T[j + (size_t)i * ldt] = -tau[i] * s;
        }
    }
}

// dlarfb('L','T','F','C'): C <- H^T C = C - V T^T V^T C.  C is m x n, V m x k.
inline void larfb_left_trans(int m, int n, int k, const double* V, int ldv, const double* T, int ldt,
                             double* C, int ldc) {
    if (m == 0 || n == 0 || k == 0) return;
    std::vector<double> W((size_t)k * n);
    for (int j = 0; j < n; ++j) {
        const double* cj = C + (size_t)j * ldc;
        for (int i = 0; i < k; ++i) {
            const double* vi = V + (size_t)i * ldv;
            double s = cj[i];
            for (int r = i + 1; r < m; ++r) s += vi[r] * cj[r];
            W[i + (size_t)j * k] = s;
        }
    }
    // W <- T^T W (T upper => T^T lower): rows bottom-up in place
    for (int j = 0; j < n; ++j)
        for (int i = k - 1; i >= 0; --i) {
            double s = 0.0;
            for (int l = 0; l <= i; ++l) s += T[l + (size_t)i * ldt] * W[l + (size_t)j * k];
            W[i + (size_t)j * k] = s;
        }
    for (int j = 0; j < n; ++j) {
        double* cj = C + (size_t)j * ldc;
        for (int i = 0; i < k; ++i) {
            const double* vi = V + (size_t)i * ldv;
            double w = W[i + (size_t)j * k];
            cj[i] -= w;
            for (int r = i + 1; r < m; ++r) cj[r] -= vi[r] * w;
        }
    }
}

// dlarfb('R','N','F','C'): C <- C H = C - C V T V^T.  C is m x n, V n x k.
inline void larfb_right_notrans(int m, int n, int k, const double* V, int ldv, const double* T, int ldt,
                                double* C, int ldc) {
    if (m == 0 || n == 0 || k == 0) return;
    std::vector<double> W((size_t)m * k);
    for (int i = 0; i < k; ++i) {
        double* wi = W.data() + (size_t)i * m;
        for (int r = 0; r < m; ++r) wi[r] = C[r + (size_t)i * ldc];
        for (int l = i + 1; l < n; ++l) {
            double v = V[l + (size_t)i * ldv];
            const double* cl = C + (size_t)l * ldc;
            for (int r = 0; r < m; ++r) wi[r] += cl[r] * v;
        }
    }
    // W <- W T (T upper): columns right-to-left in place
    for (int i = k - 1; i >= 0; --i) {
        double* wi = W.data() + (size_t)i * m;
        double tii = T[i + (size_t)i * ldt];
        for (int r = 0; r < m; ++r) wi[r] *= tii;
        for (int l = 0; l < i; ++l) {
            double t = T[l + (size_t)i * ldt];
            const double* wl = W.data() + (size_t)l * m;
            for (int r = 0; r < m; ++r) wi[r] += wl[r] * t;
        }
    }
    // C -= W V^T
    for (int l = 0; l < n; ++l) {
        double* cl = C + (size_t)l * ldc;
        int lim = std::min(l, k - 1);
        for (int i = 0; i <= lim; ++i) {
            double v = (l == i) ? 1.0 : V[l + (size_t)i * ldv];
            const double* wi = W.data() + (size_t)i * m;
            for (int r = 0; r < m; ++r) cl[r] -= wi[r] * v;
        }
    }
}

// dgeqrt2: unblocked QR of an m x n panel with its n x n T.
inline void geqrt2(int m, int n, double* A, int lda, double* T, int ldt) {
    int k = std::min(m, n);
    std::vector<double> tau(k);
    for (int j = 0; j < k; ++j) {
        double& alpha = A[j + (size_t)j * lda];
        tau[j] = larfg(m - j, alpha, &A[j + 1 + (size_t)j * lda], 1);
        if (tau[j] != 0.0) {
            for (int l = j + 1; l < n; ++l) {
                double w = A[j + (size_t)l * lda];
                for (int r = j + 1; r < m; ++r) w += A[r + (size_t)j * lda] * A[r + (size_t)l * lda];
                w *= tau[j];
                A[j + (size_t)l * lda] -= w;
                for (int r = j + 1; r < m; ++r) A[r + (size_t)l * lda] -= A[r + (size_t)j * lda] * w;
            }
        }
    }
    larft(m, k, A, lda, tau.data(), T, ldt);
}

// dgeqrt(m, n, nb): blocked QR, V and R in A, T stored nb x min(m,n) (one nb x nb
// upper-triangular block per panel), exactly LAPACK's storage.
inline void geqrt(int m, int n, int nb, double* A, int lda, double* T, int ldt) {
    int k = std::min(m, n);
    for (int i = 0; i < k; i += nb) {
        int ib = std::min(k - i, nb);
        geqrt2(m - i, ib, &A[i + (size_t)i * lda], lda, &T[(size_t)i * ldt], ldt);
        if (i + ib < n)
            larfb_left_trans(m - i, n - i - ib, ib, &A[i + (size_t)i * lda], lda, &T[(size_t)i * ldt], ldt,
                             &A[i + (size_t)(i + ib) * lda], lda);
    }
}

// dtprfb('L','T','F','C') with L = 0: [A; B] <- H^T [A; B], H = I - [I; V] T [I; V]^T.
// A is k x n, B is m x n, V is m x k.
inline void tprfb_lt(int m, int n, int k, const double* V, int ldv, const double* T, int ldt,
                     double* A, int lda, double* B, int ldb) {
    if (n == 0 || k == 0) return;
    std::vector<double> W((size_t)k * n);
    for (int j = 0; j < n; ++j) {
        const double* bj = B + (size_t)j * ldb;
        for (int i = 0; i < k; ++i) {
            const double* vi = V + (size_t)i * ldv;
            double s = A[i + (size_t)j * lda];
            for (int r = 0; r < m; ++r) s += vi[r] * bj[r];
            W[i + (size_t)j * k] = s;
        }
    }
    for (int j = 0; j < n; ++j)
        for (int i = k - 1; i >= 0; --i) {
            double s = 0.0;
            for (int l = 0; l <= i; ++l) s += T[l + (size_t)i * ldt] * W[l + (size_t)j * k];
            W[i + (size_t)j * k] = s;
        }
    for (int j = 0; j < n; ++j) {
        double* bj = B + (size_t)j * ldb;
        for (int i = 0; i < k; ++i) {
            double w = W[i + (size_t)j * k];
            A[i + (size_t)j * lda] -= w;
            const double* vi = V + (size_t)i * ldv;
            for (int r = 0; r < m; ++r) bj[r] -= vi[r] * w;
        }
    }
}

// dtpqrt2 with L = 0: QR of [A (n x n upper); B (m x n)]; V overwrites B, T is n x n.
inline void tpqrt2(int m, int n, double* A, int lda, double* B, int ldb, double* T, int ldt) {
    std::vector<double> tau(n);
    for (int j = 0; j < n; ++j) {
        tau[j] = larfg(m + 1, A[j + (size_t)j * lda], &B[(size_t)j * ldb], 1);
        if (tau[j] != 0.0) {
            for (int l = j + 1; l < n; ++l) {
                double w = A[j + (size_t)l * lda];
                for (int r = 0; r < m; ++r) w += B[r + (size_t)j * ldb] * B[r + (size_t)l * ldb];
                w *= tau[j];
                A[j + (size_t)l * lda] -= w;
                for (int r = 0; r < m; ++r) B[r + (size_t)l * ldb] -= B[r + (size_t)j * ldb] * w;
            }
        }
    }
    // T: reflector i is [e_i; B(:,i)]; e_i^T e_j = 0 for i != j, so V^T V = B^T B off-diagonal.
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) T[i + (size_t)j * ldt] = 0.0;
        T[i + (size_t)i * ldt] = tau[i];
        std::vector<double> w(i, 0.0);
        for (int j = 0; j < i; ++j) {
            double s = 0.0;
            for (int r = 0; r < m; ++r) s += B[r + (size_t)j * ldb] * B[r + (size_t)i * ldb];
            w[j] = s;
        }
        for (int j = 0; j < i; ++j) {
            double s = 0.0;
            for (int l = j; l < i; ++l) s += T[j + (size_t)l * ldt] * w[l];
            T[j + (size_t)i * ldt] = -tau[i] * s;
        }
    }
}

// dtpqrt(m, n, L=0, nb): T stored nb x n.
inline void tpqrt(int m, int n, int nb, double* A, int lda, double* B, int ldb, double* T, int ldt) {
    for (int i = 0; i < n; i += nb) {
        int ib = std::min(n - i, nb);
        tpqrt2(m, ib, &A[i + (size_t)i * lda], lda, &B[(size_t)i * ldb], ldb, &T[(size_t)i * ldt], ldt);
        if (i + ib < n)
            tprfb_lt(m, n - i - ib, ib, &B[(size_t)i * ldb], ldb, &T[(size_t)i * ldt], ldt,
                     &A[i + (size_t)(i + ib) * lda], lda, &B[(size_t)(i + ib) * ldb], ldb);
    }
}

// dtpmqrt('L','T', m, n, k, L=0, nb): [A (k x n); B (m x n)] <- Q^T [A; B].
inline void tpmqrt_lt(int m, int n, int k, int nb, const double* V, int ldv, const double* T, int ldt,
                      double* A, int lda, double* B, int ldb) {
    for (int i = 0; i < k; i += nb) {
        int ib = std::min(nb, k - i);
        tprfb_lt(m, n, ib, &V[(size_t)i * ldv], ldv, &T[(size_t)i * ldt], ldt, &A[i], lda, B, ldb);
    }
}

// Apply the blocked-T Q^T of dgeqrt (T stored nb x k) to C (m x n): C <- Q^T C.
inline void gemqrt_lt(int m, int n, int k, int nb, const double* V, int ldv, const double* T, int ldt,
                      double* C, int ldc) {
    for (int i = 0; i < k; i += nb) {
        int ib = std::min(nb, k - i);
        larfb_left_trans(m - i, n, ib, &V[i + (size_t)i * ldv], ldv, &T[(size_t)i * ldt], ldt, &C[i], ldc);
    }
}

// cblas_dtrsm(Right, Upper, NoTrans, NonUnit): B (m x n) <- B R^{-1}.
inline void trsm_right_upper(int m, int n, const double* R, int ldr, double* B, int ldb) {
    for (int j = 0; j < n; ++j) {
        for (int l = 0; l < j; ++l) {
            double rl = R[l + (size_t)j * ldr];
            if (rl == 0.0) continue;
            for (int i = 0; i < m; ++i) B[i + (size_t)j * ldb] -= B[i + (size_t)l * ldb] * rl;
        }
        double d = 1.0 / R[j + (size_t)j * ldr];
        for (int i = 0; i < m; ++i) B[i + (size_t)j * ldb] *= d;
    }
}

// cblas_dtrsv(Upper, NoTrans, NonUnit): x <- R^{-1} x.
inline void trsv_upper(int n, const double* R, int ldr, double* x) {
    for (int j = n - 1; j >= 0; --j) {
        x[j] /= R[j + (size_t)j * ldr];
        double xj = x[j];
        for (int i = 0; i < j; ++i) x[i] -= R[i + (size_t)j * ldr] * xj;
    }
}

// cblas_dtrsv(Upper, Trans, NonUnit): x <- R^{-T} x (forward substitution with R^T), as
// tQR::fwd / tScaleD::fwd call it (src/toperations.cpp:19,72).
inline void trsv_upper_trans(int n, const double* R, int ldr, double* x) {
    for (int j = 0; j < n; ++j) {
        double s = x[j];
        for (int i = 0; i < j; ++i) s -= R[i + (size_t)j * ldr] * x[i];
        x[j] = s / R[j + (size_t)j * ldr];
    }
}

// dormqr('L','T', n=1): x <- Q^T x = H(k)...H(1) x.
inline void ormqr_lt_vec(int m, int k, const double* V, int ldv, const double* tau, double* x) {
    for (int i = 0; i < k; ++i) {
        double s = x[i];
        for (int r = i + 1; r < m; ++r) s += V[r + (size_t)i * ldv] * x[r];
        s *= tau[i];
        x[i] -= s;
        for (int r = i + 1; r < m; ++r) x[r] -= V[r + (size_t)i * ldv] * s;
    }
}

// dormqr('L','N', n=1): x <- Q x = H(1)...H(k) x.
inline void ormqr_ln_vec(int m, int k, const double* V, int ldv, const double* tau, double* x) {
    for (int i = k - 1; i >= 0; --i) {
        double s = x[i];
        for (int r = i + 1; r < m; ++r) s += V[r + (size_t)i * ldv] * x[r];
        s *= tau[i];
        x[i] -= s;
        for (int r = i + 1; r < m; ++r) x[r] -= V[r + (size_t)i * ldv] * s;
    }
}

inline std::atomic<long>& geqp3_near_ties() {
    static std::atomic<long> count{0};
    return count;
}
inline double& geqp3_near_tie_band() {
    static double band = 1e-12;
    return band;
}

// dgeqp3 (column-pivoted QR; the dlaqp2 step applied to every column).  jpvt is 0-based
// on return (the reference decrements LAPACK's 1-based pivots, src/util.cpp:510-511).
// Pivot = first index of the largest partial column norm; norms are downdated with
// LAPACK's cancellation guard (recompute when temp2 <= sqrt(eps)).
// tie_rel > 0 replaces the argmax by "lowest index whose norm is >= (1 - tie_rel) * max":
// identical to LAPACK whenever the best column wins by more than tie_rel, and independent of
// rounding order when two columns tie in exact arithmetic (symmetric grids produce such ties).
inline void geqp3(int m, int n, double* A, int lda, int* jpvt, double* tau, double tie_rel = 0.0) {
    if (m == 0 || n == 0) return;
    int k = std::min(m, n);
    std::vector<double> vn1(n), vn2(n);
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < m; ++i) s += A[i + (size_t)j * lda] * A[i + (size_t)j * lda];
        vn1[j] = vn2[j] = std::sqrt(s);
        jpvt[j] = j;
    }
    const double tol3z = std::sqrt(1.1102230246251565e-16);  // sqrt(dlamch('E'))
    for (int i = 0; i < k; ++i) {
        int pvt = i;
        for (int j = i + 1; j < n; ++j)
            if (std::fabs(vn1[j]) > std::fabs(vn1[pvt])) pvt = j;
        if (tie_rel > 0.0) {
            const double thr = std::fabs(vn1[pvt]) * (1.0 - tie_rel);
            for (int j = i; j < n; ++j)
                if (std::fabs(vn1[j]) >= thr) { pvt = j; break; }
        }
        {
            // diagnostic: count pivot steps whose winner is within 1e-12 of a runner-up
            double best = std::fabs(vn1[pvt]);
            for (int j = i; j < n; ++j)
                if (j != pvt && best > 0 && std::fabs(vn1[j]) >= best * (1.0 - geqp3_near_tie_band())) { geqp3_near_ties()++; break; }
        }
        if (pvt != i) {
            for (int r = 0; r < m; ++r) std::swap(A[r + (size_t)pvt * lda], A[r + (size_t)i * lda]);
            std::swap(jpvt[pvt], jpvt[i]);
            vn1[pvt] = vn1[i];
            vn2[pvt] = vn2[i];
        }
        tau[i] = larfg(m - i, A[i + (size_t)i * lda], &A[i + 1 + (size_t)i * lda], 1);
        if (tau[i] != 0.0) {
            for (int l = i + 1; l < n; ++l) {
                double w = A[i + (size_t)l * lda];
                for (int r = i + 1; r < m; ++r) w += A[r + (size_t)i * lda] * A[r + (size_t)l * lda];
                w *= tau[i];
                A[i + (size_t)l * lda] -= w;
                for (int r = i + 1; r < m; ++r) A[r + (size_t)l * lda] -= A[r + (size_t)i * lda] * w;
            }
        }
        for (int j = i + 1; j < n; ++j) {
            if (vn1[j] != 0.0) {
                double temp = std::fabs(A[i + (size_t)j * lda]) / vn1[j];
                temp = 1.0 - temp * temp;
                temp = std::max(temp, 0.0);
                double q = vn1[j] / vn2[j];
                double temp2 = temp * q * q;
                if (temp2 <= tol3z) {
                    if (i + 1 < m) {
                        double s = 0.0;
                        for (int r = i + 1; r < m; ++r) s += A[r + (size_t)j * lda] * A[r + (size_t)j * lda];
                        vn1[j] = std::sqrt(s);
                        vn2[j] = vn1[j];
                    } else {
                        vn1[j] = 0.0;
                        vn2[j] = 0.0;
                    }
                } else {
                    vn1[j] *= std::sqrt(temp);
                }
            }
        }
    }
}

// choose_rank (src/util.cpp:830-851), rel = true.
inline int choose_rank(const double* s, int n, double tol) {
    if (tol == 0) return n;
    if (tol >= 1.0) return 0;
    if (n <= 1) return n;
    double sref = std::fabs(s[0]);
    int rank = 1;
    while (rank < n && std::fabs(s[rank]) / sref >= tol) rank++;
    return rank;
}

}  // namespace orc
