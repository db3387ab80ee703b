// GMRES over a batch of independent systems (Tree::partition_batch): one Krylov process per
// system, run in lockstep.  The SpMV and the preconditioner apply act on all systems at once
// (the factorization is block diagonal over the systems, so each segment of a vector is
// transformed on its own); the Gram-Schmidt dots, norms and updates are per-system segmented
// launches; the Givens updates run on the host per system from one small read-back per step.
// Each system follows dist_gmres (src/is.cpp:447-527) -- MGS Arnoldi, Givens rotations, x0 = 0,
// its own stopping test -- with the restart and flexible variants of Tree::gmres_dev.  A system
// that has converged keeps zero basis vectors until the others finish.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "launch.hpp"
#include "tree.hpp"
#include "tree_impl.hpp"

namespace spaqr {

int Tree::gmres_batch_dev(const double* b, double* x, int iters, double tol_, int restart, double true_tol,
                          bool flexible, std::vector<GmresResult>& res) {
    try {
        if (!factorized) throw std::runtime_error("factorize() first");
        const int ns = (int)inst_n.size();
        if (ns < 1) throw std::runtime_error("no systems");
        const int m = (restart > 0) ? std::min(restart, iters) : iters;
        if (m < 1) throw std::runtime_error("iters must be positive");
        const int hs = m + 2;  // stride of one system's Hessenberg column in H
        if (!d_seg) {
            std::vector<int> off(inst_off.begin(), inst_off.end());
            off.push_back(n);
            HIPCHK(hipMalloc(&d_seg, sizeof(int) * off.size()));
            HIPCHK(hipMemcpy(d_seg, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
        }
        // workspace: basis Q (m + 1 vectors), Z (m, flexible), w, then the per-system scalars
        const size_t nv = (size_t)(m + 1) + (flexible ? (size_t)m : 0) + 1;
        const size_t nsc = (size_t)seg_partials(ns) + (size_t)ns * hs + (size_t)ns * (m + 1) + 2 * (size_t)ns;
        double* Q = krylov_ws(sizeof(double) * ((size_t)n * nv + nsc));
        double* Z = flexible ? Q + (size_t)n * (m + 1) : nullptr;
        double* w = Q + (size_t)n * (nv - 1);
        double* part = Q + (size_t)n * nv;
        double* H = part + seg_partials(ns);
        double* coef = H + (size_t)ns * hs;
        double* rn = coef + (size_t)ns * (m + 1);
        int* keep = (int*)(rn + ns);
        auto dot = [&](const double* a, const double* bb, double* out, int stride, int sq) {
            launch_seg_dot(ns, d_seg, a, bb, part, out, stride, sq, stream_);
        };
        auto read = [&](std::vector<double>& h, const double* d, size_t cnt) {
            h.resize(cnt);
            HIPCHK(hipMemcpyAsync(h.data(), d, sizeof(double) * cnt, hipMemcpyDeviceToHost, stream_));
            sync_stream();
        };
        auto put_keep = [&](const std::vector<int>& k) {
            HIPCHK(hipMemcpyAsync(keep, k.data(), sizeof(int) * ns, hipMemcpyHostToDevice, stream_));
            sync_stream();  // the host vector may change right after
        };
        struct Sys {
            double bnorm = 0, ref = -1, last_err = 1, true_rel = 1;
            int total = 0, cycles = 0, steps = 0;
            bool done = false, run = false;
            std::vector<double> H, cs, sn, beta;
        };
        std::vector<Sys> sy(ns);
        launch_fill<double>(x, (size_t)n, 0.0, stream_);
        std::vector<double> hv;
        dot(b, b, rn, 1, 1);
        read(hv, rn, ns);
        for (int s = 0; s < ns; ++s) sy[s].bnorm = hv[s];
        bool first = true;
        std::vector<int> kp(ns, 0);
        while (true) {
            // r = b - A x (left: M^{-1}(b - A x))
            if (first) {
                HIPCHK(hipMemcpyAsync(Q, b, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
            } else {
                if (spmv_dev(x, w) != 0) throw std::runtime_error(last_error);
                HIPCHK(hipMemcpyAsync(Q, b, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
                launch_axpby(n, -1.0, w, 1.0, Q, stream_);
            }
            if (!flexible && solve_dev(Q) != 0) throw std::runtime_error(last_error);
            dot(Q, Q, rn, 1, 1);
            read(hv, rn, ns);
            int running = 0;
            for (int s = 0; s < ns; ++s) {
                Sys& S = sy[s];
                S.run = false;
                S.steps = 0;
                if (S.done) continue;
                if (!(hv[s] == hv[s])) throw std::runtime_error("batch GMRES: NaN residual");
                if (hv[s] == 0.0) {
                    S.true_rel = 0.0;
                    S.done = true;
                    continue;
                }
                if (S.ref < 0) S.ref = flexible ? S.bnorm : hv[s];
                S.run = true;
                ++running;
                S.H.assign((size_t)(m + 1) * m, 0.0);
                S.cs.assign(m, 0.0);
                S.sn.assign(m, 0.0);
                S.beta.assign(m + 1, 0.0);
                S.beta[0] = hv[s];
            }
            if (!running) break;
            for (int s = 0; s < ns; ++s) kp[s] = sy[s].run ? 1 : 0;
            put_keep(kp);
            launch_seg_scal(ns, d_seg, rn, 1, keep, Q, Q, stream_);
            for (int k = 0; k < m && running; ++k) {
                double* qk = Q + (size_t)k * n;
                double* qn = Q + (size_t)(k + 1) * n;
                if (flexible) {
                    double* zk = Z + (size_t)k * n;
                    HIPCHK(hipMemcpyAsync(zk, qk, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
                    if (solve_dev(zk) != 0) throw std::runtime_error(last_error);
                    if (spmv_dev(zk, qn) != 0) throw std::runtime_error(last_error);
                } else {
                    if (spmv_dev(qk, qn) != 0) throw std::runtime_error(last_error);
                    if (solve_dev(qn) != 0) throw std::runtime_error(last_error);
                }
                timed(P_KRYLOV, 4L * (k + 1) + 4, ns, [&] {
                    for (int i = 0; i <= k; ++i) {
                        dot(qn, Q + (size_t)i * n, H + i, hs, 0);
                        launch_seg_axpy(ns, d_seg, H + i, hs, Q + (size_t)i * n, qn, stream_);
                    }
                    dot(qn, qn, H + k + 1, hs, 1);
                    launch_seg_scal(ns, d_seg, H + k + 1, hs, keep, qn, qn, stream_);
                });
                prof_add(P_KRYLOV, 4.0 * n * (k + 1) + 3.0 * n, 8.0 * n * (2.0 * (k + 1) + 3.0) + 8.0 * n * (k + 1));
                read(hv, H, (size_t)ns * hs);
                running = 0;
                for (int s = 0; s < ns; ++s) {
                    Sys& S = sy[s];
                    if (!S.run) continue;
                    auto h = [&](int i, int j) -> double& { return S.H[i + (size_t)j * (m + 1)]; };
                    for (int i = 0; i <= k + 1; ++i) h(i, k) = hv[(size_t)s * hs + i];
                    for (int i = 0; i < k; ++i) {
                        const double tmp = S.cs[i] * h(i, k) + S.sn[i] * h(i + 1, k);
                        h(i + 1, k) = -S.sn[i] * h(i, k) + S.cs[i] * h(i + 1, k);
                        h(i, k) = tmp;
                    }
                    if (h(k, k) == 0) {
                        S.cs[k] = 0;
                        S.sn[k] = 1;
                    } else {
                        const double tt = std::sqrt(h(k, k) * h(k, k) + h(k + 1, k) * h(k + 1, k));
                        S.cs[k] = std::fabs(h(k, k)) / tt;
                        S.sn[k] = S.cs[k] * h(k + 1, k) / h(k, k);
                    }
                    h(k, k) = S.cs[k] * h(k, k) + S.sn[k] * h(k + 1, k);
                    h(k + 1, k) = 0.0;
                    S.beta[k + 1] = -S.sn[k] * S.beta[k];
                    S.beta[k] = S.cs[k] * S.beta[k];
                    S.last_err = std::fabs(S.beta[k + 1]) / S.ref;
                    S.steps = k + 1;
                    if (S.last_err <= tol_) S.run = false;  // this cycle of system s ends here
                    else ++running;
                }
                if (running && k + 1 < m) {
                    for (int s = 0; s < ns; ++s) kp[s] = sy[s].run ? 1 : 0;
                    put_keep(kp);
                }
            }
            // x += sum_j y_j basis_j per system, y from the triangular system of its own steps
            std::vector<double> cf((size_t)ns * (m + 1), 0.0);
            int kmax = 0;
            for (int s = 0; s < ns; ++s) {
                Sys& S = sy[s];
                const int kk = S.steps;
                if (kk == 0) continue;
                auto h = [&](int i, int j) -> double { return S.H[i + (size_t)j * (m + 1)]; };
                std::vector<double> y(kk, 0.0);
                for (int i = kk - 1; i >= 0; --i) {
                    y[i] = S.beta[i];
                    for (int j = i + 1; j < kk; ++j) y[i] -= h(i, j) * y[j];
                    y[i] /= h(i, i);
                }
                for (int j = 0; j < kk; ++j) cf[(size_t)s * (m + 1) + j] = y[j];
                kmax = std::max(kmax, kk);
                S.total += kk;
                S.cycles++;
            }
            HIPCHK(hipMemcpyAsync(coef, cf.data(), sizeof(double) * cf.size(), hipMemcpyHostToDevice, stream_));
            const double* basis = flexible ? Z : Q;
            for (int j = 0; j < kmax; ++j) launch_seg_axpy_coef(ns, d_seg, coef + j, m + 1, basis + (size_t)j * n, x, stream_);
            first = false;
            // true residuals
            if (spmv_dev(x, w) != 0) throw std::runtime_error(last_error);
            launch_axpby(n, 1.0, b, -1.0, w, stream_);
            dot(w, w, rn, 1, 1);
            read(hv, rn, ns);  // also waits for the coefficient upload above
            bool all = true;
            for (int s = 0; s < ns; ++s) {
                Sys& S = sy[s];
                if (S.done) continue;
                S.true_rel = hv[s] / S.bnorm;
                if (restart <= 0 || S.true_rel <= true_tol || S.total >= iters) S.done = true;
                all = all && S.done;
            }
            if (all) break;
        }
        sync_stream();
        res.assign(ns, GmresResult());
        for (int s = 0; s < ns; ++s) {
            res[s].iters = sy[s].total - 1;
            res[s].cycles = sy[s].cycles;
            res[s].prec_res = sy[s].last_err;
            res[s].true_res = sy[s].true_rel;
        }
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

}  // namespace spaqr
