// Oracle of the randomized sparsification rank path (rank modes 1 and 2 of the HIP planner).
// TEST INFRASTRUCTURE ONLY: FP64 restatement of what the GPU kernels compute, used by tests
// to compare the kernels on fixed inputs; nothing in the product links it.
//
// The path has no counterpart on the reference's hot path.  Its closest reference analogue
// is the unused adaptive range finder random_rrqr (src/util.cpp:643-715); the quantity it
// decides is the one geqp3 + choose_rank decide in sparsify_rrqr_only
// (src/ptree.cpp:373-439, src/util.cpp:830-851): the number of columns kept so that every
// discarded tail is below tol times the largest column of C.
//
// Steps (FP64, column-major, ld = rows):
//   gauss         Omega(i, j) from splitmix64 (S. Vigna, public domain) of seed ^ (j<<32 | i)
//                 and a second round of it, then Box-Muller: sqrt(-2 ln u1) cos(2 pi u2)
//   colnorm       cn[j] = ||C(:, j)||^2, cmax2 = max_j cn[j]
//   sketch        Y = C Omega(0:w, 0:s)
//   features      the predictor's 32 inputs (R diagonal decay of a QR'd 16-column sketch
//                 and structural ratios)
//   tail_select   smallest r <= rY with ||B(r:k, j)||^2 <= tol^2 cmax2 for all j, B = Q^T C;
//                 -1 when no r <= rY passes
//   sketch_rank   the pipeline: sketch, QR of Y (pivoted geqp3 or unpivoted geqrt, full
//                 width), B = Q^T C with rY = min(k, s) reflectors, tail_select
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

#include "dense.hpp"

namespace orc {

inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double gauss_entry(int i, int j, uint64_t seed) {
    const uint64_t h1 = splitmix64(seed ^ (((uint64_t)j << 32) | (uint32_t)i));
    const uint64_t h2 = splitmix64(h1);
    const double u1 = 1.0 - (double)(h1 >> 11) * 0x1.0p-53;  // (0, 1]
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;        // [0, 1)
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

inline void colnorm2(int k, int w, const double* C, double* cn, double* cmax2) {
    double mx = 0.0;
    for (int j = 0; j < w; ++j) {
        double s = 0.0;
        for (int i = 0; i < k; ++i) s += C[i + (size_t)j * k] * C[i + (size_t)j * k];
        cn[j] = s;
        mx = std::max(mx, s);
    }
    *cmax2 = mx;
}

inline void sketch(int k, int w, int s, const double* C, const double* om, int ldo, double* Y) {
    for (int c = 0; c < s; ++c)
        for (int i = 0; i < k; ++i) {
            double a = 0.0;
            for (int j = 0; j < w; ++j) a += C[i + (size_t)j * k] * om[j + (size_t)c * ldo];
            Y[i + (size_t)c * k] = a;
        }
}

// rdiag: |R_jj| of the QR of the 16-column sketch (min(16, k, w) entries).
inline void rank_features(int k, int w, const double* rdiag, const double* cn, double cmax2, int nsout, int nsin,
                          int level, double tol, float* x) {
    const double thr2 = tol * tol * cmax2;
    double tot = 0.0, count = 0.0;
    for (int j = 0; j < w; ++j) {
        tot += cn[j];
        count += (cn[j] >= thr2 && cn[j] > 0.0) ? 1.0 : 0.0;
    }
    const int kw = std::min(k, w), nd = std::min(std::min(16, k), w);
    const double r0 = nd > 0 ? std::fabs(rdiag[0]) : 0.0;
    for (int f = 0; f < 32; ++f) {
        double v = 0.0;
        if (f < 16) {
            if (f < nd && r0 > 0.0) v = std::min(1.0, std::max(0.0, 1.0 + std::log10(std::fabs(rdiag[f]) / r0 + 1e-30) / 8.0));
        } else {
            switch (f) {
                case 16: v = std::log2(k + 1.0) / 11.0; break;
                case 17: v = std::log2(w + 1.0) / 14.0; break;
                case 18: v = (double)kw / (double)std::max(std::max(k, w), 1); break;
                case 19: v = level / 16.0; break;
                case 20: v = tol > 0.0 ? -std::log10(tol) / 8.0 : 1.0; break;
                case 21: v = nsout / 32.0; break;
                case 22: v = nsin / 32.0; break;
                case 23: v = w > 0 ? count / w : 0.0; break;
                case 24: v = kw > 0 ? std::min(1.0, count / kw) : 0.0; break;
                case 25: v = tot > 0.0 ? cmax2 / tot : 0.0; break;
                case 26: v = (cmax2 > 0.0 && kw > 0) ? std::min(1.0, tot / cmax2 / kw) : 0.0; break;
                default: v = 0.0;
            }
        }
        x[f] = (float)v;
    }
}

inline int tail_select(int k, int w, const double* B, int ry, double cmax2, double tol) {
    const double thr2 = tol * tol * cmax2;
    int need = 0;
    for (int j = 0; j < w; ++j) {
        const double* col = B + (size_t)j * k;
        // tails ||B(r:k, j)||^2 for r = 0 .. k, from the bottom
        std::vector<double> tail(k + 1, 0.0);
        for (int i = k - 1; i >= 0; --i) tail[i] = tail[i + 1] + col[i] * col[i];
        int rc = -1;
        for (int r = 0; r <= std::min(ry, k); ++r)
            if (tail[r] <= thr2) { rc = r; break; }
        if (rc < 0) return -1;
        need = std::max(need, rc);
    }
    return need;
}

// The whole sketch path on one interface; returns the tail-select rank (-1: no rank <= rY).
inline int sketch_rank(int k, int w, int s, const double* C, const double* om, int ldo, double tol, bool pivoted) {
    std::vector<double> cn(w), Y((size_t)k * s), B(C, C + (size_t)k * w);
    double cmax2 = 0.0;
    colnorm2(k, w, C, cn.data(), &cmax2);
    sketch(k, w, s, C, om, ldo, Y.data());
    const int ry = std::min(k, s);
    std::vector<double> tau(ry, 0.0);
    if (pivoted) {
        std::vector<int> jp(s);
        geqp3(k, s, Y.data(), k, jp.data(), tau.data());
    } else {
        for (int j = 0; j < ry; ++j) {
            double* col = Y.data() + j + (size_t)j * k;
            double alpha = col[0];
            tau[j] = larfg(k - j, alpha, col + 1, 1);
            col[0] = 1.0;
            for (int c = j + 1; c < s; ++c) {
                double* y = Y.data() + j + (size_t)c * k;
                double d = 0.0;
                for (int i = 0; i < k - j; ++i) d += col[i] * y[i];
                d *= tau[j];
                for (int i = 0; i < k - j; ++i) y[i] -= d * col[i];
            }
            col[0] = alpha;
        }
    }
    for (int j = 0; j < w; ++j) ormqr_lt_vec(k, ry, Y.data(), k, tau.data(), B.data() + (size_t)j * k);
    return tail_select(k, w, B.data(), ry, cmax2, tol);
}

}  // namespace orc
