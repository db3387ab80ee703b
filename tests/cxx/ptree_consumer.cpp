// A C++ caller of the MI355X path written against the reference's API shape: ParTree,
// dist_gmres, the Operation list and mmio (include/spaqr_ptree.hpp over include/spaqr_gpu.h),
// with no Python in the process.  It builds the 2D 5-point Poisson problem on a 64 x 64 grid with
// the driver's preprocessing (A <- A D 10, pspaQR.cpp:132-147; RHS std::mt19937(0) +
// uniform_real_distribution(-1, 1)), passes A and the coordinates through Matrix Market files
// (sp_mmwrite / sp_mmread, dense_mmwrite / dense_mmread: the driver's input path), then
// partition -> assemble -> factorize -> dist_gmres, and checks: the files give back the same
// matrix and coordinates; true relative residual <= 1e-12; the Operation replay equals solve(x)
// bit for bit; solve(b, x) in the natural ordering equals the permuted in-place solve.  Then an
// overdetermined system through cgls (the least-squares path, ordered without coordinates),
// checked on the normal equations.  Prints one line of results and exits non-zero on any failed
// check.
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "spaqr_ptree.hpp"

struct Csc {  // compressed column storage with the accessors of Eigen::SparseMatrix
    int n = 0;
    std::vector<int> colptr, rowind;
    std::vector<double> val;
    int rows() const { return n; }
    int cols() const { return n; }
    const int* outerIndexPtr() const { return colptr.data(); }
    const int* innerIndexPtr() const { return rowind.data(); }
    const double* valuePtr() const { return val.data(); }
};

struct Coords {  // d x n column-major, as Eigen::MatrixXd
    int d = 0, n = 0;
    std::vector<double> x;
    int rows() const { return d; }
    int cols() const { return n; }
    const double* data() const { return x.data(); }
};

static Csc poisson2d(int m) {
    Csc A;
    A.n = m * m;
    A.colptr.push_back(0);
    for (int j = 0; j < A.n; ++j) {  // column j, rows in increasing order
        const int i0 = j / m, i1 = j % m;
        const int nb[5] = {j - m, j - 1, j, j + 1, j + m};
        const bool ok[5] = {i0 > 0, i1 > 0, true, i1 < m - 1, i0 < m - 1};
        for (int q = 0; q < 5; ++q)
            if (ok[q]) {
                A.rowind.push_back(nb[q]);
                A.val.push_back(q == 2 ? 4.0 : -1.0);
            }
        A.colptr.push_back((int)A.rowind.size());
    }
    for (int j = 0; j < A.n; ++j) {  // A <- A D 10, D_jj = 1 / ||A(:, j)||
        double s = 0;
        for (int k = A.colptr[j]; k < A.colptr[j + 1]; ++k) s += A.val[k] * A.val[k];
        const double d = 1.0 / std::sqrt(s);
        for (int k = A.colptr[j]; k < A.colptr[j + 1]; ++k) A.val[k] = A.val[k] * d * 10.0;
    }
    return A;
}

// An overdetermined system for the least-squares path: [P; G] with P the 5-point Poisson
// operator on an m x m grid and G its forward differences (one row per interior grid edge),
// column-scaled as the driver scales A.
static spaqr::SparseMatrix least_squares2d(int m) {
    const int n = m * m;
    std::vector<std::vector<std::pair<int, double>>> cols(n);
    for (int j = 0; j < n; ++j) {
        const int i0 = j / m, i1 = j % m;
        const int nb[5] = {j - m, j - 1, j, j + 1, j + m};
        const bool ok[5] = {i0 > 0, i1 > 0, true, i1 < m - 1, i0 < m - 1};
        for (int q = 0; q < 5; ++q)
            if (ok[q]) cols[j].push_back({nb[q], q == 2 ? 4.0 : -1.0});
    }
    int r = n;
    for (int ax = 0; ax < 2; ++ax)
        for (int i0 = 0; i0 < m; ++i0)
            for (int i1 = 0; i1 < m; ++i1) {
                const int j = i0 * m + i1;
                if ((ax == 0 && i0 == m - 1) || (ax == 1 && i1 == m - 1)) continue;
                const int h = ax == 0 ? j + m : j + 1;
                cols[h].push_back({r, 1.0});
                cols[j].push_back({r, -1.0});
                ++r;
            }
    spaqr::SparseMatrix A;
    A.nr = r;
    A.nc = n;
    A.colptr.push_back(0);
    for (int j = 0; j < n; ++j) {
        double s = 0;
        for (auto& e : cols[j]) s += e.second * e.second;
        const double d = 10.0 / std::sqrt(s);
        for (auto& e : cols[j]) {  // row indices increase within a column
            A.rowind.push_back(e.first);
            A.val.push_back(e.second * d);
        }
        A.colptr.push_back((int)A.rowind.size());
    }
    return A;
}

int main() {
    const int m = 64, n = m * m;
    Csc A = poisson2d(m);
    Coords X;
    X.d = 2;
    X.n = n;
    for (int i0 = 0; i0 < m; ++i0)
        for (int i1 = 0; i1 < m; ++i1) {
            X.x.push_back(i0);
            X.x.push_back(i1);
        }
    std::mt19937 gen(0);
    std::uniform_real_distribution<double> dis(-1.0, 1.0);
    std::vector<double> b(n);
    for (auto& v : b) v = dis(gen);

    int failed = 0;
    try {
        // through Matrix Market files, as the driver reads its inputs
        char dir[] = "/tmp/spaqr_consumer_XXXXXX";
        if (!mkdtemp(dir)) throw std::runtime_error("mkdtemp failed");
        const std::string fa = std::string(dir) + "/A.mtx", fx = std::string(dir) + "/X.mtx";
        {
            spaqr::SparseMatrix S;
            S.nr = S.nc = A.n;
            S.colptr = A.colptr;
            S.rowind = A.rowind;
            S.val = A.val;
            spaqr::DenseMatrix D;
            D.nr = X.d;
            D.nc = X.n;
            D.x = X.x;
            spaqr::mmio::sp_mmwrite(fa, S);
            spaqr::mmio::dense_mmwrite(fx, D);
            const spaqr::SparseMatrix S2 = spaqr::mmio::sp_mmread(fa);
            const spaqr::DenseMatrix D2 = spaqr::mmio::dense_mmread(fx);
            if (S2.nr != S.nr || S2.colptr != S.colptr || S2.rowind != S.rowind || S2.val != S.val || D2.x != D.x)
                failed |= 32;
            std::remove(fa.c_str());
            std::remove(fx.c_str());
            rmdir(dir);
        }
        spaqr::ParTree t(6, 0);
        t.set_tol(1e-2);
        t.set_scale(1);
        t.set_order(1.0f);
        t.set_square(1);
        t.set_nthreads(16);
        t.set_Xcoo(&X);
        t.partition(A);
        t.assemble(A);
        t.factorize();

        // the right-hand side in the tree ordering
        std::vector<double> bt(n);
        spaqr::detail::check(spaqr_tree_permute(t.handle(), 0, b.data(), bt.data()), t.handle(), "permute");

        // dist_gmres: b is overwritten with the solution
        std::vector<double> x = bt;
        const unsigned long it = spaqr::dist_gmres(A, x, t, 300, 1e-12, true);
        std::vector<double> Ax = t.spmv(x);
        double rn = 0, bn = 0;
        for (int i = 0; i < n; ++i) {
            rn += (bt[i] - Ax[i]) * (bt[i] - Ax[i]);
            bn += bt[i] * bt[i];
        }
        const double relres = std::sqrt(rn / bn);
        if (!(relres <= 1e-12)) failed |= 1;

        // the Operation replay is solve(x), kernel for kernel
        std::vector<double> x1 = bt, x2 = bt;
        t.solve(x1);
        t.apply_operations(x2);
        double dmax = 0;
        for (int i = 0; i < n; ++i) dmax = std::fmax(dmax, std::fabs(x1[i] - x2[i]));
        if (dmax != 0.0) failed |= 2;

        // solve(b, x) in the natural ordering against the in-place solve in the tree ordering
        std::vector<double> xn, x1n(n);
        t.solve(b, xn);
        spaqr::detail::check(spaqr_tree_permute(t.handle(), 1, x1.data(), x1n.data()), t.handle(), "permute");
        double dn = 0;
        for (int i = 0; i < n; ++i) dn = std::fmax(dn, std::fabs(xn[i] - x1n[i]));
        if (dn != 0.0) failed |= 4;

        // extract_x returns the last preconditioner result in the natural ordering
        std::vector<double> xe;
        t.extract_x(xe);
        double de = 0;
        for (int i = 0; i < n; ++i) de = std::fmax(de, std::fabs(xe[i] - xn[i]));
        if (de != 0.0) failed |= 8;

        // least squares through cgls (solve_nrml as its preconditioner), ordered by the graph
        // bisection (no coordinates); checked on the host: ||A^T (b - A x)|| / ||A^T b||
        const spaqr::SparseMatrix L = least_squares2d(24);
        std::vector<double> bl(L.nr);
        for (auto& v : bl) v = dis(gen);
        spaqr::ParTree tl(5, 0);
        tl.set_tol(1e-2);
        tl.set_scale(1);
        tl.set_order(1.0f);
        tl.set_square(0);
        tl.partition(L);
        tl.assemble(L);
        tl.factorize();
        std::vector<double> xl;
        long ls_it = 200;
        double ls_tol = 1e-12;
        spaqr::cgls(L, bl, xl, tl, ls_it, ls_tol);
        std::vector<double> rl = bl, gr(L.nc, 0.0), gb(L.nc, 0.0);
        for (int j = 0; j < L.nc; ++j)
            for (int k = L.colptr[j]; k < L.colptr[j + 1]; ++k) rl[L.rowind[k]] -= L.val[k] * xl[j];
        for (int j = 0; j < L.nc; ++j)
            for (int k = L.colptr[j]; k < L.colptr[j + 1]; ++k) {
                gr[j] += L.val[k] * rl[L.rowind[k]];
                gb[j] += L.val[k] * bl[L.rowind[k]];
            }
        double ng = 0, nb2 = 0;
        for (int j = 0; j < L.nc; ++j) { ng += gr[j] * gr[j]; nb2 += gb[j] * gb[j]; }
        const double ls_rel = std::sqrt(ng / nb2);
        if (!(ls_rel <= 1e-11)) failed |= 64;

        std::printf("consumer: n=%d levels=6 gmres_iters=%lu relres=%.3e ops=%zu replay_diff=%g natural_diff=%g "
                    "extract_diff=%g ls_m=%d ls_n=%d cgls_iters=%ld cgls_rel=%.3e ls_normal_rel=%.3e failed=%d\n",
                    n, it, relres, t.operations().size(), dmax, dn, de, L.nr, L.nc, ls_it, ls_tol, ls_rel, failed);
    } catch (std::exception& e) {
        std::printf("consumer: error: %s\n", e.what());
        failed |= 16;
    }
    spaqr_shutdown();
    return failed ? 1 : 0;
}
