/*
 * spaqr_ptree.hpp -- the reference's C++ ParTree / dist_gmres / Operation API over the C ABI
 * of spaqr_gpu.h, header only.
 *
 * A caller written against include/ptree.h:123-159, include/is.h:28 and
 * include/operations.h:13-19 of the reference keeps its code: the methods have the same names
 * and argument meaning.  They are templates, so the reference's Eigen types work unchanged --
 * a sparse matrix needs rows(), cols(), outerIndexPtr(), innerIndexPtr(), valuePtr() in
 * compressed column storage (Eigen::SparseMatrix<double> after makeCompressed()), a vector
 * size(), data(), resize() (Eigen::VectorXd, std::vector<double>), the coordinate matrix
 * rows(), cols(), data() (Eigen::MatrixXd, d x n column-major).
 *
 * One GPU holds the whole tree (the reference's one-rank case): nranks() is 1, the local sizes
 * are the global ones, set_nthreads / set_ttor_log have nothing to configure.
 */
#ifndef SPAQR_PTREE_HPP
#define SPAQR_PTREE_HPP

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "spaqr_gpu.h"

namespace spaqr {

/* include/operations.h:13-19: an operation applied to the vector by the preconditioner */
struct Operation {
    virtual void fwd() = 0;
    virtual void bwd() = 0;
    virtual ~Operation() {}
    virtual std::string Opname() { return "o"; }
};

namespace detail {
inline void check(int rc, const spaqr_tree* t, const char* what) {
    if (rc != 0) {
        const char* m = spaqr_last_error(t);
        throw std::runtime_error(std::string(what) + " failed: " + (m ? m : ""));
    }
}
}  // namespace detail

/* Compressed-column and column-major dense matrices with the accessors the ParTree templates
   read (the subset of Eigen::SparseMatrix / Eigen::MatrixXd they use). */
struct SparseMatrix {
    int nr = 0, nc = 0;
    std::vector<int> colptr, rowind;
    std::vector<double> val;
    int rows() const { return nr; }
    int cols() const { return nc; }
    const int* outerIndexPtr() const { return colptr.data(); }
    const int* innerIndexPtr() const { return rowind.data(); }
    const double* valuePtr() const { return val.data(); }
};
struct DenseMatrix {
    int nr = 0, nc = 0;
    std::vector<double> x;
    int rows() const { return nr; }
    int cols() const { return nc; }
    const double* data() const { return x.data(); }
};

/* include/mmio.hpp:174-377: Matrix Market files through spaqr_mm_* (the driver reads A with
   sp_mmread and the coordinates with dense_mmread, pspaQR.cpp:85,162) */
namespace mmio {
inline SparseMatrix sp_mmread(const std::string& path) {
    SparseMatrix A;
    long long nnz = 0;
    detail::check(spaqr_mm_read_sparse(path.c_str(), &A.nr, &A.nc, &nnz, nullptr, nullptr, nullptr), nullptr,
                  "sp_mmread");
    A.colptr.resize((size_t)A.nc + 1);
    A.rowind.resize((size_t)nnz);
    A.val.resize((size_t)nnz);
    detail::check(spaqr_mm_read_sparse(path.c_str(), &A.nr, &A.nc, &nnz, A.colptr.data(), A.rowind.data(),
                                       A.val.data()),
                  nullptr, "sp_mmread");
    return A;
}
inline DenseMatrix dense_mmread(const std::string& path) {
    DenseMatrix X;
    detail::check(spaqr_mm_read_dense(path.c_str(), &X.nr, &X.nc, nullptr), nullptr, "dense_mmread");
    X.x.resize((size_t)X.nr * X.nc);
    detail::check(spaqr_mm_read_dense(path.c_str(), &X.nr, &X.nc, X.x.data()), nullptr, "dense_mmread");
    return X;
}
inline std::vector<double> vector_mmread(const std::string& path) {
    DenseMatrix X = dense_mmread(path);
    if (X.nc != 1) throw std::runtime_error("vector_mmread: " + path + " is not an n x 1 array");
    return X.x;
}
inline void sp_mmwrite(const std::string& path, const SparseMatrix& A, bool symmetric = false) {
    detail::check(spaqr_mm_write_sparse(path.c_str(), A.nr, A.nc, A.colptr.data(), A.rowind.data(), A.val.data(),
                                        symmetric ? 1 : 0),
                  nullptr, "sp_mmwrite");
}
inline void dense_mmwrite(const std::string& path, const DenseMatrix& X) {
    detail::check(spaqr_mm_write_dense(path.c_str(), X.nr, X.nc, X.x.data()), nullptr, "dense_mmwrite");
}
}  // namespace mmio

namespace detail {

/* a device vector of FP64 owned by the caller (spaqr_device_alloc / free) */
class DeviceVec {
  public:
    DeviceVec() {}
    explicit DeviceVec(size_t n) { resize(n); }
    ~DeviceVec() { release(); }
    DeviceVec(const DeviceVec&) = delete;
    DeviceVec& operator=(const DeviceVec&) = delete;
    void resize(size_t n) {
        if (n <= n_ && p_) return;
        release();
        p_ = (double*)spaqr_device_alloc(sizeof(double) * (n ? n : 1));
        if (!p_) throw std::runtime_error("spaqr_device_alloc failed");
        n_ = n;
    }
    void release() {
        if (p_) spaqr_device_free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    double* get() const { return p_; }
    void put(const double* h, size_t n) { check(spaqr_memcpy_h2d(p_, h, sizeof(double) * n), nullptr, "h2d copy"); }
    void take(double* h, size_t n) const { check(spaqr_memcpy_d2h(h, p_, sizeof(double) * n), nullptr, "d2h copy"); }

  private:
    double* p_ = nullptr;
    size_t n_ = 0;
};

/* one phase of one level of the preconditioner apply, as an Operation */
class PhaseOp : public Operation {
  public:
    PhaseOp(spaqr_tree* t, int level, int phase) : t_(t), level_(level), phase_(phase) {}
    void fwd() override { check(spaqr_tree_solve_phase(t_, level_, phase_, 1), t_, "solve phase"); }
    void bwd() override { check(spaqr_tree_solve_phase(t_, level_, phase_, 0), t_, "solve phase"); }
    std::string Opname() override {
        static const char* names[3] = {"Merge", "QR", "ScaleOrthogonalD"};
        return names[phase_];
    }

  private:
    spaqr_tree* t_;
    int level_, phase_;
};
}  // namespace detail

class ParTree {
  public:
    /* include/ptree.h:135 */
    ParTree(int nlevels, int skip) : nlevels_(nlevels) {
        t_ = spaqr_tree_create(nlevels, skip);
        if (!t_) throw std::runtime_error(std::string("spaqr_tree_create: ") + spaqr_last_error(nullptr));
    }
    ~ParTree() {
        ops_.clear();
        x_.release();
        y_.release();
        if (t_) spaqr_tree_destroy(t_);
    }
    ParTree(const ParTree&) = delete;
    ParTree& operator=(const ParTree&) = delete;

    /* include/tree.h:137-143, include/ptree.h:123-125 */
    void set_tol(double tol) { detail::check(spaqr_tree_set_tol(t_, tol), t_, "set_tol"); }
    void set_skip(int skip) { detail::check(spaqr_tree_set_skip(t_, skip), t_, "set_skip"); }
    void set_scale(int s) { detail::check(spaqr_tree_set_scale(t_, s), t_, "set_scale"); }
    void set_square(int s) { detail::check(spaqr_tree_set_square(t_, s), t_, "set_square"); }
    void set_order(float o) { detail::check(spaqr_tree_set_order(t_, o), t_, "set_order"); }
    void set_verbose(int v) { detail::check(spaqr_tree_set_verbose(t_, v), t_, "set_verbose"); }
    void set_nthreads(int) {}
    void set_ttor_log(int) {}
    template <class Mat> void set_Xcoo(Mat* X) {
        detail::check(spaqr_tree_set_Xcoo(t_, X->data(), (int)X->rows(), (int)X->cols()), t_, "set_Xcoo");
    }
    /* leaf parts in place of coordinates (the METIS path, src/ptree.cpp:503-518): e.g. the
       output of METIS_PartGraphRecursive on A^T A with 2^(nlevels-1) parts; with neither
       coordinates nor parts partition() bisects the graph itself */
    void set_parts(const std::vector<int>& parts) {
        detail::check(spaqr_tree_set_parts(t_, parts.data(), (int)parts.size()), t_, "set_parts");
    }
    /* FP32 fronts (no counterpart in the reference): call before assemble */
    void set_precision(int bits) { detail::check(spaqr_tree_set_precision(t_, bits), t_, "set_precision"); }

    /* src/ptree.cpp:495, 696, 873 */
    template <class SpMat> void partition(SpMat& A) {
        n_ = (int)A.cols();
        m_ = (int)A.rows();
        detail::check(spaqr_tree_partition(t_, (int)A.rows(), (int)A.cols(), A.outerIndexPtr(), A.innerIndexPtr(),
                                           A.valuePtr()),
                      t_, "partition");
    }
    template <class SpMat> void assemble(SpMat&) {
        detail::check(spaqr_tree_assemble(t_), t_, "assemble");
        x_.resize(m_ > n_ ? m_ : n_);
        y_.resize(m_ > n_ ? m_ : n_);
        ops_.clear();
    }
    int factorize() {
        detail::check(spaqr_tree_factorize(t_), t_, "factorize");
        return 0;
    }

    /* solve(VectorXd b, VectorXd& x), src/ptree.cpp:2268: b and x in the natural ordering; for a
       rectangular A, b has m entries and x gets n (the preconditioner's least-squares answer) */
    template <class Vec> void solve(Vec b, Vec& x) const {
        if (m_ != n_) {
            std::vector<double> bt(m_), xt(n_);
            detail::check(spaqr_tree_permute_rows(t_, 0, b.data(), bt.data()), t_, "permute_rows");
            x_.put(bt.data(), m_);
            detail::check(spaqr_tree_solve_ls(t_, x_.get(), y_.get()), t_, "solve");
            y_.take(xt.data(), n_);
            x.resize(n_);
            detail::check(spaqr_tree_permute(t_, 1, xt.data(), x.data()), t_, "permute");
            return;
        }
        std::vector<double> bt(n_), xt(n_);
        detail::check(spaqr_tree_permute(t_, 0, b.data(), bt.data()), t_, "permute");
        x_.put(bt.data(), n_);
        detail::check(spaqr_tree_solve(t_, x_.get()), t_, "solve");
        x_.take(xt.data(), n_);
        x.resize(n_);
        detail::check(spaqr_tree_permute(t_, 1, xt.data(), x.data()), t_, "permute");
    }
    /* solve(VectorXd& x), src/ptree.cpp:2788: x <- M^{-1} x in the tree ordering */
    template <class Vec> void solve(Vec& x) const {
        x_.put(x.data(), n_);
        detail::check(spaqr_tree_solve(t_, x_.get()), t_, "solve");
        x_.take(x.data(), n_);
    }
    /* spmv(VectorXd x), src/ptree.cpp:3212: A x in the tree orderings (m entries) */
    template <class Vec> Vec spmv(Vec x) const {
        x_.put(x.data(), n_);
        detail::check(spaqr_tree_spmv(t_, x_.get(), y_.get()), t_, "spmv");
        Vec y(x);
        y.resize(m_);
        y_.take(y.data(), m_);
        return y;
    }
    /* solve_nrml(VectorXd b, VectorXd& x) (the sequential Tree's, src/tree.cpp:1524-1570):
       x ~ (A^T A)^{-1} b, both n entries in the natural ordering -- cgls's preconditioner */
    template <class Vec> void solve_nrml(Vec b, Vec& x) const {
        std::vector<double> bt(n_), xt(n_);
        detail::check(spaqr_tree_permute(t_, 0, b.data(), bt.data()), t_, "permute");
        x_.put(bt.data(), n_);
        detail::check(spaqr_tree_solve_nrml(t_, x_.get(), y_.get()), t_, "solve_nrml");
        y_.take(xt.data(), n_);
        x.resize(n_);
        detail::check(spaqr_tree_permute(t_, 1, xt.data(), x.data()), t_, "permute");
    }
    /* distribute_x(x, x_dist), src/ptree.cpp:3295-3304 with one rank: x <- P x, x_dist = x */
    template <class Vec> void distribute_x(Vec& x, Vec& x_dist) {
        std::vector<double> t(x.data(), x.data() + n_);
        detail::check(spaqr_tree_permute(t_, 1, t.data(), x.data()), t_, "permute");
        x_dist.resize(n_);
        for (int i = 0; i < n_; ++i) x_dist.data()[i] = x.data()[i];
    }
    /* extract_x(x), src/ptree.cpp:3264-3293 with one rank: the last preconditioner result
       (held in the solve workspace) in the natural ordering */
    template <class Vec> void extract_x(Vec& x) {
        std::vector<double> t(n_);
        detail::check(spaqr_tree_solve_store(t_, y_.get()), t_, "solve_store");
        y_.take(t.data(), n_);
        x.resize(n_);
        detail::check(spaqr_tree_permute(t_, 1, t.data(), x.data()), t_, "permute");
    }
    int nrows_local() const { return m_; }
    int ncols_local() const { return n_; }
    int nranks() const { return 1; }

    /* The preconditioner as a list of operations (the sequential Tree's ops, include/tree.h):
       per level Merge (levels above 0), QR and ScaleOrthogonalD.  apply_operations replays them
       as the reference's Tree::solve does -- fwd() in order, bwd() in reverse -- on x (tree
       ordering), which gives solve(x). */
    std::vector<std::unique_ptr<Operation>>& operations() {
        if (ops_.empty())
            for (int l = 0; l < nlevels_; ++l)
                for (int ph = 0; ph < 3; ++ph)
                    if (l > 0 || ph > 0) ops_.emplace_back(new detail::PhaseOp(t_, l, ph));
        return ops_;
    }
    template <class Vec> void apply_operations(Vec& x) {
        auto& ops = operations();
        x_.put(x.data(), n_);
        detail::check(spaqr_tree_solve_load(t_, x_.get()), t_, "solve_load");
        for (auto& op : ops) op->fwd();
        for (auto it = ops.rbegin(); it != ops.rend(); ++it) (*it)->bwd();
        detail::check(spaqr_tree_solve_store(t_, x_.get()), t_, "solve_store");
        x_.take(x.data(), n_);
    }

    spaqr_tree* handle() const { return t_; }

  private:
    spaqr_tree* t_ = nullptr;
    int nlevels_ = 0, n_ = 0, m_ = 0;
    mutable detail::DeviceVec x_, y_;
    std::vector<std::unique_ptr<Operation>> ops_;
};

/* dist_gmres(A, b, M, iter, tol, verb), src/is.cpp:447-527: left-preconditioned GMRES from
   x0 = 0 in the tree ordering; b is overwritten with the solution and the 0-based index of the
   last step is returned.  As in the reference, the iteration stops at tol only when verb is
   true (with verb false all iter steps run). */
template <class SpMat, class Vec>
unsigned long dist_gmres(const SpMat&, Vec& b, const ParTree& M, unsigned long iter, double tol, bool verb) {
    const int n = M.ncols_local();
    detail::DeviceVec bd(n), xd(n);
    bd.put(b.data(), n);
    spaqr_gmres_result r;
    detail::check(spaqr_gmres(M.handle(), bd.get(), xd.get(), (int)iter, verb ? tol : 0.0, 0, 0.0, &r, nullptr, 0,
                              nullptr),
                  M.handle(), "gmres");
    xd.take(b.data(), n);
    return (unsigned long)r.iters;
}

/* cgls(mat, rhs, x, precond, iters, tol_error, verb), src/is.cpp:274-345, with x0 = 0: CG on
   the normal equations of the least-squares system mat x ~ rhs (natural orderings; rhs m
   entries, x n) preconditioned by precond.solve_nrml, on the device.  iters: in the maximum, out
   the count; tol_error: in the tolerance on ||A^T r|| / ||A^T b||, out the value reached. */
template <class MatrixType, class Rhs, class Dest>
int cgls(const MatrixType&, const Rhs& rhs, Dest& x, const ParTree& precond, long& iters, double& tol_error,
         bool verb = false) {
    (void)verb;
    const int m = precond.nrows_local(), n = precond.ncols_local();
    std::vector<double> bt(m), xt(n);
    detail::check(spaqr_tree_permute_rows(precond.handle(), 0, rhs.data(), bt.data()), precond.handle(), "permute_rows");
    detail::DeviceVec bd(m), xd(n);
    bd.put(bt.data(), m);
    int it = 0;
    double rel = 0.0;
    detail::check(spaqr_cgls(precond.handle(), bd.get(), xd.get(), (int)iters, tol_error, &it, &rel), precond.handle(),
                  "cgls");
    xd.take(xt.data(), n);
    x.resize(n);
    detail::check(spaqr_tree_permute(precond.handle(), 1, xt.data(), x.data()), precond.handle(), "permute");
    iters = it;
    tol_error = rel;
    return it;
}

}  // namespace spaqr

#endif /* SPAQR_PTREE_HPP */
