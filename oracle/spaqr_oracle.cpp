// spaQR oracle: CPU FP64 restatement of the pspaQR ParTree path on one rank.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared against, and the
// "port" CPU baseline that bench.py times beside it; the product never links it.
//
// What it restates (reference file:line):
//   ordering      GeometricPartitionAtA / bissect_geo / getInterfacesUsingA
//                 (src/partition.cpp:241-350,514-587), initPermutation (src/tree.cpp:48-57),
//                 merge_if (src/cluster.cpp:23-32), cluster hierarchy (src/ptree.cpp:587-691)
//   assembly      ParTree::assemble (src/ptree.cpp:696-864), block2dense (src/util.cpp:853-899)
//   factorize     ParTree::factorize (src/ptree.cpp:873-2154) with the task bodies
//                 geqrt_cluster :170-195, tsqrt_edge :197-214, larfb_edge :216-221,
//                 ssrfb_edges :223-266, scale_cluster :269-286, trsm_edge :365-370,
//                 sparsify_rrqr_only :373-439, scale apply :1621-1633,
//                 sparsify apply :1867-1897, compute_new_edges :289-340,
//                 fill-in get_sparsity/alloc_fillin :866-871,140-153
//   solve         ParTree::solve(x) (src/ptree.cpp:2788-3210) with helpers :2160-2262
//   spmv          ParTree::spmv (src/ptree.cpp:3212-3262)
//   GMRES         dist_gmres (src/is.cpp:447-527) and LUSolve (src/util.cpp:1019-1032)
//
// Orders the reference leaves to its task scheduler or to pointer hashing are fixed here:
// edgesIn lists, fill-in pair order and the solve's reductions follow cluster order; the
// interiors of one level update shared blocks (ssrfb, and the solve's order_of_trans) colour
// by colour, with the DSatur colouring of dsatur_colours below and cluster order inside a
// colour (a colour's interiors share no out-neighbour, so that inner order cannot matter).
// The RRQR is full geqp3 + choose_rank (the reference's non-MKL branch,
// src/ptree.cpp:411-416).  Only square systems (rperm == cperm, c2r_count == 1).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "dense.hpp"
#include "rank_path.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// Planted mistakes, for tests that show the parity checks catch them (0: none).  Each is a
// plausible slip in restating the reference:
//   1 interface inference takes a neighbour's side from its ancestor one level too high
//     (src/partition.cpp:538 divides the separator index by 2^(lvl - nbr.lvl - 1))
//   2 the RRQR gather drops the A(c,m)_top columns of C (src/ptree.cpp:390-398)
//   3 the truncation keeps one column fewer than choose_rank (src/ptree.cpp:420-431)
//   4 the sparsify apply skips the left transform Q_c^T (src/ptree.cpp:1881-1889)
//   5 the sparsified level range also takes the second level from the top
//     (src/ptree.cpp:30, 1566, 1635 stop at nlevels - 2)
//   6 getInterfacesUsingATA counts a neighbour in the vertex's own half as right of it
//     (src/partition.cpp:659 tests pj <= pi)
//   7 sparsify_extra keeps one row fewer than choose_rank (src/ptree.cpp:480-492)
//   8 the solve skips orthogonalD_extra_fwd, Q_ex^T on the extra rows (src/ptree.cpp:2531)
static int g_fault = 0;

// f(i) for i in [0, n) on the OpenMP threads.  Tasks are independent (each writes only its own
// blocks), so the results do not depend on the thread count; the first exception is rethrown.
template <class F>
static void par_for(int n, F&& f) {
    std::exception_ptr err;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) {
        try {
            f(i);
        } catch (...) {
#pragma omp critical
            if (!err) err = std::current_exception();
        }
    }
    if (err) std::rethrow_exception(err);
}

// Relative band inside which geqp3 pivot candidates count as tied (lowest index wins); the
// HIP kernel uses the same rule, so FP64 runs pick the same pivots despite rounding order.
constexpr double kPivotTieRel = 1e-10;

// ------------------------------------------------------------------ IDs
struct SepID {
    int lvl = -1, sep = 0;
    SepID() {}
    SepID(int l, int s) : lvl(l), sep(s) {}
    bool operator==(const SepID& o) const { return lvl == o.lvl && sep == o.sep; }
    bool operator!=(const SepID& o) const { return !(*this == o); }
    bool operator<(const SepID& o) const { return lvl < o.lvl || (lvl == o.lvl && sep < o.sep); }
};

struct ClusterID {  // include/cluster.h:61-135
    SepID self, l, r;
    int part = 2, section = 0;
    bool operator==(const ClusterID& o) const {
        return self == o.self && l == o.l && r == o.r && part == o.part && section == o.section;
    }
    bool operator!=(const ClusterID& o) const { return !(*this == o); }
    bool operator<(const ClusterID& o) const {
        if (self != o.self) return self < o.self;
        if (l != o.l) return l < o.l;
        if (r != o.r) return r < o.r;
        if (section != o.section) return section < o.section;
        return part < o.part;
    }
};

static SepID merge_sep(const SepID& s) { return SepID(s.lvl + 1, s.sep / 2); }

static ClusterID merge_if(const ClusterID& c, int lvl) {  // src/cluster.cpp:23-32
    ClusterID p;
    p.self = c.self;
    p.l = c.l.lvl < lvl ? merge_sep(c.l) : c.l;
    p.r = c.r.lvl < lvl ? merge_sep(c.r) : c.r;
    p.section = lvl > 0 ? c.section / 2 : c.section;
    p.part = c.self.lvl == lvl ? 2 : c.part;
    return p;
}

static SepID find_highest_common(SepID n1, SepID n2) {  // src/partition.cpp:63-86
    int l1 = n1.lvl, l2 = n2.lvl, s1 = n1.sep, s2 = n2.sep;
    while (l1 < l2) { l1++; s1 /= 2; }
    while (l2 < l1) { l2++; s2 /= 2; }
    while (s1 != s2) { l1++; l2++; s1 /= 2; s2 /= 2; }
    return SepID(l1, s1);
}

static SepID find_lowest_common(SepID n1, SepID n2) {  // src/partition.cpp:88-100
    return n1.lvl < n2.lvl ? n1 : n2;
}

// ------------------------------------------------------------------ sparse (CSC)
struct CSC {
    int nr = 0, nc = 0;
    std::vector<int> colptr, rowind;
    std::vector<double> val;
};

static CSC transpose(const CSC& A) {
    CSC T;
    T.nr = A.nc; T.nc = A.nr;
    T.colptr.assign(A.nr + 1, 0);
    for (int k = 0; k < A.colptr[A.nc]; ++k) T.colptr[A.rowind[k] + 1]++;
    for (int i = 0; i < A.nr; ++i) T.colptr[i + 1] += T.colptr[i];
    T.rowind.resize(A.colptr[A.nc]);
    T.val.resize(A.colptr[A.nc]);
    std::vector<int> pos(T.colptr.begin(), T.colptr.end() - 1);
    for (int j = 0; j < A.nc; ++j)
        for (int k = A.colptr[j]; k < A.colptr[j + 1]; ++k) {
            int p = pos[A.rowind[k]]++;
            T.rowind[p] = j;
            T.val[p] = A.val[k];
        }
    return T;  // row indices come out sorted because j increases
}

// Structure of A^T A (columns i, j adjacent iff they share a row), diagonal included.
static CSC ata_structure(const CSC& A) {
    CSC At = transpose(A);
    CSC G;
    G.nr = G.nc = A.nc;
    G.colptr.assign(A.nc + 1, 0);
    std::vector<int> mark(A.nc, -1);
    std::vector<int> tmp;
    for (int j = 0; j < A.nc; ++j) {
        tmp.clear();
        for (int k = A.colptr[j]; k < A.colptr[j + 1]; ++k) {
            int r = A.rowind[k];
            for (int q = At.colptr[r]; q < At.colptr[r + 1]; ++q) {
                int i = At.rowind[q];
                if (mark[i] != j) { mark[i] = j; tmp.push_back(i); }
            }
        }
        std::sort(tmp.begin(), tmp.end());
        G.rowind.insert(G.rowind.end(), tmp.begin(), tmp.end());
        G.colptr[j + 1] = (int)G.rowind.size();
    }
    G.val.assign(G.rowind.size(), 1.0);
    return G;
}

// A(rperm, cperm): entry (rperm[i], cperm[j]) of A lands at (i, j).
static CSC permute(const CSC& A, const std::vector<int>& rperm, const std::vector<int>& cperm) {
    std::vector<int> rinv(A.nr);
    for (int i = 0; i < A.nr; ++i) rinv[rperm[i]] = i;
    CSC P;
    P.nr = A.nr; P.nc = A.nc;
    P.colptr.assign(A.nc + 1, 0);
    std::vector<std::pair<int, double>> col;
    for (int j = 0; j < A.nc; ++j) {
        int oj = cperm[j];
        col.clear();
        for (int k = A.colptr[oj]; k < A.colptr[oj + 1]; ++k) col.push_back({rinv[A.rowind[k]], A.val[k]});
        std::sort(col.begin(), col.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (auto& e : col) { P.rowind.push_back(e.first); P.val.push_back(e.second); }
        P.colptr[j + 1] = (int)P.rowind.size();
    }
    return P;
}

// ------------------------------------------------------------------ ordering
// Tensor coordinates of linspace_nd (src/util.cpp:901-930): id = sum x_i n^(d-1-i).
static std::vector<double> linspace_nd(int n, int d) {
    int N = 1;
    for (int i = 0; i < d; ++i) N *= n;
    std::vector<double> X((size_t)d * N);
    for (int id = 0; id < N; ++id) {
        int rem = id;
        for (int dim = d - 1; dim >= 0; --dim) {
            X[dim + (size_t)id * d] = rem % n;
            rem /= n;
        }
    }
    return X;
}

// bissect_geo (src/partition.cpp:241-279): split on the widest dimension at the value of
// the N/2-th smallest coordinate; <= goes left.  The value (not the tie order) decides.
static void bissect_geo(const std::vector<int>& dofs, std::vector<int>& parts, const std::vector<double>& X, int d,
                        int start) {
    int N = (int)dofs.size();
    if (N == 0) return;
    int bestdim = -1;
    double maxrange = -1.0;
    for (int dim = 0; dim < d; ++dim) {
        double mx = -INFINITY, mn = INFINITY;
        for (int i : dofs) { mx = std::max(mx, X[dim + (size_t)i * d]); mn = std::min(mn, X[dim + (size_t)i * d]); }
        if (mx - mn > maxrange) { bestdim = dim; maxrange = mx - mn; }
    }
    std::vector<double> v(N);
    for (int i = 0; i < N; ++i) v[i] = X[bestdim + (size_t)dofs[i] * d];
    std::nth_element(v.begin(), v.begin() + N / 2, v.end());
    double midv = v[N / 2];
    for (int j : dofs) parts[j] = (X[bestdim + (size_t)j * d] <= midv) ? start : start + 1;
}

static int part_at_lvl(int part_leaf, int lvl) {  // src/partition.cpp:54-61
    for (int l = 0; l < lvl - 1; ++l) part_leaf /= 2;
    return part_leaf;
}

// GeometricPartitionAtA + getInterfacesUsingA for a square matrix without matching.
static std::vector<ClusterID> geometric_partition(const CSC& A, int nlevels, const std::vector<double>& X, int d) {
    const int c = A.nc;
    CSC G = ata_structure(A);
    std::vector<int> parts(c, 0);
    for (int depth = 0; depth < nlevels - 1; ++depth) {
        int np = 1 << depth;
        std::vector<std::vector<int>> dofs(np);
        for (int i = 0; i < c; ++i) dofs[parts[i]].push_back(i);
        for (int k = 0; k < np; ++k) bissect_geo(dofs[k], parts, X, d, 2 * k);
    }
    std::vector<ClusterID> cmap(c);
    for (int depth = 0; depth < nlevels - 1; ++depth) {
        int l = nlevels - depth - 1;
        for (int i = 0; i < c; ++i) {
            if (cmap[i].self.lvl != -1) continue;
            int pi = part_at_lvl(parts[i], l);
            if (pi % 2 != 0) continue;
            bool sep = false;
            for (int k = G.colptr[i]; k < G.colptr[i + 1]; ++k) {
                int j = G.rowind[k];
                if (cmap[j].self.lvl == -1 && part_at_lvl(parts[j], l) == pi + 1) { sep = true; break; }
            }
            if (sep) { cmap[i].self = SepID(l, pi / 2); cmap[i].section = parts[i]; }
        }
    }
    for (int i = 0; i < c; ++i)
        if (cmap[i].self.lvl == -1) {
            cmap[i].self = cmap[i].l = cmap[i].r = SepID(0, parts[i]);
            cmap[i].section = parts[i];
        }
    // getInterfacesUsingA (src/partition.cpp:514-587); square, rmap == cmap.
    std::vector<ClusterID> rmap = cmap;
    CSC At = transpose(A);  // column i of At = row i of A
    for (int i = 0; i < A.nr; ++i) {
        SepID self = rmap[i].self;
        if (self.lvl <= 0) continue;
        SepID left(-1, 0), right(-1, 0);
        for (int k = At.colptr[i]; k < At.colptr[i + 1]; ++k) {
            int j = At.rowind[k];
            SepID nbr = cmap[j].self;
            if (nbr.lvl < self.lvl) {
                int pj = nbr.sep >> (self.lvl - nbr.lvl - (g_fault == 1 ? 0 : 1));
                if (pj % 2 == 0) left = (left.lvl == -1) ? nbr : find_lowest_common(nbr, left);
                else right = (right.lvl == -1) ? nbr : find_lowest_common(nbr, right);
            }
        }
        rmap[i].l = left;
        rmap[i].r = right;
    }
    cmap = rmap;
    for (int i = 0; i < A.nr; ++i) {
        SepID self = rmap[i].self, left = rmap[i].l, right = rmap[i].r;
        if (right == SepID()) {
            for (int k = At.colptr[i]; k < At.colptr[i + 1]; ++k) {
                int j = At.rowind[k];
                if (self == cmap[j].self && cmap[j].r != SepID())
                    right = (right == SepID()) ? cmap[j].r : find_highest_common(right, cmap[j].r);
            }
        } else if (left == SepID()) {
            for (int k = At.colptr[i]; k < At.colptr[i + 1]; ++k) {
                int j = At.rowind[k];
                if (self == cmap[j].self && cmap[j].l != SepID())
                    left = (left == SepID()) ? cmap[j].l : find_highest_common(left, cmap[j].l);
            }
        }
        rmap[i].l = left;
        rmap[i].r = right;
    }
    for (int i = 0; i < A.nr; ++i) {
        if (rmap[i].l == SepID()) rmap[i].l = rmap[i].self;
        if (rmap[i].r == SepID()) rmap[i].r = rmap[i].self;
    }
    return rmap;
}

// MetisPartition (src/partition.cpp:372-424) with the leaf parts given in place of
// partition_metis_RB's: separators on A^T A by the same rule as the geometric path, then
// getInterfacesUsingATA (src/partition.cpp:637-682) -- left starts at the vertex's own leaf,
// right at the first lower-level neighbour in the right half, both widened by
// find_highest_common.
static std::vector<ClusterID> parts_partition(const CSC& A, int nlevels, const std::vector<int>& parts) {
    const int c = A.nc;
    CSC G = ata_structure(A);
    std::vector<ClusterID> cmap(c);
    for (int depth = 0; depth < nlevels - 1; ++depth) {
        const int l = nlevels - depth - 1;
        for (int i = 0; i < c; ++i) {
            if (cmap[i].self.lvl != -1) continue;
            const int pi = part_at_lvl(parts[i], l);
            if (pi % 2 != 0) continue;
            bool sep = false;
            for (int k = G.colptr[i]; k < G.colptr[i + 1]; ++k) {
                const int j = G.rowind[k];
                if (cmap[j].self.lvl == -1 && part_at_lvl(parts[j], l) == pi + 1) { sep = true; break; }
            }
            if (sep) { cmap[i].self = SepID(l, pi / 2); cmap[i].section = parts[i]; }
        }
    }
    for (int i = 0; i < c; ++i)
        if (cmap[i].self.lvl == -1) {
            cmap[i].self = cmap[i].l = cmap[i].r = SepID(0, parts[i]);
            cmap[i].section = parts[i];
        }
    for (int i = 0; i < c; ++i) {
        const SepID self = cmap[i].self;
        if (self.lvl <= 0) continue;
        SepID left(0, parts[i]), right(-1, 0);
        const int pi = part_at_lvl(parts[i], self.lvl);
        for (int k = G.colptr[i]; k < G.colptr[i + 1]; ++k) {
            const int j = G.rowind[k];
            const SepID nbr = cmap[j].self;
            if (nbr.lvl < self.lvl) {
                const int pj = part_at_lvl(parts[j], self.lvl);
                if (g_fault == 6 ? pj < pi : pj <= pi) left = find_highest_common(nbr, left);
                else right = (right.lvl == -1) ? nbr : find_highest_common(nbr, right);
            }
        }
        if (right.lvl == -1) throw std::runtime_error("separator without a right neighbour");
        cmap[i].l = left;
        cmap[i].r = right;
        cmap[i].section = parts[i];
    }
    return cmap;
}

// initPermutation (src/tree.cpp:48-57).
static std::vector<int> init_permutation(int nlevels, const std::vector<ClusterID>& cmap) {
    std::vector<int> cperm(cmap.size());
    for (size_t i = 0; i < cperm.size(); ++i) cperm[i] = (int)i;
    std::vector<ClusterID> pm = cmap;
    auto cmp = [&pm](int i, int j) { return pm[i] < pm[j]; };
    std::stable_sort(cperm.begin(), cperm.end(), cmp);
    for (int lvl = 1; lvl < nlevels; ++lvl) {
        for (auto& s : pm) s = merge_if(s, lvl);
        std::stable_sort(cperm.begin(), cperm.end(), cmp);
    }
    return cperm;
}

// ------------------------------------------------------------------ clusters and edges
// Colours for the interiors of one level (index = position in cluster order; nb[i] = the
// orders of interior i's out-neighbours other than itself).  Two interiors conflict when they
// share an out-neighbour.  DSatur: repeatedly colour the uncoloured interior with the most
// distinct colours among its conflicts (then the most conflicts, then the lowest index) with
// the smallest colour none of its conflicts holds.
static std::vector<int> dsatur_colours(const std::vector<std::vector<int>>& nb) {
    const int V = (int)nb.size();
    std::map<int, std::vector<int>> sharing;
    for (int i = 0; i < V; ++i)
        for (int n : nb[i]) sharing[n].push_back(i);
    std::vector<std::vector<int>> adj(V);
    for (auto& kv : sharing)
        for (int a : kv.second)
            for (int b : kv.second)
                if (a != b) adj[a].push_back(b);
    for (auto& a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    std::vector<int> col(V, -1), sat(V, 0);
    std::vector<std::vector<char>> seen(V);
    std::set<std::tuple<int, int, int>> queue;  // (-saturation, -conflicts, index)
    for (int i = 0; i < V; ++i) queue.insert(std::make_tuple(0, -(int)adj[i].size(), i));
    while (!queue.empty()) {
        const int i = std::get<2>(*queue.begin());
        queue.erase(queue.begin());
        int c = 0;
        while (c < (int)seen[i].size() && seen[i][c]) ++c;
        col[i] = c;
        for (int j : adj[i]) {
            if (col[j] >= 0) continue;
            if ((int)seen[j].size() <= c) seen[j].resize(c + 1, 0);
            if (seen[j][c]) continue;
            queue.erase(std::make_tuple(-sat[j], -(int)adj[j].size(), j));
            seen[j][c] = 1;
            ++sat[j];
            queue.insert(std::make_tuple(-sat[j], -(int)adj[j].size(), j));
        }
    }
    return col;
}

struct Cluster;
struct Edge {  // include/edge.h:14-58: block A(n2, n1), rows(n2) x cols(n1)
    Cluster* n1;
    Cluster* n2;
    Mat A;
    bool spars = false;
};

struct Cluster {  // include/cluster.h:142-389
    int cstart, csize, rstart, rsize;
    ClusterID id;
    int order, merge_level;
    bool eliminated = false;
    int rank;
    int row_rank = 0;  // extra rows kept (rows - cols; after sparsify_extra, its rank)
    int csize_org, rsize_org;
    // rectangular systems: an interior's extra-row targets (cluster order per extra row, in
    // row order) and row_perm; an interface's received extra rows (interior order, start in
    // that interior's row_perm, count), sorted; the extra sparsification's reflectors
    std::vector<int> r2c;
    std::vector<int> row_perm;
    std::vector<std::array<int, 3>> extra_from;
    int tot_extra = 0;
    Mat Qex, Tex;
    std::vector<double> tauex;
    int nex = 0;  // rows of Qex: the extra rows at sparsify_extra
    Cluster* parent = nullptr;
    std::vector<Cluster*> children;
    int rposparent = -1, cposparent = -1;
    std::set<int> dist2;  // cluster orders (deterministic stand-in for the pointer set)
    std::vector<Edge*> edgesOut, edgesIn, sparsOut, sparsIn;
    Edge* eself = nullptr;
    int col_interior_deps = 0;
    // factors
    Mat T;                         // T of the diagonal geqrt (k x k, nb = k)
    std::map<int, Mat> Tmap;       // tpqrt T per neighbour order (nb x k)
    Mat Qs, Ts;                    // scaling
    Mat Qsp, Tsp;                  // sparsification
    std::vector<double> tausp;
    bool sparsified = false;
    std::vector<double> x;

    Cluster(int cs, int cz, int rs, int rz, ClusterID id_, int order_, int ml)
        : cstart(cs), csize(cz), rstart(rs), rsize(rz), id(id_), order(order_), merge_level(ml) {
        rank = cz;
        row_rank = rz - cz;
        csize_org = cz;
        rsize_org = rz;
        x.assign(rz, 0.0);
    }
    int lvl() const { return id.self.lvl; }
    int cols() const { return csize; }
    int rows() const { return rsize; }
    void reset_size(int r, int c) { rsize = r; csize = c; rank = c; row_rank = r - c; }  // src/cluster.cpp:214-218
    void set_size(int r, int c) { reset_size(r, c); x.assign(r, 0.0); }
    Edge* find_out(const Cluster* n) const {
        for (Edge* e : edgesOut) if (e->n2 == n) return e;
        return nullptr;
    }
};

struct OpTrace {  // per-level record of the dense work (shapes), for flop/byte accounting
    int level, op;  // op: 0 geqrt, 1 larfb, 2 tpqrt, 3 tpmqrt, 4 scale_geqrt, 5 scale_apply, 6 geqp3, 7 spars_apply
    int m, n, k;
};

struct Tree {
    int nlevels, skip;
    double tol = 0.1;
    int n = 0;
    int m = 0;                      // rows of A (m > n: a least-squares system)
    CSC A, App;
    std::vector<int> rperm;         // row permutation (== cperm for square systems)
    std::vector<double> X;
    int d = 0;
    std::vector<int> given_parts;  // leaf parts in place of coordinates (the METIS path)
    std::vector<int> cperm;
    std::vector<ClusterID> cpermed;
    std::vector<std::vector<Cluster*>> bottoms, interiors, interfaces;
    std::vector<std::vector<Cluster*>> update_order;  // per level: interiors by (colour, cluster order)
    std::vector<std::vector<int>> update_colour;      // their colours
    std::vector<std::unique_ptr<Cluster>> owned_c;
    std::vector<std::unique_ptr<Edge>> owned_e;
    int max_order = 0;
    int current_bottom = 0;
    std::vector<OpTrace> trace;
    std::vector<std::vector<int>> ranks_before, ranks_after;  // per level, sparsified interfaces
    double t_factor = 0.0;

    Tree(int nl, int sk) : nlevels(nl), skip(sk) {
        bottoms.resize(nl);
        interiors.resize(nl);
        interfaces.resize(nl);
        ranks_before.resize(nl);
        ranks_after.resize(nl);
    }

    Cluster* new_cluster(int cs, int cz, int rs, int rz, ClusterID id, int ml) {
        owned_c.emplace_back(new Cluster(cs, cz, rs, rz, id, max_order++, ml));
        return owned_c.back().get();
    }

    bool want_sparsify(const Cluster* c) const {  // src/ptree.cpp:29-41
        if (c->merge_level < skip || c->merge_level >= nlevels - (g_fault == 5 ? 1 : 2) || tol == 0) return false;
        if (c->lvl() > c->merge_level) {
            if (c->id.l.lvl <= c->merge_level && c->id.r.lvl <= c->merge_level) return true;
        }
        return false;
    }

    Cluster* by_order(int o) const {
        // orders at one merge level are contiguous; search all levels
        for (auto& lv : bottoms) {
            if (lv.empty()) continue;
            int b = lv[0]->order;
            if (o >= b && o < b + (int)lv.size()) return lv[o - b];
        }
        return nullptr;
    }

    // ---------------------------------------------------------- partition
    bool square() const { return m == n; }

    // form_rmap without a matching (src/tree.cpp:102-165): every row of Ap = A(:, cperm) goes
    // to the cluster whose columns carry the largest sum of squares of that row (the first
    // such cluster in column order wins ties; a zero row goes to the last column), filed
    // under that cluster's first column the row touches; rows are then ordered by that column,
    // and in row order within it.  c2r_count[j]: the rows filed under permuted column j.
    void form_rmap(const CSC& Ap, std::vector<int>& c2r_count) {
        CSC At = transpose(Ap);  // column i of At: the entries of row i, by increasing column
        std::vector<std::vector<int>> c2r(n);
        c2r_count.assign(n, 0);
        for (int i = 0; i < m; ++i) {
            double max_val = -1.0;
            int cidx = -1;
            for (int k = At.colptr[i]; k < At.colptr[i + 1];) {
                const ClusterID& cid = cpermed[At.rowind[k]];
                double val = At.val[k] * At.val[k];
                int kn = k + 1;
                while (kn < At.colptr[i + 1] && cpermed[At.rowind[kn]] == cid) {
                    val += At.val[kn] * At.val[kn];
                    ++kn;
                }
                if (val > max_val) { max_val = val; cidx = At.rowind[k]; }
                k = kn;
            }
            if (cidx == -1) cidx = n - 1;
            c2r_count[cidx]++;
            c2r[cidx].push_back(i);
        }
        rperm.clear();
        for (int j = 0; j < n; ++j) rperm.insert(rperm.end(), c2r[j].begin(), c2r[j].end());
    }

    void partition() {  // src/ptree.cpp:495-694
        if (!square() && given_parts.empty())
            throw std::runtime_error("a rectangular system needs leaf parts (the METIS path): the geometric path "
                                     "assigns rows to clusters by a bipartite matching (HSL MC64) for non-square A");
        std::vector<ClusterID> cmap =
            given_parts.empty() ? geometric_partition(A, nlevels, X, d) : parts_partition(A, nlevels, given_parts);
        cperm = init_permutation(nlevels, cmap);
        cpermed.resize(n);
        for (int k = 0; k < n; ++k) cpermed[k] = cmap[cperm[k]];
        std::vector<int> c2r_count(n, 1);
        if (square()) {
            rperm = cperm;
        } else {
            std::vector<int> ident(m);
            for (int i = 0; i < m; ++i) ident[i] = i;
            form_rmap(permute(A, ident, cperm), c2r_count);
        }
        // leaf clusters: runs of equal ClusterID, with the rows filed under their columns
        for (int k = 0, l = 0; k < n;) {
            int kn = k + 1, nr = c2r_count[k];
            while (kn < n && cpermed[kn] == cpermed[k]) nr += c2r_count[kn++];
            if (nr == 0) throw std::runtime_error("a leaf cluster got no rows");
            Cluster* s = new_cluster(k, kn - k, l, nr, cpermed[k], 0);
            bottoms[0].push_back(s);
            (s->lvl() == 0 ? interiors[0] : interfaces[0]).push_back(s);
            l += nr;
            k = kn;
        }
        for (int lvl = 1; lvl < nlevels; ++lvl) {
            auto& prev = bottoms[lvl - 1];
            size_t b = 0;
            while (b < prev.size() && prev[b]->lvl() < lvl) b++;
            for (size_t k = b; k < prev.size();) {
                ClusterID pid = merge_if(prev[k]->id, lvl);
                std::vector<Cluster*> ch;
                int cs = 0, rs = 0;
                int cst = prev[k]->cstart, rst = prev[k]->rstart;
                while (k < prev.size() && merge_if(prev[k]->id, lvl) == pid) {
                    ch.push_back(prev[k]);
                    cs += prev[k]->cols();
                    rs += prev[k]->rows();
                    k++;
                }
                Cluster* p = new_cluster(cst, cs, rst, rs, pid, lvl);
                std::sort(ch.begin(), ch.end(), [](Cluster* a, Cluster* b2) { return a->id < b2->id; });
                for (Cluster* c : ch) { c->parent = p; p->children.push_back(c); }
                bottoms[lvl].push_back(p);
                (pid.self.lvl == lvl ? interiors[lvl] : interfaces[lvl]).push_back(p);
            }
        }
    }

    // ---------------------------------------------------------- edges
    Edge* new_edge(Cluster* c1, Cluster* c2, int cur) {  // new_edgeOut, src/ptree.cpp:101-120
        owned_e.emplace_back(new Edge{c1, c2, Mat(c2->rows(), c1->cols())});
        Edge* e = owned_e.back().get();
        c1->edgesOut.push_back(e);
        bool sp = c1->lvl() > cur && c2->lvl() > cur && c1 != c2;
        e->spars = sp;
        if (sp) c1->sparsOut.push_back(e);
        if (c1 != c2) add_edge_in(c2, e, sp);
        return e;
    }
    void add_edge_in(Cluster* c, Edge* e, bool sp) {  // src/cluster.cpp:85-90
        c->edgesIn.push_back(e);
        if (sp) c->sparsIn.push_back(e);
        if (e->n1->lvl() == e->n1->merge_level) c->col_interior_deps++;
    }
    static void sort_out(Cluster* c) {
        std::sort(c->edgesOut.begin(), c->edgesOut.end(),
                  [](Edge* a, Edge* b) { return a->n2->order < b->n2->order; });
    }

    // ---------------------------------------------------------- assemble
    void assemble() {  // src/ptree.cpp:696-864
        App = permute(A, rperm, cperm);
        std::vector<Cluster*> cl(n), rl(m);  // cluster of each permuted column / row
        for (Cluster* s : bottoms[0]) {
            for (int k = s->cstart; k < s->cstart + s->cols(); ++k) cl[k] = s;
            for (int k = s->rstart; k < s->rstart + s->rows(); ++k) rl[k] = s;
        }
        for (Cluster* self : bottoms[0]) {
            std::set<int> nb;
            for (int j = self->cstart; j < self->cstart + self->cols(); ++j)
                for (int k = App.colptr[j]; k < App.colptr[j + 1]; ++k) nb.insert(rl[App.rowind[k]]->order);
            nb.insert(self->order);
            for (int o : nb) {
                Cluster* nbr = bottoms[0][o - bottoms[0][0]->order];
                owned_e.emplace_back(new Edge{self, nbr, Mat(nbr->rows(), self->cols())});
                Edge* e = owned_e.back().get();
                for (int j = 0; j < self->cols(); ++j) {
                    int col = self->cstart + j;
                    for (int k = App.colptr[col]; k < App.colptr[col + 1]; ++k) {
                        int r = App.rowind[k];
                        if (r >= nbr->rstart && r < nbr->rstart + nbr->rows()) e->A(r - nbr->rstart, j) = App.val[k];
                    }
                }
                self->edgesOut.push_back(e);
                if (self != nbr) {
                    bool sp = self->lvl() > 0 && nbr->lvl() > 0;
                    e->spars = sp;
                    add_edge_in(nbr, e, sp);
                    if (sp) self->sparsOut.push_back(e);
                } else {
                    self->eself = e;
                }
            }
            sort_out(self);
        }
        // distance-two connections (src/ptree.cpp:832-854)
        CSC G = ata_structure(App);
        std::vector<char> visited(bottoms[0].size(), 0);
        int o0 = bottoms[0][0]->order;
        for (Cluster* self : bottoms[0]) {
            std::fill(visited.begin(), visited.end(), 0);
            for (int col = self->cstart; col < self->cstart + self->cols(); ++col)
                for (int k = G.colptr[col]; k < G.colptr[col + 1]; ++k) {
                    Cluster* pn = cl[G.rowind[k]];
                    if (visited[pn->order - o0]) continue;
                    if (pn->lvl() >= self->lvl() && pn != self) self->dist2.insert(pn->order);
                    else if (pn->lvl() < self->lvl()) self->dist2.insert(pn->dist2.begin(), pn->dist2.end());
                    visited[pn->order - o0] = 1;
                }
        }
    }

    // ---------------------------------------------------------- factorize
    void factorize() {
        auto t0 = std::chrono::steady_clock::now();
        for (int ilvl = 0; ilvl < nlevels; ++ilvl) {
            current_bottom = ilvl;
            fillin();
            eliminate(ilvl);
            if (!square()) reassign(ilvl);
            scale(ilvl);
            if (ilvl >= skip && ilvl < nlevels - (g_fault == 5 ? 1 : 2) && tol != 0) sparsify(ilvl);
            if (ilvl < nlevels - 1) merge();
        }
        t_factor = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    void fillin() {  // src/ptree.cpp:885-1010
        for (Cluster* c : interiors[current_bottom]) {
            for (Edge* e : c->edgesIn) c->dist2.insert(e->n1->order);  // get_sparsity :866-871
        }
        for (Cluster* c : interiors[current_bottom]) {
            std::vector<int> d2(c->dist2.begin(), c->dist2.end());
            for (int no : d2) {
                Cluster* nn = by_order(no);
                // alloc_fillin(c, nn), src/ptree.cpp:140-153
                std::vector<Edge*> outs = c->edgesOut;
                for (Edge* e : outs) {
                    Cluster* nbr = e->n2;
                    if (!nn->find_out(nbr)) new_edge(nn, nbr, current_bottom);
                }
            }
        }
    }

    void eliminate(int ilvl) {
        auto& ints = interiors[current_bottom];
        const int ni = (int)ints.size();
        std::vector<std::vector<OpTrace>> tl(ni);
        // per interior (each touches only its own row and column blocks): geqrt_cluster
        // (src/ptree.cpp:170-195), larfb_edge on every edgesIn (:216-221), then the tsqrt chain
        // over edgesOut after self (:197-214, 1134-1141, 1453-1461)
        par_for(ni, [&](int i) {
            Cluster* c = ints[i];
            if (c->cols() == 0) throw std::runtime_error("cluster has zero cols");
            if (c->rows() < c->cols()) throw std::runtime_error("an interior has fewer rows than columns");
            int k = c->cols(), m = c->rows();
            c->T = Mat(k, k);
            geqrt(m, k, k, c->eself->A.data(), m, c->T.data(), k);
            c->eliminated = true;
            tl[i].push_back({ilvl, 0, m, k, k});
        });
        par_for(ni, [&](int i) {
            Cluster* c = ints[i];
            int k = c->cols();
            for (Edge* e : c->edgesIn) {
                larfb_left_trans(c->rows(), e->A.c, k, c->eself->A.data(), c->rows(), c->T.data(), k, e->A.data(),
                                 e->A.r);
                tl[i].push_back({ilvl, 1, c->rows(), e->A.c, k});
            }
        });
        par_for(ni, [&](int i) {
            Cluster* c = ints[i];
            int k = c->cols();
            int nb = std::min(32, k);
            if (c->edgesOut.empty() || c->edgesOut[0]->n2 != c) throw std::runtime_error("first edgeOut must be self");
            for (size_t j = 1; j < c->edgesOut.size(); ++j) {
                Edge* e = c->edgesOut[j];
                Mat T(nb, k);
                Mat R = c->eself->A.block(0, 0, k, k);
                tpqrt(e->A.r, k, nb, R.data(), k, e->A.data(), e->A.r, T.data(), nb);
                c->eself->A.set_block(0, 0, R);
                c->Tmap[e->n2->order] = std::move(T);
                tl[i].push_back({ilvl, 2, e->A.r, k, k});
            }
        });
        // trace records in the sequential order: geqrt of every interior, then larfb, then tpqrt
        for (int kind = 0; kind < 3; ++kind)
            for (int i = 0; i < ni; ++i)
                for (const auto& r : tl[i])
                    if (r.op == kind) trace.push_back(r);
        // ssrfb_edges: for every m in edgesIn(c), the chain over n in edgesOut(c)
        // (src/ptree.cpp:223-266, 1366-1406); interiors colour by colour.
        {
            std::vector<std::vector<int>> nb(ints.size());
            for (size_t i = 0; i < ints.size(); ++i)
                for (size_t j = 1; j < ints[i]->edgesOut.size(); ++j) nb[i].push_back(ints[i]->edgesOut[j]->n2->order);
            const std::vector<int> col = dsatur_colours(nb);
            std::vector<int> idx(ints.size());
            for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return col[a] < col[b]; });
            if ((int)update_order.size() <= current_bottom) update_order.resize(current_bottom + 1);
            update_order[current_bottom].clear();
            if ((int)update_colour.size() <= current_bottom) update_colour.resize(current_bottom + 1);
            update_colour[current_bottom].clear();
            for (int i : idx) {
                update_order[current_bottom].push_back(ints[i]);
                update_colour[current_bottom].push_back(col[i]);
            }
            if (getenv("SPAQR_ORACLE_COLOURS")) {
                int nc = 0;
                for (int c : col) nc = std::max(nc, c + 1);
                fprintf(stderr, "[oracle] level %d: %zu interiors, %d colours\n", current_bottom, ints.size(), nc);
            }
        }
        // colour by colour; the interiors of one colour share no out-neighbour, so they update
        // disjoint blocks and run in parallel
        const auto& uo = update_order[current_bottom];
        std::vector<std::vector<OpTrace>> ut(uo.size());
        for (size_t g0 = 0; g0 < uo.size();) {
            size_t g1 = g0 + 1;
            while (g1 < uo.size() && update_colour[current_bottom][g1] == update_colour[current_bottom][g0]) ++g1;
            par_for((int)(g1 - g0), [&](int ii) {
            Cluster* c = uo[g0 + ii];
            auto& trc = ut[g0 + ii];
            int k = c->cols();
            int nb = std::min(32, k);
            for (Edge* cm : c->edgesIn) {
                Cluster* m = cm->n1;
                Mat top = cm->A.block(0, 0, k, cm->A.c);
                for (size_t j = 1; j < c->edgesOut.size(); ++j) {
                    Edge* cn = c->edgesOut[j];
                    Edge* mn = m->find_out(cn->n2);
                    if (!mn) throw std::runtime_error("missing fill-in block");
                    const Mat& T = c->Tmap.at(cn->n2->order);
                    tpmqrt_lt(mn->A.r, mn->A.c, k, nb, cn->A.data(), cn->A.r, T.data(), nb, top.data(), k,
                              mn->A.data(), mn->A.r);
                    trc.push_back({ilvl, 3, mn->A.r, mn->A.c, k});
                }
                cm->A.set_block(0, 0, top);
            }
            });
            g0 = g1;
        }
        for (const auto& v : ut) trace.insert(trace.end(), v.begin(), v.end());
    }

    // Row reassignment of a rectangular system (src/ptree.cpp:1149-1356; the row rule is
    // Cluster::update_extra_rnorm, src/cluster.cpp:151-160).  After an interior c is
    // eliminated, rows cols(c) .. rows(c) of Q_c^T A(c, :) touch no eliminated column: each
    // goes to the out-neighbour m of c (one that is also an in-neighbour, is_cnbr) whose block
    // A(c, m) holds the largest norm of that row (strictly larger replaces, so the first in
    // cluster order keeps a tie; all-zero rows stay with c and are dropped -- the reference's
    // scheduler-ordered loop drops them too when c comes after its neighbours).  row_perm
    // stable-sorts the extra rows by target.  Every interface then appends the rows it
    // received, in interior order, to its diagonal block and to the blocks of its in-edges
    // from non-eliminated clusters (shift_tf + Edge::merge, :1316-1356, include/edge.h:49-57);
    // the rows are zero in a block whose source has no edge into c.
    void reassign(int ilvl) {
        (void)ilvl;
        for (Cluster* c : interiors[current_bottom]) {
            const int k = c->cols(), ex = c->rows() - k;
            c->r2c.assign(std::max(ex, 0), c->order);
            c->row_perm.clear();
            if (ex <= 0) continue;
            std::vector<double> best(ex, 0.0);
            std::vector<Edge*> ins(c->edgesIn.begin(), c->edgesIn.end());
            std::sort(ins.begin(), ins.end(), [](Edge* a, Edge* b) { return a->n1->order < b->n1->order; });
            for (Edge* e : ins) {
                if (!c->find_out(e->n1)) continue;  // is_cnbr
                for (int i = 0; i < ex; ++i) {
                    double s2 = 0.0;
                    for (int j = 0; j < e->A.c; ++j) s2 += e->A(k + i, j) * e->A(k + i, j);
                    const double nrm = std::sqrt(s2);
                    if (nrm > best[i]) { best[i] = nrm; c->r2c[i] = e->n1->order; }
                }
            }
            c->row_perm.resize(ex);
            for (int i = 0; i < ex; ++i) c->row_perm[i] = i;
            std::stable_sort(c->row_perm.begin(), c->row_perm.end(),
                             [&](int a, int b) { return c->r2c[a] < c->r2c[b]; });
            for (Edge* eo : c->edgesOut) {
                Cluster* nn = eo->n2;
                if (nn == c) continue;
                int i0 = 0;
                while (i0 < ex && c->r2c[c->row_perm[i0]] != nn->order) ++i0;
                int i1 = i0;
                while (i1 < ex && c->r2c[c->row_perm[i1]] == nn->order) ++i1;
                nn->extra_from.push_back({c->order, i0 < ex ? i0 : 0, i1 - i0});
                nn->tot_extra += i1 - i0;
            }
        }
        for (Cluster* nn : interfaces[current_bottom]) {
            if (nn->tot_extra == 0) continue;
            std::sort(nn->extra_from.begin(), nn->extra_from.end());
            std::vector<Edge*> grow{nn->eself};
            for (Edge* e : nn->edgesIn)
                if (e->n1->lvl() != e->n1->merge_level) grow.push_back(e);
            for (Edge* e : grow) {
                Cluster* n1 = e->n1;
                Mat G(e->A.r + nn->tot_extra, e->A.c);
                G.set_block(0, 0, e->A);
                int row = e->A.r;
                for (const auto& t : nn->extra_from) {
                    Cluster* c = by_order(t[0]);
                    const int rinc = t[2];
                    Edge* src = n1->find_out(c);
                    if (src)
                        for (int i = 0; i < rinc; ++i) {
                            const int sr = c->cols() + c->row_perm[t[1] + i];
                            for (int j = 0; j < G.c; ++j) G(row + i, j) = src->A(sr, j);
                        }
                    row += rinc;
                }
                e->A = std::move(G);
            }
            nn->reset_size(nn->rows() + nn->tot_extra, nn->cols());
        }
    }

    void scale(int ilvl) {
        auto& itf = interfaces[current_bottom];
        const int nf = (int)itf.size();
        std::vector<std::vector<OpTrace>> tl(nf), ta(nf);
        // scale_cluster (src/ptree.cpp:269-286), one interface per task
        par_for(nf, [&](int i) {
            Cluster* c = itf[i];
            if (!want_sparsify(c)) return;
            Edge* e = c->eself;
            int k = c->cols(), m = c->rows();
            if (m < k) throw std::runtime_error("a scaled interface has fewer rows than columns");
            c->Qs = e->A;
            c->Ts = Mat(k, k);
            geqrt(m, k, k, c->Qs.data(), m, c->Ts.data(), k);
            for (int j = 0; j < k; ++j)
                for (int i2 = 0; i2 < m; ++i2) e->A(i2, j) = (i2 == j) ? 1.0 : 0.0;
            tl[i].push_back({ilvl, 4, m, k, k});
        });
        // scale apply on every interface-interface edge (src/ptree.cpp:1621-1633); each edge belongs
        // to the sparsOut list of one interface, the task that updates it
        par_for(nf, [&](int i) {
            Cluster* n1 = itf[i];
            for (Edge* e : n1->sparsOut) {
                Cluster* c = e->n2;
                if (want_sparsify(c))
                    larfb_left_trans(e->A.r, e->A.c, c->cols(), c->Qs.data(), c->Qs.r, c->Ts.data(), c->Ts.r,
                                     e->A.data(), e->A.r);
                if (want_sparsify(n1)) trsm_right_upper(e->A.r, e->A.c, n1->Qs.data(), n1->Qs.r, e->A.data(), e->A.r);
                if (want_sparsify(c) || want_sparsify(n1)) ta[i].push_back({ilvl, 5, e->A.r, e->A.c, 0});
            }
        });
        for (const auto& v : tl) trace.insert(trace.end(), v.begin(), v.end());
        for (const auto& v : ta) trace.insert(trace.end(), v.begin(), v.end());
    }

    void sparsify(int ilvl) {
        auto& itf = interfaces[current_bottom];
        const int nf = (int)itf.size();
        std::vector<std::vector<OpTrace>> tr(nf), ta(nf);
        // sparsify_rrqr_only (src/ptree.cpp:373-439), one interface per task: each reads its
        // coupling blocks and rewrites only its own factors and diagonal block
        par_for(nf, [&](int ii) {
            Cluster* c = itf[ii];
            if (!want_sparsify(c)) return;
            int k = c->cols();
            int smp = 0, spm = 0;
            for (Edge* e : c->sparsOut) smp += e->A.r;
            if (g_fault != 2)
                for (Edge* e : c->sparsIn) spm += e->A.c;
            Mat C(k, smp + spm);
            int is = 0;
            for (Edge* e : c->sparsOut) {
                for (int j = 0; j < e->A.r; ++j)
                    for (int i = 0; i < k; ++i) C(i, is + j) = e->A(j, i);
                is += e->A.r;
            }
            if (g_fault != 2)
                for (Edge* e : c->sparsIn) {
                    for (int j = 0; j < e->A.c; ++j)
                        for (int i = 0; i < k; ++i) C(i, is + j) = e->A(i, j);
                    is += e->A.c;
                }
            int w = C.c;
            int mn = std::min(k, w);
            std::vector<int> jpvt(w);
            std::vector<double> tau(mn, 0.0);
            int rank = 0;
            if (w > 0 && k > 0) {
                geqp3(k, w, C.data(), k, jpvt.data(), tau.data(), kPivotTieRel);
                std::vector<double> rii(mn);
                for (int i = 0; i < mn; ++i) rii[i] = C(i, i);
                rank = choose_rank(rii.data(), mn, tol);
                if (g_fault == 3 && rank > 1) rank--;
            } else {
                // geqp3 returns without touching C (src/util.cpp:504-505); diagonal is empty
                rank = choose_rank(nullptr, 0, tol);
            }
            tr[ii].push_back({ilvl, 6, k, w, rank});
            c->Qsp = C.block(0, 0, k, rank);
            c->tausp.assign(tau.begin(), tau.begin() + rank);
            c->Tsp = Mat(rank, rank);
            larft(k, rank, c->Qsp.data(), k, c->tausp.data(), c->Tsp.data(), rank);
            c->sparsified = true;
            c->rank = rank;
            int row_rank = 0;
            if (!square()) {
                // sparsify_extra (src/ptree.cpp:441-493): the same RRQR on the extra rows of
                // the cluster's row blocks, C_ex = [A(c, m)(k:rows, :) ...] over sparsIn
                const int ex = c->rows() - k;
                int spm = 0;
                for (Edge* e : c->sparsIn) spm += e->A.c;
                Mat Ce(ex, spm);
                int js = 0;
                for (Edge* e : c->sparsIn) {
                    for (int j = 0; j < e->A.c; ++j)
                        for (int i = 0; i < ex; ++i) Ce(i, js + j) = e->A(k + i, j);
                    js += e->A.c;
                }
                const int mne = std::min(ex, spm);
                std::vector<int> jp(std::max(spm, 1));
                std::vector<double> te(std::max(mne, 1), 0.0);
                if (ex > 0 && spm > 0) {
                    geqp3(ex, spm, Ce.data(), std::max(ex, 1), jp.data(), te.data(), kPivotTieRel);
                    std::vector<double> rii(mne);
                    for (int i = 0; i < mne; ++i) rii[i] = Ce(i, i);
                    row_rank = choose_rank(rii.data(), mne, tol);
                } else {
                    row_rank = choose_rank(nullptr, 0, tol);
                }
                if (g_fault == 7 && row_rank > 1) row_rank--;
                tr[ii].push_back({ilvl, 8, ex, spm, row_rank});
                c->nex = ex;
                c->Qex = Ce.block(0, 0, ex, row_rank);
                c->tauex.assign(te.begin(), te.begin() + row_rank);
                c->Tex = Mat(row_rank, row_rank);
                larft(ex, row_rank, c->Qex.data(), std::max(ex, 1), c->tauex.data(), c->Tex.data(), std::max(row_rank, 1));
            }
            // the pivot block becomes [I_rank; 0] ((rank + row_rank) x rank): src/ptree.cpp:435-438
            // (square) and sparsify_resize_tf :1835-1850 (rectangular)
            c->eself->A = Mat(rank + row_rank, rank);
            for (int i = 0; i < rank; ++i) c->eself->A(i, i) = 1.0;
            c->reset_size(rank + row_rank, rank);
        });
        for (int ii = 0; ii < nf; ++ii)
            for (const auto& r : tr[ii]) {
                trace.push_back(r);
                if (r.op != 6) continue;
                ranks_before[ilvl].push_back(r.m);
                ranks_after[ilvl].push_back(r.k);
            }
        // sparsify apply (src/ptree.cpp:1867-1897) on every interface-interface edge; each edge
        // belongs to the sparsOut list of one interface, the task that rewrites it
        par_for(nf, [&](int ii) {
            Cluster* n = itf[ii];
            for (Edge* e : n->sparsOut) {
                Cluster* c = e->n2;
                Mat Acopy = e->A;
                if (want_sparsify(n)) {
                    larfb_right_notrans(Acopy.r, Acopy.c, n->Qsp.c, n->Qsp.data(), n->Qsp.r, n->Tsp.data(), n->Tsp.r,
                                        Acopy.data(), Acopy.r);
                    e->A = Acopy;
                }
                e->A.conservative_resize(c->rank + c->row_rank, n->rank);
                if (want_sparsify(c) && g_fault != 4) {
                    Mat A11 = Acopy.block(0, 0, c->Qsp.r, n->rank);
                    larfb_left_trans(A11.r, A11.c, c->Qsp.c, c->Qsp.data(), c->Qsp.r, c->Tsp.data(), c->Tsp.r,
                                     A11.data(), A11.r);
                    for (int j = 0; j < n->rank; ++j)
                        for (int i = 0; i < c->rank; ++i) e->A(i, j) = A11(i, j);
                    if (!square()) {  // the extra rows: Q_ex^T A(c,n)(k:rows, :), src/ptree.cpp:1891-1894
                        Mat A21 = Acopy.block(c->Qsp.r, 0, c->nex, n->rank);
                        if (c->row_rank > 0)
                            larfb_left_trans(A21.r, A21.c, c->row_rank, c->Qex.data(), c->Qex.r, c->Tex.data(),
                                             c->Tex.r, A21.data(), A21.r);
                        for (int j = 0; j < n->rank; ++j)
                            for (int i = 0; i < c->row_rank; ++i) e->A(c->rank + i, j) = A21(i, j);
                    }
                }
                if (want_sparsify(n) || want_sparsify(c)) ta[ii].push_back({ilvl, 7, Acopy.r, Acopy.c, 0});
            }
        });
        for (const auto& v : ta) trace.insert(trace.end(), v.begin(), v.end());
    }

    void merge() {  // src/ptree.cpp:1940-2123 (one rank)
        for (Cluster* c : interiors[current_bottom]) c->eliminated = true;
        current_bottom++;
        for (Cluster* snew : bottoms[current_bottom]) {
            int rs = 0, cs = 0;
            for (Cluster* so : snew->children) {
                so->rposparent = rs;
                so->cposparent = cs;
                rs += so->rows();
                cs += so->cols();
            }
            snew->set_size(rs, cs);
            snew->rsize_org = rs;
            snew->csize_org = cs;
        }
        // compute_new_edges (src/ptree.cpp:289-340): snew in cluster order, neighbours sorted
        for (Cluster* snew : bottoms[current_bottom]) {
            std::set<int> nbrs;
            for (Cluster* so : snew->children) {
                for (Edge* eo : so->edgesOut)
                    if (!eo->n2->eliminated) nbrs.insert(eo->n2->parent->order);
                for (int d2o : so->dist2) {
                    Cluster* d2c = by_order(d2o);
                    if (d2c->lvl() >= current_bottom) snew->dist2.insert(d2c->parent->order);
                }
            }
            for (int o : nbrs) {
                Cluster* nn = by_order(o);
                Edge* e = new_edge(snew, nn, current_bottom);
                if (snew == nn) snew->eself = e;
            }
            sort_out(snew);
            for (Cluster* so : snew->children)
                for (Edge* eo : so->edgesOut) {
                    Cluster* no = eo->n2;
                    if (no->eliminated) continue;
                    Edge* f = snew->find_out(no->parent);
                    if (eo->A.r != no->rows() || eo->A.c != so->cols()) throw std::runtime_error("merge size mismatch");
                    f->A.set_block(no->rposparent, so->cposparent, eo->A);
                    eo->A = Mat();  // the block now lives in the parent
                }
        }
    }

    // ---------------------------------------------------------- solve (preconditioner)
    // splitD_fwd / splitD_bwd (src/ptree.cpp:2223-2242): the kept rows of a sparsified
    // rectangular cluster, x = [x(0:rank); x_extra; x_fine] and back.
    static void split_fwd(Cluster* c) {
        const int x1 = c->csize_org - c->rank, x2 = (int)c->x.size() - c->csize_org;
        std::vector<double> f(c->x.begin() + c->rank, c->x.begin() + c->csize_org);
        std::vector<double> e(c->x.begin() + c->csize_org, c->x.end());
        std::copy(e.begin(), e.end(), c->x.begin() + c->rank);
        std::copy(f.begin(), f.end(), c->x.begin() + c->rank + x2);
        (void)x1;
    }
    static void split_bwd(Cluster* c) {
        const int x1 = c->csize_org - c->rank, x2 = (int)c->x.size() - c->csize_org;
        std::vector<double> e(c->x.begin() + c->rank, c->x.begin() + c->rank + x2);
        std::vector<double> f(c->x.begin() + c->rank + x2, c->x.begin() + c->rank + x2 + x1);
        std::copy(f.begin(), f.end(), c->x.begin() + c->rank);
        std::copy(e.begin(), e.end(), c->x.begin() + c->csize_org);
    }
    // every cluster's vector at its full length: rows at creation plus the extra rows it receives
    void size_vectors() {
        for (auto& up : owned_c) up->x.assign(up->rsize_org + up->tot_extra, 0.0);
    }

    // Forward half of solve(b, x) (src/ptree.cpp:2267-2598): b (m entries, permuted rows).
    void solve_fwd(const double* b) {
        size_vectors();
        for (Cluster* c : bottoms[0]) std::copy(b + c->rstart, b + c->rstart + c->rsize_org, c->x.begin());
        for (int L = 0; L < nlevels; ++L) {
            if (L > 0)
                for (Cluster* p : bottoms[L]) {  // merge_fwd :2244-2252
                    int k = 0;
                    for (Cluster* ch : p->children)
                        for (int i = 0; i < ch->rows(); ++i) p->x[k++] = ch->x[i];
                }
            for (Cluster* c : interiors[L]) {  // fwd_geqrt :2846-2851
                int m2 = c->eself->A.r, k = c->eself->A.c;
                gemqrt_lt(m2, 1, k, k, c->eself->A.data(), m2, c->T.data(), k, c->x.data(), m2);
            }
            for (Cluster* c : update_order[L]) {  // fwd_tsqrt :2160-2186, order_of_trans = colour order
                int k = c->cols(), nb = std::min(32, k);
                for (size_t j = 1; j < c->edgesOut.size(); ++j) {
                    Edge* e = c->edgesOut[j];
                    Cluster* nn = e->n2;
                    const Mat& T = c->Tmap.at(nn->order);
                    tpmqrt_lt(e->A.r, 1, k, nb, e->A.data(), e->A.r, T.data(), nb, c->x.data(), k, nn->x.data(),
                              e->A.r);
                }
            }
            if (!square())
                for (Cluster* nn : interfaces[L]) {  // fwd_reassign :2395-2431 and merge_x (src/cluster.cpp:321-328)
                    int row = nn->rsize_org;
                    for (const auto& t : nn->extra_from) {
                        Cluster* c = by_order(t[0]);
                        for (int i = 0; i < t[2]; ++i) nn->x[row + i] = c->x[c->cols() + c->row_perm[t[1] + i]];
                        row += t[2];
                    }
                }
            for (Cluster* c : interfaces[L]) {  // fwd_spars :2957-2963 (+ :2529-2533)
                if (!want_sparsify(c)) continue;
                int m2 = c->Qs.r, k = c->Qs.c;
                gemqrt_lt(m2, 1, k, k, c->Qs.data(), m2, c->Ts.data(), k, c->x.data(), m2);
                if (!square() && g_fault != 8)  // orthogonalD_extra_fwd :2206-2211
                    ormqr_lt_vec(c->nex, c->row_rank, c->Qex.data(), std::max(c->nex, 1), c->tauex.data(),
                                 c->x.data() + c->csize_org);
                ormqr_lt_vec(c->csize_org, c->Qsp.c, c->Qsp.data(), c->Qsp.r, c->tausp.data(), c->x.data());
                if (!square()) split_fwd(c);
            }
        }
    }

    // Backward half (src/ptree.cpp:2600-3210) from the vectors as they stand; x (n entries,
    // permuted columns) gets each leaf's first cols_org entries.
    void solve_bwd(double* xv) {
        for (int L = nlevels - 1; L >= 0; --L) {
            for (Cluster* c : interfaces[L]) {  // bwd_spars :3146-3174
                if (want_sparsify(c)) {
                    if (!square()) split_bwd(c);
                    ormqr_ln_vec(c->csize_org, c->Qsp.c, c->Qsp.data(), c->Qsp.r, c->tausp.data(), c->x.data());
                    trsv_upper(c->csize_org, c->Qs.data(), c->Qs.r, c->x.data());
                }
            }
            for (Cluster* c : interfaces[L])
                for (Edge* e : c->edgesOut) {
                    Cluster* n2 = e->n2;
                    if (n2->lvl() != c->merge_level) continue;
                    int s2 = e->A.c, k2 = n2->cols();
                    for (int j = 0; j < s2; ++j) {
                        double xj = c->x[j];
                        for (int i = 0; i < k2; ++i) n2->x[i] -= e->A(i, j) * xj;
                    }
                }
            for (Cluster* c : interiors[L])  // bwd_trsv :2188-2192
                trsv_upper(c->cols(), c->eself->A.data(), c->eself->A.r, c->x.data());
            if (L > 0)
                for (Cluster* p : bottoms[L]) {  // bwd_merge :3105-3123
                    int k = 0;
                    for (Cluster* ch : p->children)
                        for (int i = 0; i < ch->cols(); ++i) ch->x[i] = p->x[k++];
                }
        }
        for (Cluster* c : bottoms[0])
            for (int i = 0; i < c->csize_org; ++i) xv[c->cstart + i] = c->x[i];
    }

    // ParTree::solve(b, x) (src/ptree.cpp:2267): b has m entries (permuted rows), x n entries
    // (permuted columns); square systems may pass the same array.
    void solve(const double* b, double* x) {
        solve_fwd(b);
        solve_bwd(x);
    }
    void solve(double* xv) { solve(xv, xv); }

    // The transpose of solve_bwd: y (n entries, permuted columns) into the vectors, then level by
    // level upwards the transposes of the backward operations in reverse order -- the parents
    // gather their children's first cols entries (tMerge::fwd, src/toperations.cpp:88-96), the
    // interiors solve with R^T and subtract R_cn^T x_c from each in-neighbour (tQR::fwd :12-27),
    // the sparsified interfaces solve with R_s^T (tScaleD::fwd :71-74), apply Q_sp^T
    // (tOrthogonalD::fwd :79-81) and, for a rectangular system, splitD_fwd with zero extra
    // entries (nothing reads them on the way back).
    void solve_bwd_transpose(const double* y) {
        size_vectors();
        for (Cluster* c : bottoms[0]) std::copy(y + c->cstart, y + c->cstart + c->csize_org, c->x.begin());
        for (int L = 0; L < nlevels; ++L) {
            if (L > 0)
                for (Cluster* p : bottoms[L]) {
                    int k = 0;
                    for (Cluster* ch : p->children)
                        for (int i = 0; i < ch->cols(); ++i) p->x[k++] = ch->x[i];
                }
            for (Cluster* c : interiors[L]) trsv_upper_trans(c->cols(), c->eself->A.data(), c->eself->A.r, c->x.data());
            for (Cluster* c : interfaces[L])
                for (Edge* e : c->edgesOut) {
                    Cluster* n2 = e->n2;
                    if (n2->lvl() != c->merge_level) continue;
                    const int s2 = e->A.c, k2 = n2->cols();
                    for (int j = 0; j < s2; ++j) {
                        double acc = 0.0;
                        for (int i = 0; i < k2; ++i) acc += e->A(i, j) * n2->x[i];
                        c->x[j] -= acc;
                    }
                }
            for (Cluster* c : interfaces[L]) {
                if (!want_sparsify(c)) continue;
                trsv_upper_trans(c->csize_org, c->Qs.data(), c->Qs.r, c->x.data());
                ormqr_lt_vec(c->csize_org, c->Qsp.c, c->Qsp.data(), c->Qsp.r, c->tausp.data(), c->x.data());
                if (!square()) {
                    std::fill(c->x.begin() + c->csize_org, c->x.end(), 0.0);
                    split_fwd(c);
                }
            }
        }
    }

    // solve_nrml (src/tree.cpp:1524-1570) on this tree: x = M^{-1} M^{-T} y, M^{-1} the backward
    // half of the solve -- an approximate (A^T A)^{-1} y for CGLS.
    void solve_nrml(const double* y, double* x) {
        solve_bwd_transpose(y);
        solve_bwd(x);
    }

    void spmv(const double* x, double* y) const {  // src/ptree.cpp:3212-3262: y (m) = App x (n)
        for (int i = 0; i < m; ++i) y[i] = 0.0;
        for (int j = 0; j < n; ++j)
            for (int k = App.colptr[j]; k < App.colptr[j + 1]; ++k) y[App.rowind[k]] += App.val[k] * x[j];
    }
    void spmv_t(const double* r, double* y) const {  // y (n) = App^T r (m), the adjoint in cgls
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = App.colptr[j]; k < App.colptr[j + 1]; ++k) s += App.val[k] * r[App.rowind[k]];
            y[j] = s;
        }
    }
};

static double dot(const std::vector<double>& a, const double* b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// dist_gmres (src/is.cpp:447-527) with x0 = 0, MGS Arnoldi and Givens; break on the
// preconditioned residual as the reference driver does (verb = true).  restart > 0 is an
// extension (restarted GMRES, x0 = current iterate) not present in the reference.
struct GmresResult { int iters; double prec_res; };

static GmresResult gmres_cycle(Tree& t, const std::vector<double>& b, std::vector<double>& x, int iter, double tol,
                               std::vector<double>* hist) {
    const int n = t.n;
    std::vector<double> r(n), tmp(n);
    // r = M^{-1}(b - A x)
    t.spmv(x.data(), tmp.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - tmp[i];
    t.solve(r.data());
    double r_norm = std::sqrt(dot(r, r.data()));
    if (r_norm == 0.0) return {0, 0.0};
    std::vector<double> sn(iter, 0.0), cs(iter, 0.0), beta(iter + 1, 0.0);
    std::vector<std::vector<double>> Q;
    Q.push_back(r);
    for (auto& v : Q[0]) v /= r_norm;
    beta[0] = r_norm;
    std::vector<double> H((size_t)(iter + 1) * iter, 0.0);
    auto h = [&](int i, int j) -> double& { return H[i + (size_t)j * (iter + 1)]; };
    int k = 0;
    double err = 1.0;
    for (; k < iter; ++k) {
        std::vector<double> q(n);
        t.spmv(Q[k].data(), q.data());
        t.solve(q.data());
        for (int i = 0; i < k + 1; ++i) {
            h(i, k) = dot(q, Q[i].data());
            for (int l = 0; l < n; ++l) q[l] -= h(i, k) * Q[i][l];
        }
        h(k + 1, k) = std::sqrt(dot(q, q.data()));
        for (auto& v : q) v /= h(k + 1, k);
        Q.push_back(q);
        for (int i = 0; i < k; ++i) {
            double temp = cs[i] * h(i, k) + sn[i] * h(i + 1, k);
            h(i + 1, k) = -sn[i] * h(i, k) + cs[i] * h(i + 1, k);
            h(i, k) = temp;
        }
        if (h(k, k) == 0) { cs[k] = 0; sn[k] = 1; }
        else {
            double tt = std::sqrt(h(k, k) * h(k, k) + h(k + 1, k) * h(k + 1, k));
            cs[k] = std::fabs(h(k, k)) / tt;
            sn[k] = cs[k] * h(k + 1, k) / h(k, k);
        }
        h(k, k) = cs[k] * h(k, k) + sn[k] * h(k + 1, k);
        h(k + 1, k) = 0.0;
        beta[k + 1] = -sn[k] * beta[k];
        beta[k] = cs[k] * beta[k];
        err = std::fabs(beta[k + 1]) / r_norm;
        if (hist) hist->push_back(err);
        if (err <= tol) break;
    }
    k = std::min(k, iter - 1);
    // LUSolve (src/util.cpp:1019-1032) and x += Q y
    std::vector<double> y(k + 1, 0.0);
    for (int i = k; i >= 0; --i) {
        y[i] = beta[i];
        for (int j = i + 1; j < k + 1; ++j) y[i] -= h(i, j) * y[j];
        y[i] /= h(i, i);
    }
    for (int j = 0; j <= k; ++j)
        for (int l = 0; l < n; ++l) x[l] += Q[j][l] * y[j];
    return {k, err * r_norm};
}

// cgls (src/is.cpp:274-345) with x0 = 0 and the tree's solve_nrml as the preconditioner:
// CG on A^T A x = A^T b, stopping when ||A^T (b - A x)||^2 < tol^2 ||A^T b||^2.  Returns the
// iteration count as the reference counts it; *rel = ||A^T r|| / ||A^T b||.
static int cgls(Tree& t, const double* b, double* x, int max_iters, double tol, double* rel) {
    const int m = t.m, n = t.n;
    std::vector<double> r(b, b + m), s(n), p(n), z(n), q(m);
    std::fill(x, x + n, 0.0);
    t.spmv_t(b, s.data());
    const double rhs2 = dot(s, s.data());
    if (rhs2 == 0.0) { *rel = 0.0; return 0; }
    const double threshold = tol * tol * rhs2;
    double res2 = rhs2;
    if (res2 < threshold) { *rel = std::sqrt(res2 / rhs2); return 0; }
    t.solve_nrml(s.data(), p.data());
    double abs_new = dot(s, p.data());
    int i = 0;
    while (i < max_iters) {
        t.spmv(p.data(), q.data());
        const double alpha = abs_new / dot(q, q.data());
        for (int j = 0; j < n; ++j) x[j] += alpha * p[j];
        for (int j = 0; j < m; ++j) r[j] -= alpha * q[j];
        t.spmv_t(r.data(), s.data());
        res2 = dot(s, s.data());
        if (res2 < threshold) break;
        t.solve_nrml(s.data(), z.data());
        const double abs_old = abs_new;
        abs_new = dot(s, z.data());
        const double beta = abs_new / abs_old;
        for (int j = 0; j < n; ++j) p[j] = z[j] + beta * p[j];
        ++i;
    }
    *rel = std::sqrt(res2 / rhs2);
    return i;
}

}  // namespace orc

// ====================================================================== C API (ctypes)
using namespace orc;

extern "C" {

void* orc_tree_create(int n, const int* colptr, const int* rowind, const double* val, int cn, int cd, int nlevels,
                      double tol, int skip) {
    Tree* t = new Tree(nlevels, skip);
    t->tol = tol;
    t->n = n;
    t->m = n;
    t->A.nr = t->A.nc = n;
    t->A.colptr.assign(colptr, colptr + n + 1);
    t->A.rowind.assign(rowind, rowind + colptr[n]);
    t->A.val.assign(val, val + colptr[n]);
    t->X = linspace_nd(cn, cd);
    t->d = cd;
    return t;
}

// The same tree ordered by given leaf parts (the reference's METIS path) instead of coordinates.
void* orc_tree_create_parts(int n, const int* colptr, const int* rowind, const double* val, const int* parts,
                            int nlevels, double tol, int skip) {
    Tree* t = (Tree*)orc_tree_create(n, colptr, rowind, val, 1, 1, nlevels, tol, skip);
    t->X.clear();
    t->d = 0;
    t->given_parts.assign(parts, parts + n);
    return t;
}

// A rectangular system (m x n, m >= n: least squares) ordered by given leaf parts.
void* orc_tree_create_rect(int m, int n, const int* colptr, const int* rowind, const double* val, const int* parts,
                           int nlevels, double tol, int skip) {
    Tree* t = new Tree(nlevels, skip);
    t->tol = tol;
    t->n = n;
    t->m = m;
    t->A.nr = m;
    t->A.nc = n;
    t->A.colptr.assign(colptr, colptr + n + 1);
    t->A.rowind.assign(rowind, rowind + colptr[n]);
    t->A.val.assign(val, val + colptr[n]);
    t->given_parts.assign(parts, parts + n);
    return t;
}

void orc_tree_destroy(void* h) { delete (Tree*)h; }

int orc_nrows(void* h) { return ((Tree*)h)->m; }
void orc_get_rperm(void* h, int* out) {
    Tree* t = (Tree*)h;
    std::memcpy(out, t->rperm.data(), sizeof(int) * t->m);
}
// Per cluster (order): rstart, rows at creation, extra rows received, row_rank, and for an
// interior the number of its extra rows that went to another cluster.
void orc_get_rows(void* h, int* out /* 5 * ncl */) {
    Tree* t = (Tree*)h;
    for (auto& up : t->owned_c) {
        Cluster* c = up.get();
        int* o = out + 5 * c->order;
        o[0] = c->rstart; o[1] = c->rsize_org; o[2] = c->tot_extra; o[3] = c->row_rank;
        int moved = 0;
        for (int v : c->r2c) moved += (v != c->order);
        o[4] = moved;
    }
}
// Extra-row targets of an interior (cluster order per extra row, in row order); returns the count.
int orc_get_r2c(void* h, int order, int* out, int cap) {
    Tree* t = (Tree*)h;
    Cluster* c = t->owned_c[order].get();
    const int ex = (int)c->r2c.size();
    if (ex > cap) return -1;
    std::copy(c->r2c.begin(), c->r2c.end(), out);
    return ex;
}
void orc_solve_ls(void* h, const double* b, double* x) { ((Tree*)h)->solve(b, x); }
void orc_solve_nrml(void* h, const double* y, double* x) { ((Tree*)h)->solve_nrml(y, x); }
void orc_spmv_t(void* h, const double* r, double* y) { ((Tree*)h)->spmv_t(r, y); }
int orc_cgls(void* h, const double* b, double* x, int iters, double tol, double* rel) {
    try {
        return cgls(*(Tree*)h, b, x, iters, tol, rel);
    } catch (std::exception& e) {
        fprintf(stderr, "oracle: %s\n", e.what());
        return -1;
    }
}

int orc_partition_assemble(void* h) {
    try {
        Tree* t = (Tree*)h;
        t->partition();
        t->assemble();
        return 0;
    } catch (std::exception& e) {
        fprintf(stderr, "oracle: %s\n", e.what());
        return -1;
    }
}

int orc_factorize(void* h) {
    try {
        ((Tree*)h)->factorize();
        return 0;
    } catch (std::exception& e) {
        fprintf(stderr, "oracle: %s\n", e.what());
        return -1;
    }
}

double orc_factor_seconds(void* h) { return ((Tree*)h)->t_factor; }

void orc_get_cperm(void* h, int* out) {
    Tree* t = (Tree*)h;
    std::memcpy(out, t->cperm.data(), sizeof(int) * t->n);
}

// Cluster table: per cluster (order) -> merge_level, self.lvl, self.sep, l.lvl, l.sep, r.lvl,
// r.sep, section, part, cstart, rows_at_creation(cols_org), parent order, final rank (or -1).
int orc_num_clusters(void* h) { return (int)((Tree*)h)->owned_c.size(); }
void orc_get_clusters(void* h, int* out /* 13 * ncl */) {
    Tree* t = (Tree*)h;
    for (auto& up : t->owned_c) {
        Cluster* c = up.get();
        int* o = out + 13 * c->order;
        o[0] = c->merge_level; o[1] = c->id.self.lvl; o[2] = c->id.self.sep;
        o[3] = c->id.l.lvl; o[4] = c->id.l.sep; o[5] = c->id.r.lvl; o[6] = c->id.r.sep;
        o[7] = c->id.section; o[8] = c->id.part; o[9] = c->cstart; o[10] = c->csize_org;
        o[11] = c->parent ? c->parent->order : -1;
        o[12] = c->sparsified ? c->rank : -1;
    }
}

// Per-interior factor blocks for parity: diagonal R (upper triangle of A(c,c)).
int orc_get_diag_R(void* h, int order, double* out, int maxn) {
    Tree* t = (Tree*)h;
    Cluster* c = t->owned_c[order].get();
    if (!c->eself) return -1;
    int k = c->eself->A.c;
    if ((size_t)k * k > (size_t)maxn) return -2;
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) out[i + (size_t)j * k] = (i <= j) ? c->eself->A(i, j) : 0.0;
    return k;
}

void orc_solve(void* h, double* x) { ((Tree*)h)->solve(x); }
void orc_spmv(void* h, const double* x, double* y) { ((Tree*)h)->spmv(x, y); }

// GMRES on App x = b (permuted space).  restart <= 0: one cycle of `iters` (the reference).
// Returns total iterations (0-based k of the last cycle + cycles*restart, as counted by the
// reference's returned k); writes x and the preconditioned-residual history.
int orc_gmres(void* h, const double* b, double* x, int iters, double tol, int restart, double true_tol,
              double* hist, int hist_cap, int* hist_len) {
    Tree* t = (Tree*)h;
    std::vector<double> bb(b, b + t->n), xx(t->n, 0.0), hv;
    int total = 0;
    if (restart <= 0) {
        GmresResult r = gmres_cycle(*t, bb, xx, iters, tol, &hv);
        total = r.iters;
    } else {
        int done = 0;
        double bn = std::sqrt(dot(bb, bb.data()));
        while (done < iters) {
            int cyc = std::min(restart, iters - done);
            GmresResult r = gmres_cycle(*t, bb, xx, cyc, tol, &hv);
            done += r.iters + 1;
            std::vector<double> ax(t->n);
            t->spmv(xx.data(), ax.data());
            double rn = 0.0;
            for (int i = 0; i < t->n; ++i) rn += (bb[i] - ax[i]) * (bb[i] - ax[i]);
            if (std::sqrt(rn) / bn <= true_tol) break;
        }
        total = done - 1;
    }
    std::memcpy(x, xx.data(), sizeof(double) * t->n);
    int nh = std::min((int)hv.size(), hist_cap);
    for (int i = 0; i < nh; ++i) hist[i] = hv[i];
    *hist_len = nh;
    return total;
}

// Trace of dense work: 5 ints per record (level, op, m, n, k).
int orc_trace_len(void* h) { return (int)((Tree*)h)->trace.size(); }
void orc_get_trace(void* h, int* out) {
    Tree* t = (Tree*)h;
    for (size_t i = 0; i < t->trace.size(); ++i) {
        auto& r = t->trace[i];
        out[5 * i] = r.level; out[5 * i + 1] = r.op; out[5 * i + 2] = r.m; out[5 * i + 3] = r.n; out[5 * i + 4] = r.k;
    }
}

// ---------------- dense kernels for unit tests (column-major, lda = rows)
void orc_geqrt(int m, int n, int nb, double* A, double* T) { geqrt(m, n, nb, A, m, T, nb); }
void orc_tpqrt(int m, int n, int nb, double* A, double* B, double* T) { tpqrt(m, n, nb, A, n, B, m, T, nb); }
void orc_tpmqrt_lt(int m, int n, int k, int nb, const double* V, const double* T, double* A, double* B) {
    tpmqrt_lt(m, n, k, nb, V, m, T, nb, A, k, B, m);
}
void orc_larft(int m, int k, const double* V, const double* tau, double* T) { larft(m, k, V, m, tau, T, k); }
void orc_larfb_lt(int m, int n, int k, const double* V, const double* T, double* C) {
    larfb_left_trans(m, n, k, V, m, T, k, C, m);
}
void orc_larfb_rn(int m, int n, int k, const double* V, const double* T, double* C) {
    larfb_right_notrans(m, n, k, V, n, T, k, C, m);
}
void orc_geqp3(int m, int n, double* A, int* jpvt, double* tau) { geqp3(m, n, A, m, jpvt, tau); }
void orc_geqp3_tie(int m, int n, double* A, int* jpvt, double* tau, double tie_rel) {
    geqp3(m, n, A, m, jpvt, tau, tie_rel);
}
// C (m x n) <- Q^T C with Q from geqrt(m, k, nb) (V m x k, T nb x k)
void orc_gemqrt_lt(int m, int n, int k, int nb, const double* V, const double* T, double* C) {
    gemqrt_lt(m, n, k, nb, V, m, T, nb, C, m);
}
int orc_choose_rank(const double* s, int n, double tol) { return choose_rank(s, n, tol); }
void orc_set_fault(int f) { g_fault = f; }
int orc_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
// rank path (rank_path.hpp)
void orc_gauss(int rows, int cols, uint64_t seed, double* om) {
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i) om[i + (size_t)j * rows] = gauss_entry(i, j, seed);
}
double orc_colnorm(int k, int w, const double* C, double* cn) {
    double mx = 0.0;
    colnorm2(k, w, C, cn, &mx);
    return mx;
}
void orc_sketch(int k, int w, int s, const double* C, const double* om, int ldo, double* Y) { sketch(k, w, s, C, om, ldo, Y); }
void orc_rank_features(int k, int w, const double* rdiag, const double* cn, double cmax2, int nsout, int nsin, int level,
                       double tol, float* x) {
    rank_features(k, w, rdiag, cn, cmax2, nsout, nsin, level, tol, x);
}
int orc_tail_select(int k, int w, const double* B, int ry, double cmax2, double tol) {
    return tail_select(k, w, B, ry, cmax2, tol);
}
int orc_sketch_rank(int k, int w, int s, const double* C, const double* om, int ldo, double tol, int pivoted) {
    return sketch_rank(k, w, s, C, om, ldo, tol, pivoted != 0);
}
void orc_set_near_tie_band(double b) { geqp3_near_tie_band() = b; }
long orc_larfg_tiny_alpha(int reset) {
    long v = larfg_tiny_alpha();
    if (reset) larfg_tiny_alpha() = 0;
    return v;
}
long orc_geqp3_near_ties(int reset) {
    long v = geqp3_near_ties();
    if (reset) geqp3_near_ties() = 0;
    return v;
}
void orc_trsm_right_upper(int m, int n, const double* R, double* B) { trsm_right_upper(m, n, R, n, B, m); }
void orc_trsv_upper(int n, const double* R, double* x) { trsv_upper(n, R, n, x); }
void orc_ormqr_lt(int m, int k, const double* V, const double* tau, double* x) { ormqr_lt_vec(m, k, V, m, tau, x); }
void orc_ormqr_ln(int m, int k, const double* V, const double* tau, double* x) { ormqr_ln_vec(m, k, V, m, tau, x); }

}  // extern "C"
