// Level planner for the MI355X spaQR path (host side of the reference's ParTree,
// src/ptree.cpp).  Symbolic work (ordering, hierarchy, fill-in, merge bookkeeping) runs on
// the host over flat arrays; each tree level's numeric work is a short sequence of grouped
// batched launches (kernels.hip) over descriptor tables uploaded once per phase.
//
// Reference behaviour restated here (file:line):
//   partition      GeometricPartitionAtA src/partition.cpp:282-350, bissect_geo :241-279,
//                  getInterfacesUsingA :514-587, initPermutation src/tree.cpp:48-57,
//                  merge_if src/cluster.cpp:23-32, hierarchy src/ptree.cpp:587-691
//   assemble       src/ptree.cpp:696-864
//   factorize      src/ptree.cpp:873-2154 (level loop), task bodies :170-493, 1621-1633,
//                  1867-1897, merge :1940-2123 / compute_new_edges :289-340
//   solve          src/ptree.cpp:2788-3210
//   spmv           src/ptree.cpp:3212-3262
//   GMRES          dist_gmres src/is.cpp:447-527
// Orders that the reference leaves to its task scheduler are fixed to cluster order:
// interiors of a level touching a shared block run in colour layers where every interior
// comes after all lower-ordered interiors it conflicts with, which reproduces sequential
// cluster order exactly.
#include "tree.hpp"
#include "symbolic.hpp"

#include <algorithm>
#include <array>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <queue>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>

#include "kernels.hpp"
#include "launch.hpp"
#include "rrqr.hpp"

#include "status.hpp"
#include "tree_impl.hpp"

namespace spaqr {

namespace {
struct NumericalFailure : std::runtime_error {
    using std::runtime_error::runtime_error;
};
}  // namespace

// Reads the per-level status flags written by the pivot checks so far (the stream has been
// synchronised) and ends the factorization on the first level that raised one.
void Tree::check_levels(int upto) {
    if (!d_lflags) return;
    std::vector<int> f(nlevels), first(nlevels);
    HIPCHK(hipMemcpy(f.data(), d_lflags, sizeof(int) * nlevels, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(first.data(), d_lfirst, sizeof(int) * nlevels, hipMemcpyDeviceToHost));
    for (int L = 0; L <= std::min(upto, nlevels - 1); ++L)
        if (f[L]) {
            std::string what = (f[L] & kInfoNonFinite) ? "a non-finite entry" : "a zero pivot";
            if ((f[L] & kInfoZeroPivot) && pivot_rel > 0) what = "a pivot at or below the set threshold";
            throw NumericalFailure("numerical failure: " + what + " in the R factor of interior cluster " +
                                   std::to_string(first[L]) + " at level " + std::to_string(L) +
                                   " (spaqr_tree_get_info lists every flagged cluster)");
        }
}

int Tree::get_info(std::vector<int>& out) {
    out.assign(cl.size(), 0);
    if (d_cinfo) HIPCHK(hipMemcpy(out.data(), d_cinfo, sizeof(int) * cl.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < out.size() && i < h_info.size(); ++i) out[i] |= h_info[i];
    return (int)out.size();
}

bool ClusterID::operator<(const ClusterID& o) const {  // include/cluster.h:128-134
    if (self != o.self) return self < o.self;
    if (l != o.l) return l < o.l;
    if (r != o.r) return r < o.r;
    if (section != o.section) return section < o.section;
    return part < o.part;
}

namespace {

SepID up(const SepID& s) { return SepID(s.lvl + 1, s.sep / 2); }

ClusterID merge_if(const ClusterID& c, int lvl) {
    ClusterID p;
    p.self = c.self;
    p.l = c.l.lvl < lvl ? up(c.l) : c.l;
    p.r = c.r.lvl < lvl ? up(c.r) : c.r;
    p.section = lvl > 0 ? c.section / 2 : c.section;
    p.part = (c.self.lvl == lvl) ? 2 : c.part;
    return p;
}

SepID highest_common(SepID a, SepID b) {
    while (a.lvl < b.lvl) { a.lvl++; a.sep /= 2; }
    while (b.lvl < a.lvl) { b.lvl++; b.sep /= 2; }
    while (a.sep != b.sep) { a.lvl++; b.lvl++; a.sep /= 2; b.sep /= 2; }
    return a;
}

SepID lowest_common(SepID a, SepID b) { return a.lvl < b.lvl ? a : b; }

HostCSC csc_transpose(const HostCSC& A) {
    HostCSC T;
    T.nr = A.nc;
    T.nc = A.nr;
    const int nnz = A.colptr[A.nc];
    T.colptr.assign(A.nr + 1, 0);
    for (int k = 0; k < nnz; ++k) T.colptr[A.rowind[k] + 1]++;
    for (int i = 0; i < A.nr; ++i) T.colptr[i + 1] += T.colptr[i];
    T.rowind.resize(nnz);
    T.val.resize(nnz);
    std::vector<int> pos(T.colptr.begin(), T.colptr.end() - 1);
    for (int j = 0; j < A.nc; ++j)
        for (int k = A.colptr[j]; k < A.colptr[j + 1]; ++k) {
            int p = pos[A.rowind[k]]++;
            T.rowind[p] = j;
            T.val[p] = A.val[k];
        }
    return T;
}

// adjacency of A^T A (shared rows), diagonal included, sorted
HostCSC ata_pattern(const HostCSC& A) {
    HostCSC At = csc_transpose(A);
    HostCSC G;
    G.nr = G.nc = A.nc;
    G.colptr.assign(A.nc + 1, 0);
    std::vector<int> mark(A.nc, -1), tmp;
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
    return G;
}

HostCSC permute_rc(const HostCSC& A, const std::vector<int>& rperm, const std::vector<int>& cperm) {  // A(rperm, cperm)
    std::vector<int> inv(A.nr);
    for (int i = 0; i < A.nr; ++i) inv[rperm[i]] = i;
    HostCSC P;
    P.nr = A.nr;
    P.nc = A.nc;
    P.colptr.assign(A.nc + 1, 0);
    std::vector<std::pair<int, double>> col;
    for (int j = 0; j < A.nc; ++j) {
        const int oj = cperm[j];
        col.clear();
        for (int k = A.colptr[oj]; k < A.colptr[oj + 1]; ++k) col.emplace_back(inv[A.rowind[k]], A.val[k]);
        std::sort(col.begin(), col.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (auto& e : col) { P.rowind.push_back(e.first); P.val.push_back(e.second); }
        P.colptr[j + 1] = (int)P.rowind.size();
    }
    return P;
}

}  // namespace

// ============================================================================ arena
void* Tree::Arena::alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    while (bytes > left) {
        // move on to the next retained chunk if it is large enough, else allocate a new one
        if (next < chunks.size() && chunks[next].second >= bytes) {
            cur = (char*)chunks[next].first;
            left = chunks[next].second;
            ++next;
            break;
        }
        if (next < chunks.size()) { ++next; continue; }  // too small for this request: skip it
        const size_t want = std::max(bytes, (size_t)256 << 20);
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, want));
        chunks.push_back({p, want});
        next = chunks.size();
        cur = (char*)p;
        left = want;
    }
    void* r = cur;
    cur += bytes;
    left -= bytes;
    total += bytes;
    return r;
}

void Tree::Arena::reset() {
    next = 0;
    cur = nullptr;
    left = 0;
    total = 0;
}

void Tree::Arena::release() {
    for (auto& c : chunks) (void)hipFree(c.first);
    chunks.clear();
    reset();
}

char* Tree::PinnedArena::alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    while (bytes > left) {
        if (next < chunks.size() && chunks[next].second >= bytes) {
            cur = chunks[next].first;
            left = chunks[next].second;
            ++next;
            break;
        }
        if (next < chunks.size()) { ++next; continue; }
        const size_t want = std::max(bytes, (size_t)64 << 20);
        void* p = nullptr;
        HIPCHK(hipHostMalloc(&p, want, hipHostMallocDefault));
        chunks.push_back({(char*)p, want});
        next = chunks.size();
        cur = (char*)p;
        left = want;
    }
    char* r = cur;
    cur += bytes;
    left -= bytes;
    return r;
}

void Tree::PinnedArena::release() {
    for (auto& c : chunks) (void)hipHostFree(c.first);
    chunks.clear();
    reset();
}

void* Tree::upload(const void* h, size_t bytes) {
    void* d = desc_arena.alloc(bytes);
    char* p = staging.alloc(bytes);
    constexpr size_t kPiece = (size_t)1 << 20;
    if (bytes >= 8 * kPiece) {  // large descriptor lists: copy 1 MiB pieces on several threads
        const int pieces = (int)((bytes + kPiece - 1) / kPiece);
        parallel_ranges(pieces, 1, [&](int b, int e, int) {
            const size_t lo = (size_t)b * kPiece, hi = std::min(bytes, (size_t)e * kPiece);
            memcpy(p + lo, (const char*)h + lo, hi - lo);
        });
    } else {
        memcpy(p, h, bytes);
    }
    h2d(d, p, bytes);
    return d;
}

// Host-to-device descriptor copy: queued on the tree's stream, or, while that stream is being
// captured into a graph, done at once on a separate stream, so that the graph holds only the
// kernels and device copies and its replays read the descriptors where they already are.
void Tree::h2d(void* d, const void* h, size_t bytes) {
    if (!capturing) {
        HIPCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream_));
        return;
    }
    if (!up_stream_) HIPCHK(hipStreamCreateWithFlags(&up_stream_, hipStreamNonBlocking));
    HIPCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, up_stream_));
    HIPCHK(hipStreamSynchronize(up_stream_));
}

// ---------------------------------------------------------------------------- graphs
std::vector<long long> Tree::plan_signature() const {
    long long tb = 0;
    static_assert(sizeof(tb) == sizeof(tol), "tolerance bits");
    std::memcpy(&tb, &tol, sizeof(tb));
    std::vector<long long> s = {prec, asm_gen, tb, skip, rank_mode};
    for (int p = 0; p < T_NPARAM; ++p) s.push_back(tune[p]);
    return s;
}

void Tree::seg_begin() {
    if (!graph_rec) return;
    HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
    capturing = true;
}

void Tree::seg_end(int* ranks_dev, int nranks, int* err_dev) {
    if (!graph_rec || !capturing) return;
    hipGraph_t g = nullptr;
    capturing = false;
    HIPCHK(hipStreamEndCapture(stream_, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    gplan.exec.push_back(ex);
    gplan.rb_ranks_dev.push_back(ranks_dev);
    gplan.rb_err_dev.push_back(err_dev);
    gplan.rb_ranks.emplace_back((size_t)std::max(nranks, 0), 0);
    HIPCHK(hipGraphLaunch(ex, stream_));
}

void Tree::drop_graphs() {
    if (capturing) {  // an error while recording: end the capture and discard it
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        capturing = false;
    }
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (auto ex : gplan.exec) (void)hipGraphExecDestroy(ex);
    gplan = GraphPlan();
}

int Tree::set_graphs(bool on) {
    if (!on) drop_graphs();
    graphs = on;
    return 0;
}

template <typename S>
int Tree::replay_t() {
    auto t0 = std::chrono::steady_clock::now();
    sync_stream();
    factorized = false;
    for (size_t i = 0; i < gplan.exec.size(); ++i) {
        HIPCHK(hipGraphLaunch(gplan.exec[i], stream_));
        if (!gplan.rb_ranks_dev[i]) continue;
        std::vector<int> r(gplan.rb_ranks[i].size());
        int err = 0;
        HIPCHK(hipMemcpyAsync(r.data(), gplan.rb_ranks_dev[i], sizeof(int) * r.size(), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(&err, gplan.rb_err_dev[i], sizeof(int), hipMemcpyDeviceToHost, stream_));
        sync_stream();
        if (err) throw std::runtime_error("cooperative geqp3: a barrier wait exceeded its spin limit");
        if (r != gplan.rb_ranks[i]) return 1;  // the plan no longer fits this matrix
    }
    int chain_err = 0;
    HIPCHK(hipMemcpyAsync(&chain_err, gplan.chain_err, sizeof(int), hipMemcpyDeviceToHost, stream_));
    sync_stream();
    if (chain_err) throw std::runtime_error("cooperative tpqrt: a wait exceeded its spin limit");
    check_levels(nlevels - 1);
    factorized = true;
    ++graph_replays;
    t_factor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

void Tree::sync_stream() {
    if (!stream_) {  // nothing was ever issued
        staging.reset();
        return;
    }
    HIPCHK(hipStreamSynchronize(stream_));
    staging.reset();  // every upload queued so far has completed
    for (auto& p : prof_pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        prof[p.first].ms += ms;
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    prof_pending.clear();
}

void Tree::prof_reset() {
    for (auto& p : prof) p = ProfEntry();
}

int Tree::sync() {
    try {
        sync_stream();
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -100;
    }
}

// ============================================================================ tree
struct Tree::SolveSchedule {
    size_t off = 0;  // workspace entries assigned so far
    int built = 0;   // levels whose schedule is built
    // algorithmic work of one application: factor entries read (V, T, R, Qs, Ts, Qsp, Tsp and
    // the off-diagonal blocks of the back substitution) and vector entries read and written
    double flops = 0.0, bytes = 0.0;
    struct Level {
        void* merge_fwd = nullptr; int n_merge_fwd = 0;
        void* q = nullptr; int n_q = 0;
        // per colour, the chains by kernel: 0 one wave (k <= 64, rows <= 64), 1 / 2 pipelined
        // workgroup (k <= 512, rows <= 256 / 512), 3 plain workgroup
        std::vector<std::array<std::pair<void*, int>, 4>> chain;
        void* chain_links = nullptr;
        void* spars = nullptr; int n_spars = 0;
        void* trsv = nullptr; int n_trsv = 0;
        void* trsv_links = nullptr;
        void* merge_bwd = nullptr; int n_merge_bwd = 0;
        // rectangular systems: the reassigned rows' gather into the interfaces (fwd_reassign +
        // merge_x), Q_ex^T on the extra rows, and splitD as moves through the scratch region
        // (forward: positions sp_a -> sp_b; backward: sp_b -> sp_a)
        void* rg_si = nullptr; void* rg_di = nullptr; int n_rg = 0;
        void* qex = nullptr; int n_qex = 0;
        void* sp_a = nullptr; void* sp_b = nullptr; int n_sp = 0;
        // the transposed backward pass (solve_nrml): children's first cols entries into the
        // parents, transposed couplings per interface, and (rectangular) splitD_fwd with zero
        // extra entries (ts_a: source position or -1)
        void* merge_tr = nullptr; int n_merge_tr = 0;
        void* gemvt = nullptr; int n_gemvt = 0;
        void* gemvt_links = nullptr;
        void* ts_a = nullptr; void* ts_b = nullptr; int n_ts = 0;
    };
    std::vector<Level> lv;
    // rectangular systems: the level-0 loads of b (rows) and y (columns) and the store of x, as
    // moves between the scratch region (m entries, the FP64 vector converted) and the segments
    void* scratch = nullptr;
    void* iota = nullptr;  // 0, 1, 2, ... (the scratch side of every move)
    void* ld_si = nullptr; void* ld_di = nullptr; int n_ld = 0;
    void* lc_si = nullptr; void* lc_di = nullptr; int n_lc = 0;
    void* st_si = nullptr; void* st_di = nullptr; int n_st = 0;
    bool tr_built = false;  // the transposed pass's tasks (built on solve_nrml's first use)
};

Tree::Tree(int nl, int sk) : nlevels(nl), skip(sk) {
    bottoms.resize(nl);
    interiors.resize(nl);
    interfaces.resize(nl);
}

// The stream is made on the first call that uses the GPU (assemble): ordering and the cluster
// hierarchy are host work and run on a machine without one.
void Tree::ensure_stream() {
    if (!stream_) HIPCHK(hipStreamCreate(&stream_));
}

// Caller-given leaf parts of one system (positions off .. off + n): one per column, each in
// [0, 2^(nlevels-1)), the numbering of a recursive bisection (part k of depth d splits into 2k
// and 2k + 1; METIS_PartGraphRecursive with 2^(nlevels-1) parts numbers them so).
static void check_parts(const std::vector<int>& parts, int off, int n, int nlevels) {
    if ((long long)parts.size() < (long long)off + n) throw std::runtime_error("one leaf part per column");
    const int np = 1 << std::max(0, nlevels - 1);
    for (int i = off; i < off + n; ++i)
        if (parts[i] < 0 || parts[i] >= np)
            throw std::runtime_error("leaf part " + std::to_string(parts[i]) + " of column " + std::to_string(i - off) +
                                     " is outside [0, 2^(levels-1))");
}

// The matrix a tree accepts (the reference asserts these: src/is.cpp:460 on a NaN residual,
// src/ptree.cpp:171-175 exits on an empty interior): finite values, no empty column.
static void check_matrix(const HostCSC& A) {
    for (int j = 0; j < A.nc; ++j) {
        bool nonzero = false;
        for (int k = A.colptr[j]; k < A.colptr[j + 1]; ++k) {
            if (!std::isfinite(A.val[k]))
                throw std::runtime_error("A(" + std::to_string(A.rowind[k]) + ", " + std::to_string(j) + ") is not finite");
            if (A.rowind[k] < 0 || A.rowind[k] >= A.nr) throw std::runtime_error("row index out of range in column " + std::to_string(j));
            nonzero = nonzero || A.val[k] != 0.0;
        }
        if (!nonzero) throw std::runtime_error("column " + std::to_string(j) + " of A is zero: A is singular");
    }
}

Tree::~Tree() {
    drop_graphs();
    if (up_stream_) (void)hipStreamDestroy(up_stream_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    delete sched;
    arena.release();
    desc_arena.release();
    staging.release();
    if (d_rowptr) (void)hipFree(d_rowptr);
    if (d_col) (void)hipFree(d_col);
    if (d_val) (void)hipFree(d_val);
    if (d_tptr) (void)hipFree(d_tptr);
    if (d_tcol) (void)hipFree(d_tcol);
    if (d_tval) (void)hipFree(d_tval);
    if (d_scratch) (void)hipFree(d_scratch);
    if (d_blocks0) (void)hipFree(d_blocks0);
    if (xw) (void)hipFree(xw);
    if (d_krylov) (void)hipFree(d_krylov);
    if (d_seg) (void)hipFree(d_seg);
    if (d_mlp_w) (void)hipFree(d_mlp_w);
    if (d_omega) (void)hipFree(d_omega);
    if (d_omega_bf) (void)hipFree(d_omega_bf);
    if (stream_) (void)hipStreamDestroy(stream_);
}

bool Tree::want_sparsify(const Cluster& c) const {  // src/ptree.cpp:29-41, with the levels of c's system
    if (c.merge_level < skip || c.merge_level >= inst_lv[c.inst] - 2 || tol == 0) return false;
    if (c.lvl() > c.merge_level && c.id.l.lvl <= c.merge_level && c.id.r.lvl <= c.merge_level) return true;
    return false;
}

int Tree::find_out(const Cluster& c, int n2) const {
    for (int e : c.out)
        if (ed[e].n2 == n2) return e;
    return -1;
}

void Tree::sort_out(Cluster& c) {
    std::sort(c.out.begin(), c.out.end(), [this](int a, int b) { return ed[a].n2 < ed[b].n2; });
}

// Colours of one level's interiors (two interiors conflict when they share an out-neighbour):
// DSatur -- the uncoloured interior with the most distinct colours among its conflicts goes
// next (ties: more conflicts, then earlier in cluster order) and takes the smallest colour no
// conflict holds.  Interiors then update shared blocks colour by colour.  Returns the number
// of colours.
// The (out-neighbour, interior index) pairs the colouring works from: the only part that reads
// the edge list.
std::vector<std::pair<int, int>> Tree::colour_pairs(const std::vector<int>& ints) const {
    std::vector<std::pair<int, int>> by_nb;
    for (int i = 0; i < (int)ints.size(); ++i)
        for (size_t j = 1; j < cl[ints[i]].out.size(); ++j) by_nb.push_back({ed[cl[ints[i]].out[j]].n2, i});
    return by_nb;
}

int Tree::colour_interiors(const std::vector<int>& ints) {
    std::vector<int> col;
    const int ncol = dsatur_colours(colour_pairs(ints), (int)ints.size(), col);
    for (int i = 0; i < (int)ints.size(); ++i) cl[ints[i]].color = col[i];
    return ncol;
}

// Interiors sharing an out-neighbour conflict; DSatur over that graph (V interiors, pairs from
// colour_pairs).  Touches no tree state, so it may run beside the fill-in.
int Tree::dsatur_colours(std::vector<std::pair<int, int>> by_nb, int V, std::vector<int>& col) {
    // interiors by out-neighbour, then conflicts as sorted unique lists
    std::sort(by_nb.begin(), by_nb.end());
    std::vector<std::vector<int>> adj(V);
    for (size_t a = 0; a < by_nb.size();) {
        size_t b = a;
        while (b < by_nb.size() && by_nb[b].first == by_nb[a].first) ++b;
        for (size_t x = a; x < b; ++x)
            for (size_t y = a; y < b; ++y)
                if (x != y) adj[by_nb[x].second].push_back(by_nb[y].second);
        a = b;
    }
    for (auto& a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    // DSatur: the next interior is the uncoloured one of most distinct neighbour colours, then
    // most conflicts, then lowest index.  The last two keys are fixed, so they become one rank
    // (ord), and the uncoloured interiors sit in one min-heap of ranks per saturation; an entry is
    // stale once its interior is coloured or has moved up, and is dropped when it surfaces.
    col.assign(V, -1);
    std::vector<int> sat(V, 0), ord(V), rk(V);
    for (int i = 0; i < V; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) {
        return adj[x].size() != adj[y].size() ? adj[x].size() > adj[y].size() : x < y;
    });
    for (int r = 0; r < V; ++r) rk[ord[r]] = r;
    using MinHeap = std::priority_queue<int, std::vector<int>, std::greater<int>>;
    std::vector<MinHeap> heap(1);
    {
        std::vector<int> all(ord.size());
        for (int r = 0; r < V; ++r) all[r] = r;
        heap[0] = MinHeap(std::greater<int>(), std::move(all));
    }
    std::vector<std::vector<char>> seen(V);
    int ncol = 0, top = 0;
    for (int done = 0; done < V; ++done) {
        int i = -1;
        for (; top >= 0; --top) {
            MinHeap& h = heap[top];
            while (!h.empty() && (col[ord[h.top()]] >= 0 || sat[ord[h.top()]] != top)) h.pop();
            if (!h.empty()) {
                i = ord[h.top()];
                h.pop();
                break;
            }
        }
        if (i < 0) throw std::runtime_error("interior colouring lost an interior");
        int c = 0;
        while (c < (int)seen[i].size() && seen[i][c]) ++c;
        col[i] = c;
        ncol = std::max(ncol, c + 1);
        for (int j : adj[i]) {
            if (col[j] >= 0) continue;
            if ((int)seen[j].size() <= c) seen[j].resize(c + 1, 0);
            if (seen[j][c]) continue;
            seen[j][c] = 1;
            const int sj = ++sat[j];
            if ((int)heap.size() <= sj) heap.resize(sj + 1);
            heap[sj].push(rk[j]);
            top = std::max(top, sj);
        }
    }
    return ncol;
}

int Tree::new_edge(int c1, int c2) {  // new_edgeOut, src/ptree.cpp:101-120
    Edge e;
    e.n1 = c1;
    e.n2 = c2;
    e.rows = cl[c2].rows;
    e.cols = cl[c1].cols;
    e.spars = cl[c1].lvl() > current_bottom && cl[c2].lvl() > current_bottom && c1 != c2;
    int id = (int)ed.size();
    ed.push_back(e);
    cl[c1].out.push_back(id);
    if (e.spars) cl[c1].sout.push_back(id);
    if (c1 != c2) {
        cl[c2].in.push_back(id);
        if (e.spars) cl[c2].sin.push_back(id);
    } else {
        cl[c1].self_edge = id;
    }
    return id;
}

// ---------------------------------------------------------------------------- partition
// Leaf parts of a graph without coordinates, in place of METIS_PartGraphRecursive
// (src/partition.cpp:354-370): recursive bisection of the graph G (A^T A) into 2^(nlevels-1)
// parts, part k of depth d splitting into 2k and 2k + 1 (the numbering the separator rules read).
// Each part's induced subgraph is ordered by level structures: component by component (in the
// order of their smallest vertex), a breadth-first search from the vertex the first search
// reaches last (a pseudo-peripheral vertex); the first ceil(N/2) vertices of that order form
// part 2k.  Neighbours are visited in increasing order, so the result is deterministic.
static void graph_parts(const HostCSC& G, int nlevels, std::vector<int>& parts) {
    const int n = G.nc;
    parts.assign(n, 0);
    std::vector<int> mark(n, 0), done(n, 0), order, comp, bfs, next(n, 0);
    int stamp = 0;
    auto search = [&](int src, int k, std::vector<int>& out) {
        out.clear();
        ++stamp;
        mark[src] = stamp;
        out.push_back(src);
        for (size_t h = 0; h < out.size(); ++h) {
            const int v = out[h];
            for (int q = G.colptr[v]; q < G.colptr[v + 1]; ++q) {
                const int u = G.rowind[q];
                if (parts[u] == k && mark[u] != stamp) {
                    mark[u] = stamp;
                    out.push_back(u);
                }
            }
        }
    };
    for (int depth = 0; depth < nlevels - 1; ++depth) {
        const int np = 1 << depth;
        std::vector<std::vector<int>> dofs(np);
        for (int i = 0; i < n; ++i) dofs[parts[i]].push_back(i);
        for (int k = 0; k < np; ++k) {
            const auto& d = dofs[k];
            const int N = (int)d.size();
            if (N == 0) continue;
            order.clear();
            for (int s : d) {
                if (done[s]) continue;
                search(s, k, comp);           // the component of s
                search(comp.back(), k, bfs);  // level structure from its far end
                for (int v : bfs) done[v] = 1;
                order.insert(order.end(), bfs.begin(), bfs.end());
            }
            for (int v : d) done[v] = 0;
            const int left = N - N / 2;
            for (int q = 0; q < N; ++q) next[order[q]] = (q < left) ? 2 * k : 2 * k + 1;
        }
        parts.swap(next);  // the searches above read this depth's parts only
    }
}

// Ordering of one square system (A n x n): its column permutation and the cluster ID of every
// permuted position.  Leaf parts come from the coordinates X (d x n, geometric bisection), else
// from the caller (gparts), else from graph_parts; separators follow the same rule on every path;
// interfaces are inferred from the rows of A on the geometric path (getInterfacesUsingA) and
// from A^T A on the others (the METIS path: getInterfacesUsingATA, src/partition.cpp:637-682).
static void order_system(const HostCSC& A, const double* Xc, int xdim, const int* gparts, int nlevels,
                         std::vector<int>& cperm_, std::vector<ClusterID>& cpermed, std::vector<int>* parts_out) {
    const int n = A.nc;
    auto X = [&](size_t idx) { return Xc[idx]; };
    static const bool dbg = getenv("SPAQR_DEBUG_HOST") != nullptr;
    auto tp = std::chrono::steady_clock::now();
    auto lap = [&](const char* what) {
        if (!dbg) return;
        const auto nw = std::chrono::steady_clock::now();
        fprintf(stderr, "[partition] %s %.1f ms\n", what, std::chrono::duration<double, std::milli>(nw - tp).count());
        tp = nw;
    };
    HostCSC G = ata_pattern(A);
    lap("A^T A pattern");
    std::vector<int> parts(n, 0);
    std::vector<double> vals;
    if (!Xc && gparts) parts.assign(gparts, gparts + n);
    else if (!Xc) graph_parts(G, nlevels, parts);
    // recursive geometric bisection (src/partition.cpp:241-304)
    for (int depth = 0; Xc && depth < nlevels - 1; ++depth) {
        const int np = 1 << depth;
        std::vector<std::vector<int>> dofs(np);
        for (int i = 0; i < n; ++i) dofs[parts[i]].push_back(i);
        for (int k = 0; k < np; ++k) {
            const auto& d = dofs[k];
            const int N = (int)d.size();
            if (N == 0) continue;
            int best = -1;
            double range = -1.0;
            for (int dim = 0; dim < xdim; ++dim) {
                double lo = INFINITY, hi = -INFINITY;
                for (int i : d) { double v = X(dim + (size_t)i * xdim); lo = std::min(lo, v); hi = std::max(hi, v); }
                if (hi - lo > range) { range = hi - lo; best = dim; }
            }
            vals.resize(N);
            for (int i = 0; i < N; ++i) vals[i] = X(best + (size_t)d[i] * xdim);
            std::nth_element(vals.begin(), vals.begin() + N / 2, vals.end());
            const double mid = vals[N / 2];
            for (int i : d) parts[i] = (X(best + (size_t)i * xdim) <= mid) ? 2 * k : 2 * k + 1;
        }
    }
    lap("bisection");
    // separators: left-part vertices with an AtA neighbour in the right part (:306-333)
    std::vector<ClusterID> cmap(n);
    for (int depth = 0; depth < nlevels - 1; ++depth) {
        const int l = nlevels - depth - 1;
        const int sh = l - 1;
        for (int i = 0; i < n; ++i) {
            if (cmap[i].self.lvl != -1) continue;
            const int pi = parts[i] >> sh;
            if (pi & 1) continue;
            bool sep = false;
            for (int k = G.colptr[i]; k < G.colptr[i + 1] && !sep; ++k) {
                int j = G.rowind[k];
                if (cmap[j].self.lvl == -1 && (parts[j] >> sh) == pi + 1) sep = true;
            }
            if (sep) { cmap[i].self = SepID(l, pi / 2); cmap[i].section = parts[i]; }
        }
    }
    for (int i = 0; i < n; ++i)
        if (cmap[i].self.lvl == -1) {
            cmap[i].self = cmap[i].l = cmap[i].r = SepID(0, parts[i]);
            cmap[i].section = parts[i];
        }
    lap("separators");
    if (parts_out) *parts_out = parts;
    if (!Xc) {
        // interfaces from A^T A (src/partition.cpp:637-682): left starts at the vertex's own leaf
        // part, right at the first lower-level neighbour in the right half
        std::vector<ClusterID> out = cmap;
        for (int i = 0; i < n; ++i) {
            const SepID self = cmap[i].self;
            if (self.lvl <= 0) continue;
            SepID left(0, parts[i]), right(-1, 0);
            const int pi = parts[i] >> (self.lvl - 1);
            for (int k = G.colptr[i]; k < G.colptr[i + 1]; ++k) {
                const int j = G.rowind[k];
                const SepID nb = cmap[j].self;
                if (nb.lvl >= self.lvl) continue;
                if ((parts[j] >> (self.lvl - 1)) <= pi) left = highest_common(nb, left);
                else right = (right.lvl == -1) ? nb : highest_common(nb, right);
            }
            if (right.lvl == -1) throw std::runtime_error("a separator vertex has no neighbour on its right");
            out[i].l = left;
            out[i].r = right;
        }
        cmap = out;
    } else {
    // interfaces from the rows of A (src/partition.cpp:514-587), rmap == cmap (square)
    HostCSC At = csc_transpose(A);
    std::vector<ClusterID> rmap = cmap;
    for (int i = 0; i < n; ++i) {
        const SepID self = rmap[i].self;
        if (self.lvl <= 0) continue;
        SepID left(-1, 0), right(-1, 0);
        for (int k = At.colptr[i]; k < At.colptr[i + 1]; ++k) {
            const SepID nb = cmap[At.rowind[k]].self;
            if (nb.lvl < self.lvl) {
                int pj = nb.sep >> (self.lvl - nb.lvl - 1);
                if ((pj & 1) == 0) left = (left.lvl == -1) ? nb : lowest_common(nb, left);
                else right = (right.lvl == -1) ? nb : lowest_common(nb, right);
            }
        }
        rmap[i].l = left;
        rmap[i].r = right;
    }
    cmap = rmap;
    for (int i = 0; i < n; ++i) {
        const SepID self = rmap[i].self;
        SepID left = rmap[i].l, right = rmap[i].r;
        if (right == SepID()) {
            for (int k = At.colptr[i]; k < At.colptr[i + 1]; ++k) {
                const ClusterID& cj = cmap[At.rowind[k]];
                if (self == cj.self && cj.r != SepID()) right = (right == SepID()) ? cj.r : highest_common(right, cj.r);
            }
        } else if (left == SepID()) {
            for (int k = At.colptr[i]; k < At.colptr[i + 1]; ++k) {
                const ClusterID& cj = cmap[At.rowind[k]];
                if (self == cj.self && cj.l != SepID()) left = (left == SepID()) ? cj.l : highest_common(left, cj.l);
            }
        }
        rmap[i].l = left;
        rmap[i].r = right;
    }
    for (int i = 0; i < n; ++i) {
        if (rmap[i].l == SepID()) rmap[i].l = rmap[i].self;
        if (rmap[i].r == SepID()) rmap[i].r = rmap[i].self;
    }
    cmap = rmap;
    }
    lap("interfaces");
    // initPermutation (src/tree.cpp:48-57): stable sorts by the successively merged IDs.  The
    // first is a comparison sort; after it, a vertex's merged ID at every level is a function of
    // its level-0 ID (merge_if reads the ID only), so each later pass is a stable counting sort
    // by the rank of that value among the few distinct level-0 IDs' merged values: the same
    // permutation, O(n) per level.
    cperm_.resize(n);
    for (int i = 0; i < n; ++i) cperm_[i] = i;
    std::stable_sort(cperm_.begin(), cperm_.end(), [&](int a, int b) { return cmap[a] < cmap[b]; });
    std::vector<int> d0(n);
    std::vector<ClusterID> u;  // distinct level-0 IDs, increasing
    for (int k = 0; k < n; ++k) {
        const ClusterID& c = cmap[cperm_[k]];
        if (u.empty() || u.back() < c) u.push_back(c);
        d0[cperm_[k]] = (int)u.size() - 1;
    }
    const int nq = (int)u.size();
    std::vector<int> qord(nq), qrank(nq), cnt, nxt(n);
    for (int lvl = 1; lvl < nlevels; ++lvl) {
        for (auto& c : u) c = merge_if(c, lvl);
        for (int q = 0; q < nq; ++q) qord[q] = q;
        std::stable_sort(qord.begin(), qord.end(), [&](int a, int b) { return u[a] < u[b]; });
        int r = 0;
        for (int j = 0; j < nq; ++j) {
            if (j > 0 && u[qord[j - 1]] < u[qord[j]]) ++r;
            qrank[qord[j]] = r;
        }
        cnt.assign((size_t)r + 2, 0);
        for (int k = 0; k < n; ++k) cnt[qrank[d0[cperm_[k]]] + 1]++;
        for (int j = 0; j <= r; ++j) cnt[j + 1] += cnt[j];
        for (int k = 0; k < n; ++k) nxt[cnt[qrank[d0[cperm_[k]]]]++] = cperm_[k];
        cperm_.swap(nxt);
    }
    cpermed.resize(n);
    for (int k = 0; k < n; ++k) cpermed[k] = cmap[cperm_[k]];
    lap("permutation");
}

// The cluster hierarchy from the ordering (src/ptree.cpp:587-691): level-0 clusters are the runs
// of equal IDs, parents are grouped per level by merge_if.  Systems of a batch are kept apart:
// each owns a contiguous range of positions, and its clusters exist up to its own top level.
void Tree::make_hierarchy() {
    cl.clear();
    for (auto& v : bottoms) v.clear();
    for (auto& v : interiors) v.clear();
    for (auto& v : interfaces) v.clear();
    int rnext = 0;
    for (size_t s = 0; s < inst_n.size(); ++s) {
        const int k1 = inst_off[s] + inst_n[s];
        for (int k = inst_off[s]; k < k1;) {
            int kn = k + 1;
            while (kn < k1 && cpermed[kn] == cpermed[k]) kn++;
            Cluster c;
            c.order = (int)cl.size();
            c.merge_level = 0;
            c.inst = (int)s;
            c.id = cpermed[k];
            c.cstart = k;
            c.rstart = rnext;
            c.cols = c.cols_org = c.rank = kn - k;
            int nr = 0;
            for (int q = k; q < kn; ++q) nr += c2r_count[q];
            if (nr == 0) throw std::runtime_error("a leaf cluster got no rows");
            c.rows = c.rows_org = nr;
            c.row_rank = nr - c.cols;
            rnext += nr;
            cl.push_back(c);
            bottoms[0].push_back(c.order);
            (c.lvl() == 0 ? interiors[0] : interfaces[0]).push_back(c.order);
            k = kn;
        }
    }
    for (int lvl = 1; lvl < nlevels; ++lvl) {
        const auto prev = bottoms[lvl - 1];
        for (size_t b0 = 0; b0 < prev.size();) {
            const int s = cl[prev[b0]].inst;
            size_t e0 = b0;
            while (e0 < prev.size() && cl[prev[e0]].inst == s) e0++;
            if (lvl < inst_lv[s]) {
                size_t b = b0;
                while (b < e0 && cl[prev[b]].lvl() < lvl) b++;
                for (size_t k = b; k < e0;) {
                    const ClusterID pid = merge_if(cl[prev[k]].id, lvl);
                    std::vector<int> ch;
                    int sz = 0, rz = 0;
                    const int cst = cl[prev[k]].cstart, rst = cl[prev[k]].rstart;
                    while (k < e0 && merge_if(cl[prev[k]].id, lvl) == pid) {
                        ch.push_back(prev[k]);
                        sz += cl[prev[k]].cols;
                        rz += cl[prev[k]].rows;
                        k++;
                    }
                    Cluster p;
                    p.order = (int)cl.size();
                    p.merge_level = lvl;
                    p.inst = s;
                    p.id = pid;
                    p.cstart = cst;
                    p.rstart = rst;
                    p.cols = p.cols_org = p.rank = sz;
                    p.rows = p.rows_org = rz;
                    p.row_rank = rz - sz;
                    std::sort(ch.begin(), ch.end(), [this](int a, int b2) { return cl[a].id < cl[b2].id; });
                    p.children = ch;
                    for (int c : ch) cl[c].parent = p.order;
                    cl.push_back(p);
                    bottoms[lvl].push_back(p.order);
                    (pid.self.lvl == lvl ? interiors[lvl] : interfaces[lvl]).push_back(p.order);
                }
            }
            b0 = e0;
        }
    }
    cl_part = cl;
}

// form_rmap without a matching (src/tree.cpp:102-165; ParTree::partition :531-544): each row of
// A(:, cperm) is filed under the cluster whose columns carry the largest sum of squares of the
// row (the first in column order on a tie; a zero row goes to the last column), at that
// cluster's first column the row touches; rperm lists the rows by that column, then by row.
void Tree::form_rmap() {
    c2r_count.assign(n, 1);
    if (m == n) {
        rperm_ = cperm_;
        return;
    }
    std::vector<int> cinv(n);
    for (int k = 0; k < n; ++k) cinv[cperm_[k]] = k;
    const HostCSC At = csc_transpose(A0);  // rows of A, natural column indices
    std::vector<int> filed(m);
    std::vector<std::pair<int, double>> ent;
    c2r_count.assign(n, 0);
    for (int i = 0; i < m; ++i) {
        ent.clear();
        for (int k = At.colptr[i]; k < At.colptr[i + 1]; ++k) ent.emplace_back(cinv[At.rowind[k]], At.val[k]);
        std::sort(ent.begin(), ent.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        double best = -1.0;
        int cidx = -1;
        for (size_t q = 0; q < ent.size();) {
            const ClusterID& cid = cpermed[ent[q].first];
            double v = ent[q].second * ent[q].second;
            size_t qn = q + 1;
            while (qn < ent.size() && cpermed[ent[qn].first] == cid) {
                v += ent[qn].second * ent[qn].second;
                ++qn;
            }
            if (v > best) { best = v; cidx = ent[q].first; }
            q = qn;
        }
        if (cidx < 0) cidx = n - 1;
        filed[i] = cidx;
        c2r_count[cidx]++;
    }
    std::vector<int> start(n + 1, 0);
    for (int j = 0; j < n; ++j) start[j + 1] = start[j] + c2r_count[j];
    rperm_.assign(m, 0);
    for (int i = 0; i < m; ++i) rperm_[start[filed[i]]++] = i;
}

int Tree::partition(const HostCSC& A) {
    try {
        if (A.nr < A.nc) throw std::runtime_error("A has fewer rows than columns");
        // the reference's driver sets square from the shape (pspaQR.cpp:192); its non-square
        // path on a square A is not built here
        if ((A.nr == A.nc) != (square == 1))
            throw std::runtime_error(square ? "A is rectangular: set_square(0)" : "A is square: set_square(1)");
        const bool geo = !X.empty() && xdim > 0;
        if (A.nr != A.nc && geo)
            throw std::runtime_error("a rectangular system is ordered without coordinates (leaf parts or the graph "
                                     "bisection): the geometric path assigns rows by a bipartite matching");
        if (geo && (long long)X.size() != (long long)xdim * A.nc) throw std::runtime_error("coordinate array size mismatch");
        for (double v : X)
            if (!std::isfinite(v)) throw std::runtime_error("coordinates must be finite");
        if (!geo && !user_parts.empty()) check_parts(user_parts, 0, A.nc, nlevels);
        check_matrix(A);
        auto tp0 = std::chrono::steady_clock::now();
        A0 = A;
        n = A.nc;
        m = A.nr;
        inst_n.assign(1, n);
        inst_lv.assign(1, nlevels);
        inst_off.assign(1, 0);
        symdev.reset();
        if (device_symbolic && geo && m == n && nlevels > 1) {
            order_device(A);
            if (getenv("SPAQR_DEBUG_HOST"))
                fprintf(stderr, "[partition/device] total %.1f ms\n",
                        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp0).count());
        } else {
            order_system(A, geo ? X.data() : nullptr, xdim, (!geo && !user_parts.empty()) ? user_parts.data() : nullptr,
                         nlevels, cperm_, cpermed, &leaf_parts);
        }
        form_rmap();
        make_hierarchy();
        t_partition_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp0).count();
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

int Tree::partition_batch(const HostCSC& A, const std::vector<int>& sizes, const std::vector<int>& levels) {
    try {
        if (A.nr != A.nc) throw std::runtime_error("only square systems are supported");
        if (sizes.empty() || sizes.size() != levels.size()) throw std::runtime_error("one size and one level count per system");
        int total = 0, lmax = 0;
        for (size_t s = 0; s < sizes.size(); ++s) {
            if (sizes[s] < 1 || levels[s] < 1) throw std::runtime_error("system sizes and level counts must be positive");
            total += sizes[s];
            lmax = std::max(lmax, levels[s]);
        }
        if (total != A.nc) throw std::runtime_error("the system sizes must add up to the matrix size");
        check_matrix(A);
        if (lmax != nlevels) throw std::runtime_error("the tree's level count must equal the largest system's");
        const bool geo = !X.empty() && xdim > 0;
        n = A.nc;
        if (geo && (long long)X.size() != (long long)xdim * n) throw std::runtime_error("coordinate array size mismatch");
        const bool given = !geo && !user_parts.empty();
        if (given) {
            if ((int)user_parts.size() != n) throw std::runtime_error("one leaf part per column of the batch");
            for (int s = 0, off = 0; s < (int)sizes.size(); off += sizes[s], ++s) check_parts(user_parts, off, sizes[s], levels[s]);
        }
        inst_n = sizes;
        inst_lv = levels;
        inst_off.assign(sizes.size() + 1, 0);
        for (size_t s = 0; s < sizes.size(); ++s) inst_off[s + 1] = inst_off[s] + sizes[s];
        for (size_t s = 0; s < sizes.size(); ++s)  // no coupling between systems
            for (int j = inst_off[s]; j < inst_off[s + 1]; ++j)
                for (int k = A.colptr[j]; k < A.colptr[j + 1]; ++k)
                    if (A.rowind[k] < inst_off[s] || A.rowind[k] >= inst_off[s + 1])
                        throw std::runtime_error("the batch matrix must be block diagonal over the systems");
        A0 = A;
        cperm_.assign(n, 0);
        cpermed.assign(n, ClusterID());
        leaf_parts.assign(n, 0);
        parallel_ranges((int)sizes.size(), 1, [&](int b0, int e0, int) {
            for (int s = b0; s < e0; ++s) {
                const int off = inst_off[s], ns = sizes[s];
                HostCSC As;
                As.nr = As.nc = ns;
                As.colptr.assign(ns + 1, 0);
                for (int j = 0; j < ns; ++j) {
                    for (int k = A.colptr[off + j]; k < A.colptr[off + j + 1]; ++k) {
                        As.rowind.push_back(A.rowind[k] - off);
                        As.val.push_back(A.val[k]);
                    }
                    As.colptr[j + 1] = (int)As.rowind.size();
                }
                std::vector<int> cp, pt;
                std::vector<ClusterID> ids;
                order_system(As, geo ? X.data() + (size_t)off * xdim : nullptr, xdim,
                             given ? user_parts.data() + off : nullptr, levels[s], cp, ids, &pt);
                for (int k = 0; k < ns; ++k) {
                    cperm_[off + k] = off + cp[k];
                    cpermed[off + k] = ids[k];
                    leaf_parts[off + k] = pt[k];
                }
            }
        });
        inst_off.pop_back();
        m = n;
        symdev.reset();
        form_rmap();
        make_hierarchy();
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

// ---------------------------------------------------------------------------- assemble
int Tree::assemble() {
    try {
        if (cl_part.empty()) throw std::runtime_error("partition() first");
        ensure_stream();
        auto ta0 = std::chrono::steady_clock::now();
        if (symdev && !rect() && n_systems() == 1) {
            assemble_device();
            t_assemble_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta0).count();
            return 0;
        }
        cl = cl_part;  // a factorization since partition() has resized and flagged clusters
        App = permute_rc(A0, rperm_, cperm_);
        ed.clear();
        for (auto& c : cl) { c.out.clear(); c.in.clear(); c.sout.clear(); c.sin.clear(); c.self_edge = -1; c.dist2.clear(); }
        current_bottom = 0;
        std::vector<int> owner(n), rowner(m);  // leaf cluster of each permuted column / row
        for (int o : bottoms[0]) {
            for (int k = cl[o].cstart; k < cl[o].cstart + cl[o].cols; ++k) owner[k] = o;
            for (int k = cl[o].rstart; k < cl[o].rstart + cl[o].rows; ++k) rowner[k] = o;
        }
        // edges with dense blocks (src/ptree.cpp:728-775)
        h_blocks0.clear();
        h_block_off.clear();
        for (int s : bottoms[0]) {
            std::set<int> nb;
            const Cluster& cs = cl[s];
            for (int j = cs.cstart; j < cs.cstart + cs.cols; ++j)
                for (int k = App.colptr[j]; k < App.colptr[j + 1]; ++k) nb.insert(rowner[App.rowind[k]]);
            nb.insert(s);
            for (int o : nb) {
                int id = new_edge(s, o);
                Edge& e = ed[id];
                size_t off = h_blocks0.size();
                h_block_off.push_back(off);
                h_blocks0.resize(off + (size_t)e.rows * e.cols, 0.0);
                const int r0 = cl[o].rstart;
                for (int j = 0; j < e.cols; ++j) {
                    const int col = cs.cstart + j;
                    for (int k = App.colptr[col]; k < App.colptr[col + 1]; ++k) {
                        int r = App.rowind[k];
                        if (r >= r0 && r < r0 + e.rows) h_blocks0[off + (r - r0) + (size_t)j * e.rows] = App.val[k];
                    }
                }
            }
            sort_out(cl[s]);
        }
        // distance-two connections (src/ptree.cpp:832-854)
        HostCSC G = ata_pattern(App);
        std::vector<char> seen(bottoms[0].size(), 0);
        for (int s : bottoms[0]) {
            std::fill(seen.begin(), seen.end(), 0);
            std::set<int> d2;
            for (int col = cl[s].cstart; col < cl[s].cstart + cl[s].cols; ++col)
                for (int k = G.colptr[col]; k < G.colptr[col + 1]; ++k) {
                    int p = owner[G.rowind[k]];
                    if (seen[p]) continue;
                    if (cl[p].lvl() >= cl[s].lvl() && p != s) d2.insert(p);
                    else if (cl[p].lvl() < cl[s].lvl()) d2.insert(cl[p].dist2.begin(), cl[p].dist2.end());
                    seen[p] = 1;
                }
            cl[s].dist2.assign(d2.begin(), d2.end());
        }
        cl0 = cl;
        ed0 = ed;
        // device copies: pristine level-0 blocks in the factor precision, CSR of App for SpMV
        if (d_blocks0) { (void)hipFree(d_blocks0); d_blocks0 = nullptr; }
        const size_t nel = h_blocks0.size();
        const size_t es = (prec == 32) ? 4 : 8;
        HIPCHK(hipMalloc(&d_blocks0, std::max<size_t>(nel * es, 256)));
        prec_asm = prec;
        ++asm_gen;
        drop_graphs();
        if (prec == 32) {
            std::vector<float> f(h_blocks0.begin(), h_blocks0.end());
            HIPCHK(hipMemcpy(d_blocks0, f.data(), nel * 4, hipMemcpyHostToDevice));
        } else {
            HIPCHK(hipMemcpy(d_blocks0, h_blocks0.data(), nel * 8, hipMemcpyHostToDevice));
        }
        HostCSC R = csc_transpose(App);  // CSC of App^T == CSR of App
        if (d_rowptr) (void)hipFree(d_rowptr);
        if (d_col) (void)hipFree(d_col);
        if (d_val) (void)hipFree(d_val);
        HIPCHK(hipMalloc(&d_rowptr, sizeof(int) * (m + 1)));
        HIPCHK(hipMalloc(&d_col, sizeof(int) * std::max<size_t>(R.rowind.size(), 1)));
        HIPCHK(hipMalloc(&d_val, sizeof(double) * std::max<size_t>(R.val.size(), 1)));
        HIPCHK(hipMemcpy(d_rowptr, R.colptr.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_col, R.rowind.data(), sizeof(int) * R.rowind.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_val, R.val.data(), sizeof(double) * R.val.size(), hipMemcpyHostToDevice));
        d_nnz = (long long)R.val.size();
        // the adjoint's CSR is App's CSC
        if (d_tptr) (void)hipFree(d_tptr);
        if (d_tcol) (void)hipFree(d_tcol);
        if (d_tval) (void)hipFree(d_tval);
        d_tptr = nullptr;
        d_tcol = nullptr;
        d_tval = nullptr;
        if (rect()) {
            HIPCHK(hipMalloc(&d_tptr, sizeof(int) * (n + 1)));
            HIPCHK(hipMalloc(&d_tcol, sizeof(int) * std::max<size_t>(App.rowind.size(), 1)));
            HIPCHK(hipMalloc(&d_tval, sizeof(double) * std::max<size_t>(App.val.size(), 1)));
            HIPCHK(hipMemcpy(d_tptr, App.colptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_tcol, App.rowind.data(), sizeof(int) * App.rowind.size(), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_tval, App.val.data(), sizeof(double) * App.val.size(), hipMemcpyHostToDevice));
        }
        if (!d_scratch) HIPCHK(hipMalloc(&d_scratch, sizeof(double) * (dot_blocks + 4096)));
        n_blocks0 = h_blocks0.size();
        factorized = false;
        t_assemble_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta0).count();
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

// The geometric ordering on the device (SymbolicDevice::order); the ranks of the distinct
// level-0 IDs follow initPermutation (src/tree.cpp:48-57): its successive stable sorts by the
// IDs merged up to each level order the vertices by the top level's merged ID, then the next
// level's, ..., then the level-0 ID, then the natural index.
void Tree::order_device(const HostCSC& A) {
    ensure_stream();
    symdev.reset(new SymbolicDevice());
    auto rank_ids = [this](const std::vector<DevId>& ids) {
        const int nd = (int)ids.size();
        std::vector<std::vector<ClusterID>> chain(nd, std::vector<ClusterID>(nlevels));
        for (int q = 0; q < nd; ++q) {
            ClusterID c;
            c.self = SepID(ids[q].slvl, ids[q].ssep);
            c.l = SepID(ids[q].llvl, ids[q].lsep);
            c.r = SepID(ids[q].rlvl, ids[q].rsep);
            c.section = ids[q].section;
            chain[q][0] = c;
            for (int lvl = 1; lvl < nlevels; ++lvl) chain[q][lvl] = merge_if(chain[q][lvl - 1], lvl);
        }
        std::vector<int> ord(nd);
        for (int q = 0; q < nd; ++q) ord[q] = q;
        std::sort(ord.begin(), ord.end(), [&](int a, int b) {
            for (int lvl = nlevels - 1; lvl >= 0; --lvl)
                if (chain[a][lvl] != chain[b][lvl]) return chain[a][lvl] < chain[b][lvl];
            return a < b;
        });
        std::vector<int> rank(nd);
        for (int q = 0; q < nd; ++q) rank[ord[q]] = q;
        return rank;
    };
    std::vector<int> did;
    std::vector<DevId> ids;
    symdev->order(A.nc, A.colptr.data(), A.rowind.data(), A.val.data(), X.data(), xdim, nlevels, rank_ids, cperm_, did,
                  ids, leaf_parts, stream_);
    cpermed.resize(n);
    for (int k = 0; k < n; ++k) {
        const DevId& d = ids[did[cperm_[k]]];
        ClusterID c;
        c.self = SepID(d.slvl, d.ssep);
        c.l = SepID(d.llvl, d.lsep);
        c.r = SepID(d.rlvl, d.rsep);
        c.section = d.section;
        cpermed[k] = c;
    }
}

// ParTree::assemble (src/ptree.cpp:696-864) from the device ordering's matrix: the edges
// (every level-0 cluster pair joined by a nonzero, in the order the host phase creates them),
// the blocks scattered straight into the factor precision, App's CSR for SpMV, and the
// distance-two connections from A^T A's cluster pairs by the same rule as the host phase.
void Tree::assemble_device() {
    static const bool dbg = getenv("SPAQR_DEBUG_HOST") != nullptr;
    auto tp = std::chrono::steady_clock::now();
    auto lap = [&](const char* what) {
        if (!dbg) return;
        sync_stream();
        const auto nw = std::chrono::steady_clock::now();
        fprintf(stderr, "[assemble/device] %s %.1f ms\n", what, std::chrono::duration<double, std::milli>(nw - tp).count());
        tp = nw;
    };
    cl = cl_part;
    ed.clear();
    for (auto& c : cl) { c.out.clear(); c.in.clear(); c.sout.clear(); c.sin.clear(); c.self_edge = -1; c.dist2.clear(); }
    current_bottom = 0;
    const int n0 = (int)bottoms[0].size();
    std::vector<int> owner(n), cstart(n0), cols(n0);
    for (int q = 0; q < n0; ++q) {
        const int o = bottoms[0][q];
        if (o != q) throw std::runtime_error("level-0 clusters out of order");
        cstart[q] = cl[o].cstart;
        cols[q] = cl[o].cols;
        for (int k = cl[o].cstart; k < cl[o].cstart + cl[o].cols; ++k) owner[k] = o;
    }
    std::vector<long long> keys;
    lap("tables");
    symdev->edges(cperm_, owner, n0, keys, stream_);
    lap("edge keys");
    h_blocks0.clear();
    h_block_off.clear();
    std::vector<long long> off;
    off.reserve(keys.size());
    size_t nel = 0;
    for (size_t q = 0; q < keys.size(); ++q) {
        const int s = (int)(keys[q] / n0), o = (int)(keys[q] % n0);
        const int id = new_edge(s, o);
        h_block_off.push_back(nel);
        off.push_back((long long)nel);
        nel += (size_t)ed[id].rows * ed[id].cols;
        if (q + 1 == keys.size() || keys[q + 1] / n0 != s) sort_out(cl[s]);
    }
    n_blocks0 = nel;
    lap("edges (host)");
    if (d_blocks0) { (void)hipFree(d_blocks0); d_blocks0 = nullptr; }
    const size_t es = (prec == 32) ? 4 : 8;
    HIPCHK(hipMalloc(&d_blocks0, std::max<size_t>(nel * es, 256)));
    prec_asm = prec;
    ++asm_gen;
    drop_graphs();
    if (prec == 32) symdev->scatter_blocks<float>(off, cstart, cols, (float*)d_blocks0, nel, stream_);
    else symdev->scatter_blocks<double>(off, cstart, cols, (double*)d_blocks0, nel, stream_);
    lap("blocks");
    // distance-two connections (src/ptree.cpp:832-854)
    std::vector<long long> pairs;
    symdev->ata_pairs(pairs, stream_);
    lap("A^T A pairs");
    // The host phase visits the clusters in order and a lower-level neighbour p contributes its
    // own list, which is final when p comes before s and still empty when it comes after: so
    // s depends only on earlier clusters of lower levels, and the clusters of one level (in
    // increasing level) are independent of each other.
    {
        std::vector<size_t> first(n0 + 1, pairs.size());
        for (size_t q = pairs.size(); q-- > 0;) first[pairs[q] / n0] = q;
        for (int s = n0 - 1; s >= 0; --s) first[s] = std::min(first[s], first[s + 1]);
        int lmax = 0;
        for (int s = 0; s < n0; ++s) lmax = std::max(lmax, cl[s].lvl());
        std::vector<std::vector<int>> by_lvl(lmax + 1);
        for (int s = 0; s < n0; ++s) by_lvl[cl[s].lvl()].push_back(s);
        for (const auto& group : by_lvl)
            parallel_ranges((int)group.size(), 64, [&](int b, int e, int) {
                std::vector<int> d2;
                for (int g = b; g < e; ++g) {
                    const int s = group[g];
                    d2.clear();
                    for (size_t q = first[s]; q < first[s + 1]; ++q) {
                        const int p = (int)(pairs[q] % n0);
                        if (cl[p].lvl() >= cl[s].lvl() && p != s) d2.push_back(p);
                        else if (cl[p].lvl() < cl[s].lvl() && p < s)
                            d2.insert(d2.end(), cl[p].dist2.begin(), cl[p].dist2.end());
                    }
                    std::sort(d2.begin(), d2.end());
                    d2.erase(std::unique(d2.begin(), d2.end()), d2.end());
                    cl[s].dist2 = d2;
                }
            });
    }
    lap("distance two (host)");
    cl0 = cl;
    ed0 = ed;
    lap("pristine copies");
    if (d_rowptr) (void)hipFree(d_rowptr);
    if (d_col) (void)hipFree(d_col);
    if (d_val) (void)hipFree(d_val);
    d_rowptr = nullptr;
    d_col = nullptr;
    d_val = nullptr;
    symdev->csr(&d_rowptr, &d_col, &d_val, stream_);
    lap("csr");
    d_nnz = (long long)A0.colptr[n];
    if (d_tptr) (void)hipFree(d_tptr);
    if (d_tcol) (void)hipFree(d_tcol);
    if (d_tval) (void)hipFree(d_tval);
    d_tptr = nullptr;
    d_tcol = nullptr;
    d_tval = nullptr;
    if (!d_scratch) HIPCHK(hipMalloc(&d_scratch, sizeof(double) * (dot_blocks + 4096)));
    factorized = false;
}

// ---------------------------------------------------------------------------- fill-in
// get_sparsity + alloc_fillin (src/ptree.cpp:866-871, 140-153); returns nothing, new
// edges are appended to ed (their device blocks are allocated by the caller).
// Fill-in, first half: each interior's distance-2 list gains the sources of its in-edges.
// Returns whether the level's interiors are independent: no interior has another interior of
// this level in its distance-2 list.  Then fillin_edges changes no interior's out list (it may
// add in-edges to interiors, blocks that are zero until the trailing update), so the
// elimination's panel work can be queued before the new edges are made.
bool Tree::fillin_dist2() {
    const auto& ints = interiors[current_bottom];
    parallel_ranges((int)ints.size(), 64, [&](int b, int e, int) {
        for (int i = b; i < e; ++i) {
            Cluster& C = cl[ints[i]];
            std::vector<int>& d2 = C.dist2;
            for (int e2 : C.in) d2.push_back(ed[e2].n1);
            std::sort(d2.begin(), d2.end());
            d2.erase(std::unique(d2.begin(), d2.end()), d2.end());
        }
    });
    std::vector<char> is_int(cl.size(), 0);
    for (int c : ints) is_int[c] = 1;
    for (int c : ints)
        for (int nn : cl[c].dist2)
            if (nn != c && is_int[nn]) return false;
    return true;
}

// Fill-in, second half: the new blocks.  Independent interiors: the new blocks are fixed by the
// edges as they stand: candidates (nn, nbr) built per interior in parallel, the first occurrence
// of each kept (in interior, neighbour, out-edge order), then the edges made in that order.
// Otherwise the interiors run one after another.
void Tree::fillin_edges(bool independent) {
    static const bool dbg = getenv("SPAQR_DEBUG_FILLIN") != nullptr;
    auto tf = std::chrono::steady_clock::now();
    double tm[6] = {0};
    int ti = 0;
    auto tick = [&] {
        const auto n = std::chrono::steady_clock::now();
        tm[ti++] = std::chrono::duration<double, std::milli>(n - tf).count();
        tf = n;
    };
    const auto& ints = interiors[current_bottom];
    const int nint = (int)ints.size();
    tick();
    if (!independent) {
        // edge existence through a stamp array: mark[n2] == stamp when nn already has an edge to n2
        std::vector<int> mark(cl.size(), -1);
        int stamp = 0;
        for (int c : ints) {
            const std::vector<int> outs = cl[c].out;
            for (int nn : cl[c].dist2) {
                ++stamp;
                for (int e : cl[nn].out) mark[ed[e].n2] = stamp;
                for (int e : outs) {
                    const int nbr = ed[e].n2;
                    if (mark[nbr] != stamp) {
                        new_edge(nn, nbr);
                        mark[nbr] = stamp;
                    }
                }
            }
        }
        return;
    }
    using Cand = std::pair<int, int>;  // (nn, nbr)
    const int T = host_threads();
    // per builder slot: its candidates in order, and their indices filed by nn % T (the
    // thread that keeps the first occurrence of every pair of that nn)
    std::vector<std::vector<Cand>> cand(T);
    std::vector<std::vector<std::vector<int>>> byt(T, std::vector<std::vector<int>>(T));
    const int ns = parallel_ranges(nint, 64, [&](int b, int e, int slot) {
        std::vector<int> mark(cl.size(), -1);
        int stamp = 0;
        auto& out = cand[slot];
        auto& bt = byt[slot];
        for (int i = b; i < e; ++i) {
            const Cluster& C = cl[ints[i]];
            for (int nn : C.dist2) {
                ++stamp;
                for (int e2 : cl[nn].out) mark[ed[e2].n2] = stamp;
                for (int e2 : C.out) {
                    const int nbr = ed[e2].n2;
                    if (mark[nbr] != stamp) {
                        bt[nn % T].push_back((int)out.size());
                        out.push_back({nn, nbr});
                    }
                }
            }
        }
    });
    cand.resize(ns);
    tick();
    std::vector<std::vector<char>> keep(ns);
    size_t ncand = 0;
    for (int s2 = 0; s2 < ns; ++s2) {
        keep[s2].assign(cand[s2].size(), 0);
        ncand += cand[s2].size();
    }
    parallel_ranges(T, ncand < 16384 ? T : 1, [&](int b, int e, int) {
        for (int t = b; t < e; ++t) {
            size_t nt = 0;
            for (int s2 = 0; s2 < ns; ++s2) nt += byt[s2][t].size();
            FlatKeySet seen(nt);
            for (int s2 = 0; s2 < ns; ++s2)
                for (int j : byt[s2][t]) {
                    const Cand& c = cand[s2][j];
                    if (seen.insert(((unsigned long long)c.first << 32) | (unsigned)c.second)) keep[s2][j] = 1;
                }
        }
    });
    tick();
    // the edges in candidate order, as new_edge makes them: entries and out lists by source,
    // in lists by target (each list appended in id order), sources and targets split over
    // the threads by cluster number
    const size_t base = ed.size();
    std::vector<size_t> koff(ns + 1, 0);
    for (int s2 = 0; s2 < ns; ++s2)
        koff[s2 + 1] = koff[s2] + (size_t)std::count(keep[s2].begin(), keep[s2].end(), (char)1);
    const size_t nnew = koff[ns];
    std::vector<Cand> made(nnew);
    std::vector<std::vector<std::vector<int>>> obs(ns, std::vector<std::vector<int>>(T)), ibs = obs;
    parallel_ranges(ns, 1, [&](int b, int e, int) {
        for (int s2 = b; s2 < e; ++s2) {
            size_t i = koff[s2];
            for (size_t j = 0; j < cand[s2].size(); ++j)
                if (keep[s2][j]) {
                    const Cand c = cand[s2][j];
                    obs[s2][c.first % T].push_back((int)i);
                    if (c.first != c.second) ibs[s2][c.second % T].push_back((int)i);
                    made[i++] = c;
                }
        }
    });
    ed.resize(base + nnew);
    tick();
    parallel_ranges(T, nnew < 16384 ? T : 1, [&](int b, int e, int) {
        for (int t = b; t < e; ++t)
            for (int s2 = 0; s2 < ns; ++s2)
                for (int i : obs[s2][t]) {
                    const int c1 = made[i].first, c2 = made[i].second, id = (int)(base + i);
                    Edge& E = ed[id];
                    E.n1 = c1;
                    E.n2 = c2;
                    E.rows = cl[c2].rows;
                    E.cols = cl[c1].cols;
                    E.spars = cl[c1].lvl() > current_bottom && cl[c2].lvl() > current_bottom && c1 != c2;
                    Cluster& C1 = cl[c1];
                    C1.out.push_back(id);
                    if (E.spars) C1.sout.push_back(id);
                    if (c1 == c2) C1.self_edge = id;
                }
    });
    parallel_ranges(T, nnew < 16384 ? T : 1, [&](int b, int e, int) {
        for (int t = b; t < e; ++t)
            for (int s2 = 0; s2 < ns; ++s2)
                for (int i : ibs[s2][t]) {
                    const int id = (int)(base + i);
                    Cluster& C2 = cl[made[i].second];
                    C2.in.push_back(id);
                    if (ed[id].spars) C2.sin.push_back(id);
                }
    });
    tick();
    if (dbg)
        fprintf(stderr, "[fillin L%d] %zu candidates, %zu edges: setup %.2f cand %.2f dedup %.2f made %.2f lists %.2f ms\n",
                current_bottom, ncand, nnew, tm[0], tm[1], tm[2], tm[3], tm[4]);
}

double Tree::factor_flops() const {
    double f = 0;
    for (auto& s : stats) f += s.flops;
    return f;
}

// ---------------------------------------------------------------------------- factorize
template <typename S>
int Tree::factorize_t() {
    auto t0 = std::chrono::steady_clock::now();
    // the pristine structure; element-wise on the host threads when the cluster count is unchanged
    // (every cluster's lists keep their buffers)
    if (cl.size() == cl0.size())
        parallel_ranges((int)cl.size(), 512, [&](int b, int e, int) {
            for (int i = b; i < e; ++i) cl[i] = cl0[i];
        });
    else
        cl = cl0;
    ed.assign(ed0.begin(), ed0.end());
    current_bottom = 0;
    sync_stream();  // earlier work on the retained chunks has finished before they are reused
    arena.reset();
    desc_arena.reset();
    stats.clear();
    factorized = false;
    solve_sched_begin<S>();
    seg_begin();
    // level-0 blocks: copy of the pristine assembled blocks
    {
        S* base = (S*)arena.alloc(std::max<size_t>(n_blocks0, 1) * sizeof(S));
        HIPCHK(hipMemcpyAsync(base, d_blocks0, n_blocks0 * sizeof(S), hipMemcpyDeviceToDevice, stream_));
        for (size_t i = 0; i < ed.size(); ++i) ed[i].A = base + h_block_off[i];
    }
    d_chain_err = (int*)arena.alloc(sizeof(int));
    launch_fill<int>(d_chain_err, 1, 0, stream_);
    // status: per cluster, and per level (flags, first cluster)
    d_cinfo = (int*)arena.alloc(sizeof(int) * std::max<size_t>(cl.size(), 1));
    d_lflags = (int*)arena.alloc(sizeof(int) * nlevels);
    d_lfirst = (int*)arena.alloc(sizeof(int) * nlevels);
    launch_fill<int>(d_cinfo, cl.size(), 0, stream_);
    launch_fill<int>(d_lflags, (size_t)nlevels, 0, stream_);
    launch_fill<int>(d_lfirst, (size_t)nlevels, INT_MAX, stream_);
    h_info.assign(cl.size(), 0);
    // host-phase timers (SPAQR_DEBUG_HOST=1 prints them per factorization)
    static const bool dbg_host = getenv("SPAQR_DEBUG_HOST") != nullptr;
    double hms[24] = {0};
    hms[15] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();  // tables
    auto hclock = [] { return std::chrono::steady_clock::now(); };
    auto hacc = [&](int i, std::chrono::steady_clock::time_point& t) {
        auto n = std::chrono::steady_clock::now();
        hms[i] += std::chrono::duration<double, std::milli>(n - t).count();
        t = n;
    };
    auto zalloc = [&](size_t nel) -> S* {
        S* p = (S*)arena.alloc(std::max<size_t>(nel, 1) * sizeof(S));
        launch_fill<S>(p, nel, S(0), stream_);
        return p;
    };
    auto E = [&](int id) -> Edge& { return ed[id]; };
    const double SZ = (double)sizeof(S);
    auto Ap = [&](int id) -> S* { return (S*)ed[id].A; };
    // Blocks of at most 64 x 64 go to the one-wave kernels (SPAQR_SMALL_TASKS=0 keeps every
    // task on the workgroup kernels); the rest stay on the workgroup kernels.
    static const bool small_env = getenv("SPAQR_SMALL_TASKS") == nullptr || atoi(getenv("SPAQR_SMALL_TASKS")) != 0;
    const bool small_tasks = small_env && tuned(T_SMALL_TASKS, 1) == 1;
    const int big_front = tuned(T_BIG_FRONT, kBigFront);
    const int larfb_cols = tuned(T_LARFB_COLS, 128);
    const int trsm_rows = tuned(T_TRSM_ROWS, sizeof(S) == 4 ? 128 : 256);
    // SPAQR_DEBUG_SHAPES=1: flop-weighted shapes of the workgroup-kernel apply tasks, printed at the end
    static const bool dbg_shapes = getenv("SPAQR_DEBUG_SHAPES") != nullptr;
    std::map<std::string, std::array<double, 2>> shapes;  // key -> {tasks, flops}
    auto bucket = [](int v) { int b = 16; while (b < v) b *= 2; return b; };
    auto shape_add = [&](const char* cls, int k, int m, int n, double fl) {
        char key[96];
        snprintf(key, sizeof key, "%-7s k<=%-5d m<=%-5d n<=%-4d", cls, bucket(k), bucket(m), bucket(n));
        auto& e = shapes[key];
        e[0] += 1;
        e[1] += fl;
    };
    struct LarfbBatch {
        LarfbTask<S>* d[4] = {nullptr, nullptr, nullptr, nullptr};  // buckets 16 / 32 / 64, workgroup
        int n[4] = {0, 0, 0, 0};
        LarfbTask<S>* dr[4] = {nullptr, nullptr, nullptr, nullptr};  // C-resident FP32 (larfb: m, trsm: k2 buckets)
        int nr[4] = {0, 0, 0, 0};
        bool trsm = false;
        int launches() const {
            return (n[0] > 0) + (n[1] > 0) + (n[2] > 0) + (n[3] > 0) + (nr[0] > 0) + (nr[1] > 0) + (nr[2] > 0) + (nr[3] > 0);
        }
    };
    // larfb chunks go straight into launch buckets: 0-2 one-wave by size, 3-6 C-resident, 7
    // workgroup; Q^T C and C R^{-1} chunks into separate bucket sets (two launches)
    using LarfbBuckets = std::array<std::vector<LarfbTask<S>>, 8>;
    auto larfb_bucket = [&](const LarfbTask<S>& t, bool trsm) -> int {
        if (small_tasks) {
            const int d = trsm ? ((t.m <= 64 && t.k2 <= 64) ? t.k2 : -1) : ((t.m <= 64 && t.n <= 64) ? t.m : -1);
            if (d >= 0) return d <= 16 ? 0 : (d <= 32 ? 1 : 2);
        }
        const int q = trsm ? trsm_res_bucket(t) : larfb_res_bucket(t);
        return q < 0 ? 7 : 3 + q;
    };
    // the chunks of one task, as split_larfb cuts them, filed by bucket
    auto file_larfb = [&](const LarfbTask<S>& t, LarfbBuckets& cols, LarfbBuckets& tr) {
        if (t.V) {
            const int cw = (sizeof(S) == 4 && t.m > 512) ? std::min(64, larfb_cols) : larfb_cols;
            for (int c0 = 0; c0 < t.n; c0 += cw) {
                LarfbTask<S> u = t;
                u.C = t.C + (size_t)c0 * t.ldc;
                u.n = std::min(cw, t.n - c0);
                u.R = nullptr;
                u.k2 = 0;
                cols[larfb_bucket(u, false)].push_back(u);
            }
        }
        if (t.R) {
            const int rc = trsm_rows;
            for (int r0 = 0; r0 < t.m; r0 += rc) {
                LarfbTask<S> u = t;
                u.V = nullptr;
                u.T = nullptr;
                u.C = t.C + r0;
                u.m = std::min(rc, t.m - r0);
                tr[larfb_bucket(u, true)].push_back(u);
            }
        }
    };
    auto prep_larfb_parts = [&](const std::vector<LarfbBuckets>& parts, bool trsm) {
        const auto u = upload_buckets(parts);
        LarfbBatch b;
        b.trsm = trsm;
        for (int q = 0; q < 3; ++q) { b.d[q] = u[q].first; b.n[q] = u[q].second; }
        for (int q = 0; q < 4; ++q) { b.dr[q] = u[3 + q].first; b.nr[q] = u[3 + q].second; }
        b.d[3] = u[7].first;
        b.n[3] = u[7].second;
        if (dbg_shapes)
            for (const auto& pb : parts)
                for (int q = 3; q < 8; ++q)
                    for (const auto& t : pb[q])
                        shape_add(trsm ? "larfb+t" : "larfb", t.V ? t.k : t.k2, t.m, t.n, 4.0 * t.m * t.n * t.k);
        return b;
    };
    struct GeqrtBatch {
        GeqrtTask<S>* d[4] = {nullptr, nullptr, nullptr, nullptr};  // buckets 16 / 32 / 64, workgroup
        int n[4] = {0, 0, 0, 0};
        int launches() const { return (n[0] > 0) + (n[1] > 0) + (n[2] > 0) + (n[3] > 0); }
    };
    auto prep_geqrt = [&](const std::vector<GeqrtTask<S>>& v) {
        GeqrtBatch b;
        std::vector<GeqrtTask<S>> sm[3], big;
        if (small_tasks)
            split_small(v, [&](const GeqrtTask<S>& t) { return (t.m <= 64 && t.n <= 64) ? t.m : -1; }, sm, big);
        else
            big = v;
        for (int q = 0; q < 3; ++q) { b.d[q] = upload_vec(sm[q]); b.n[q] = (int)sm[q].size(); }
        b.d[3] = upload_vec(big);
        b.n[3] = (int)big.size();
        return b;
    };
    auto run_geqrt = [&](const GeqrtBatch& b) {
        for (int q = 0; q < 3; ++q) launch_geqrt_wave<S>(b.d[q], b.n[q], 16 << q, stream_);
        launch_geqrt<S>(b.d[3], b.n[3], stream_);
    };
    auto run_larfb = [&](const LarfbBatch& b) {
        for (int q = 0; q < 3; ++q) launch_larfb_small<S>(b.d[q], b.n[q], 16 << q, b.trsm, stream_);
        if constexpr (sizeof(S) == 4)
            for (int q = 0; q < 3; ++q) {
                if (b.trsm) launch_trsm_res(b.dr[q], b.nr[q], 4 << q, stream_);
                else launch_larfb_res(b.dr[q], b.nr[q], 4 << q, stream_);
            }
        if constexpr (sizeof(S) == 4) launch_larfb_res2(b.dr[3], b.nr[3], stream_);
        launch_larfb<S>(b.d[3], b.n[3], stream_);
    };
    struct SparsBatch {
        SparsApplyTask<S>* d[4] = {nullptr, nullptr, nullptr, nullptr};
        int n[4] = {0, 0, 0, 0};
        SparsApplyTask<S>* dr[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // C-resident FP32
        int nr[6] = {0, 0, 0, 0, 0, 0};
        int launches() const {
            int c = 0;
            for (int q = 0; q < 4; ++q) c += n[q] > 0;
            for (int q = 0; q < 6; ++q) c += nr[q] > 0;
            return c;
        }
    };
    // Sparsify tasks go straight from the builders' per-thread lists into their launch buckets
    // (0-2 one-wave by size, 3-8 C-resident, 9 workgroup) and from there, each bucket once,
    // into pinned staging and one upload.
    auto spars_bucket = [&](const SparsApplyTask<S>& t) -> int {
        if (rect()) return 9;  // the generic kernel for the rectangular shapes (kc < mc, row offsets)
        if (small_tasks) {
            int d = -1;
            if (t.ident) d = t.rc <= 64 ? 16 : -1;
            else if (t.pad == 1) d = (t.mc <= 64 && t.kn <= 64) ? t.kn : -1;
            else d = (t.rn <= 64 && t.kc <= 64) ? t.kc : -1;
            if (d >= 0) return d <= 16 ? 0 : (d <= 32 ? 1 : 2);
        }
        const int q = spars_res_bucket(t);
        return q < 0 ? 9 : 3 + q;
    };
    // parts[i][q]: builder slot i's tasks of bucket q; one threaded pass copies every slot's
    // buckets into place (one thread start per batch), then each bucket is one upload
    using SparsBuckets = std::array<std::vector<SparsApplyTask<S>>, 10>;
    auto prep_spars_parts = [&](const std::vector<SparsBuckets>& parts) {
        const auto u = upload_buckets(parts);
        SparsBatch b;
        for (int q = 0; q < 3; ++q) { b.d[q] = u[q].first; b.n[q] = u[q].second; }
        for (int q = 0; q < 6; ++q) { b.dr[q] = u[3 + q].first; b.nr[q] = u[3 + q].second; }
        b.d[3] = u[9].first;
        b.n[3] = u[9].second;
        return b;
    };
    auto run_spars = [&](const SparsBatch& b) {
        for (int q = 0; q < 3; ++q) launch_spars_small<S>(b.d[q], b.n[q], 16 << q, stream_);
        if constexpr (sizeof(S) == 4)
            for (int q = 0; q < 6; ++q) launch_spars_res(b.dr[q], b.nr[q], 4 << (q % 3), q < 3, stream_);
        launch_spars_apply<S>(b.d[3], b.n[3], stream_);
    };

    for (int L = 0; L < nlevels; ++L) {
        auto ht = hclock();
        current_bottom = L;
        LevelStats st;
        st.level = L;
        // fill-in blocks (src/ptree.cpp:885-1010), zeroed on the stream.  With independent
        // interiors they are made after the elimination's panel work is queued, while it runs.
        auto fill_in = [&](bool independent) {
            const size_t ne_before = ed.size();
            fillin_edges(independent);
            size_t tot = 0;
            for (size_t i = ne_before; i < ed.size(); ++i) tot += (size_t)ed[i].rows * ed[i].cols;
            S* blk = zalloc(tot);
            size_t off = 0;
            for (size_t i = ne_before; i < ed.size(); ++i) {
                ed[i].A = blk + off;
                off += (size_t)ed[i].rows * ed[i].cols;
            }
        };
        const bool independent = fillin_dist2();
        if (!independent) fill_in(false);
        hacc(0, ht);
        const auto& ints = interiors[L];
        const auto& itf = interfaces[L];
        st.n_interiors = (int)ints.size();
        st.n_interfaces = (int)itf.size();

        // ---- elimination descriptors: T blocks allocated in interior order first (a bump
        // allocator), then the tasks built over the interiors in parallel into per-slot lists
        // (links at per-interior offsets, larfb chunks straight into launch buckets),
        // concatenated in slot order: the sequential order.
        std::vector<int> lnk0(ints.size() + 1, 0);
        for (size_t i = 0; i < ints.size(); ++i) {
            const int c = ints[i];
            Cluster& C = cl[c];
            if (C.cols == 0) throw std::runtime_error("cluster has zero cols");
            if (C.rows < C.cols) throw std::runtime_error("an interior has fewer rows than columns");
            if (C.out.empty() || ed[C.out[0]].n2 != c) throw std::runtime_error("first out-edge of an interior must be itself");
            st.max_front = std::max(st.max_front, C.cols);
            C.T = arena.alloc((size_t)NB * C.cols * sizeof(S));
            int nl = 0;
            for (size_t j = 1; j < C.out.size(); ++j) {
                const int ns = link_slices<S>(ed[C.out[j]].rows);
                ed[C.out[j]].T = arena.alloc((size_t)NB * C.cols * sizeof(S) * ns);
                nl += ns;
            }
            lnk0[i + 1] = lnk0[i] + nl;
        }
        struct ElimPart {
            std::vector<GeqrtTask<S>> geq;
            std::vector<TpqrtTask<S>> tpq;
            LarfbBuckets lc, lt;
            long nlfb = 0;
            double flops = 0.0;
            double pf[3] = {0.0, 0.0, 0.0}, pb[3] = {0.0, 0.0, 0.0};  // geqrt, larfb, tpqrt
        };
        std::vector<ElimPart> eparts(host_threads());
        std::vector<TpqrtLink<S>> tpql((size_t)lnk0[ints.size()]);
        const int nep = parallel_ranges((int)ints.size(), 64, [&](int b, int e, int slot) {
            ElimPart& P = eparts[slot];
            for (int i = b; i < e; ++i) {
                const Cluster& C = cl[ints[i]];
                const Edge& es = ed[C.self_edge];
                P.geq.push_back({Ap(C.self_edge), (S*)C.T, es.rows, es.cols, es.rows, 0});
                const double fg = flops_geqrt<S>(es.rows, es.cols);
                P.flops += fg;
                P.pf[0] += fg;
                P.pb[0] += SZ * (2.0 * es.rows * es.cols + NB * es.cols);
                for (int e2 : C.in) {
                    const Edge& ei = ed[e2];
                    file_larfb({Ap(C.self_edge), (S*)C.T, Ap(e2), nullptr, ei.rows, ei.cols, C.cols, es.rows, ei.rows, 0, 0, 0},
                               P.lc, P.lt);
                    ++P.nlfb;
                    const double f = 4.0 * ei.rows * ei.cols * C.cols - 2.0 * ei.cols * C.cols * C.cols;
                    P.flops += f;
                    P.pf[1] += f;
                    P.pb[1] += SZ * ((double)ei.rows * C.cols + NB * C.cols + 2.0 * ei.rows * ei.cols);
                }
                TpqrtTask<S> tk{Ap(C.self_edge), C.cols, es.rows, lnk0[i], lnk0[i + 1] - lnk0[i]};
                int li = lnk0[i];
                for (size_t j = 1; j < C.out.size(); ++j) {
                    const Edge& eo = ed[C.out[j]];
                    const int ns = link_slices<S>(eo.rows), sz = (eo.rows + ns - 1) / ns;
                    for (int sl = 0; sl < ns; ++sl)  // a tall block: one link per row slice, T blocks side by side
                        tpql[li++] = {Ap(C.out[j]) + (size_t)sl * sz, (S*)eo.T + (size_t)sl * NB * C.cols,
                                      std::min(sz, eo.rows - sl * sz), eo.rows};
                    const double f = 2.0 * eo.rows * C.cols * C.cols;
                    P.flops += f;
                    P.pf[2] += f;
                    P.pb[2] += SZ * (2.0 * eo.rows * C.cols + (double)C.cols * C.cols + NB * C.cols);
                }
                P.tpq.push_back(tk);
            }
        });
        eparts.resize(nep);
        std::vector<GeqrtTask<S>> geq;
        std::vector<TpqrtTask<S>> tpq;
        long nlfb = 0;
        for (const auto& P : eparts) {
            geq.insert(geq.end(), P.geq.begin(), P.geq.end());
            tpq.insert(tpq.end(), P.tpq.begin(), P.tpq.end());
            nlfb += P.nlfb;
            st.flops += P.flops;
            prof_add(P_GEQRT, P.pf[0], P.pb[0]);
            prof_add(P_LARFB, P.pf[1], P.pb[1]);
            prof_add(P_TPQRT, P.pf[2], P.pb[2]);
        }
        hacc(1, ht);
        // ---- upload and launch the elimination's panel work; the trailing-update and scaling
        // descriptors are built while it runs
        TpqrtLink<S>* d_tpql = nullptr;
        {
            std::vector<GeqrtTask<S>> geq_small, geq_big;
            for (const auto& t : geq) (std::min(t.m, t.n) >= big_front ? geq_big : geq_small).push_back(t);
            std::vector<TpqrtTask<S>> tpq_wave, tpq_small, tpq_big, tpq_coop;
            static const bool coop_env = getenv("SPAQR_TPQRT_COOP") == nullptr || atoi(getenv("SPAQR_TPQRT_COOP")) != 0;
            const bool tpqrt_coop_on = coop_env && tuned(T_TPQRT_COOP, 1) == 1;
            // small chains (k <= 64, links of <= 64 rows) go one per wave; the rest stay on the
            // one-workgroup kernel: with the register-resident panel QR the host-driven panel
            // loop (tpqrt_big) costs more in launch gaps than it gains
            static const bool tpqrt_waves = getenv("SPAQR_TPQRT_WAVE") == nullptr || atoi(getenv("SPAQR_TPQRT_WAVE")) != 0;
            for (const auto& t : tpq) {
                bool small = tpqrt_waves && t.k <= 64;
                for (int j = 0; j < t.nlinks && small; ++j) small = tpql[t.link0 + j].rows <= 64;
                if (small) tpq_wave.push_back(t);
                else if (tpqrt_coop_on && t.k > NB && t.nlinks > 1) tpq_coop.push_back(t);  // several panels to pipeline
                else tpq_small.push_back(t);
            }
            const GeqrtBatch geq_b = prep_geqrt(geq_small);
            std::vector<LarfbBuckets> lfb_c(nep), lfb_t(nep);
            for (int i = 0; i < nep; ++i) {
                lfb_c[i] = std::move(eparts[i].lc);
                lfb_t[i] = std::move(eparts[i].lt);
            }
            const LarfbBatch lfb_b = prep_larfb_parts(lfb_c, false), lfb_tb = prep_larfb_parts(lfb_t, true);
            TpqrtTask<S>* d_tpq = upload_vec(tpq_small);
            TpqrtTask<S>* d_tpqw = upload_vec(tpq_wave);
            d_tpql = upload_vec(tpql);
            if (!geq.empty()) {
                int nl = 0;
                timed(P_GEQRT, 0, (long)geq.size(), [&] {
                    run_geqrt(geq_b);
                    nl = geq_b.launches() + geqrt_big<S>(geq_big);
                });
                if (profiling) prof[P_GEQRT].launches += nl;
            }
            if (nlfb)
                timed(P_LARFB, lfb_b.launches() + lfb_tb.launches(), nlfb, [&] {
                    run_larfb(lfb_b);
                    run_larfb(lfb_tb);
                });
            if (!tpq.empty()) {
                int nl = 0;
                timed(P_TPQRT, 0, (long)tpq.size(), [&] {
                    launch_tpqrt_wave<S>(d_tpqw, d_tpql, (int)tpq_wave.size(), stream_);
                    launch_tpqrt<S>(d_tpq, d_tpql, (int)tpq_small.size(), stream_);
                    nl = !tpq_wave.empty() + !tpq_small.empty() + tpqrt_big<S>(tpq_big, tpql) +
                         tpqrt_coop<S>(tpq_coop, d_tpql);
                });
                if (profiling) prof[P_TPQRT].launches += nl;
            }
            // the pivots of every interior's R, now final
            std::vector<PivotCheck<S>> pc;
            pc.reserve(ints.size());
            for (int c : ints) {
                const Edge& es = ed[cl[c].self_edge];
                pc.push_back({(const S*)es.A, cl[c].cols, es.rows, c, L});
            }
            launch_check_pivots<S>(upload_vec(pc), (int)pc.size(), pivot_rel, d_cinfo, d_lflags, d_lfirst, stream_);
            HIPCHK(hipGetLastError());
        }
        hacc(2, ht);
        // colour layers for the shared-block updates: interiors that share an out-neighbour get
        // different colours, and the colours run in order (the update order of the level).
        // The interiors' out lists are final here, so the colouring and then the fill-in of
        // independent interiors run while the panel work does.
        int ncolors = 0;
        if (independent) {
            // the colouring (sequential) runs on its own thread beside the fill-in (parallel); it
            // reads only the pairs taken here, before the fill-in grows the edge list
            std::vector<int> col;
            std::exception_ptr cerr;
            std::thread th([&, pairs = colour_pairs(ints)]() mutable {
                try {
                    ncolors = dsatur_colours(std::move(pairs), (int)ints.size(), col);
                } catch (...) {
                    cerr = std::current_exception();
                }
            });
            try {
                fill_in(true);
            } catch (...) {
                th.join();
                throw;
            }
            th.join();
            if (cerr) std::rethrow_exception(cerr);
            for (size_t i = 0; i < ints.size(); ++i) cl[ints[i]].color = col[i];
            hacc(0, ht);
        } else {
            ncolors = colour_interiors(ints);
            hacc(10, ht);
        }
        st.colors = ncolors;
        // Trailing-update chains, built over the interiors in parallel: links into one array at
        // per-slot offsets (counted first), tasks straight into their colour's launch buckets
        // (0-2 one-wave by link rows, 3-6 C-resident, 7 workgroup), all in interior order.
        const int TILE = tuned(T_TPMQRT_TILE, 64);
        using TpmBuckets = std::array<std::vector<TpmqrtTask<S>>, 8>;
        const int nint = (int)ints.size();
        const int NS = std::min(host_threads(), std::max(1, nint / 64));
        std::vector<long> lcount(NS + 1, 0);
        parallel_ranges(nint, 64, [&](int b, int e, int slot) {
            long n = 0;
            for (int i = b; i < e; ++i) {
                const Cluster& C = cl[ints[i]];
                long nl = 0;
                for (size_t j = 1; j < C.out.size(); ++j) nl += link_slices<S>(ed[C.out[j]].rows);
                n += (long)C.in.size() * nl;
            }
            lcount[slot + 1] = n;
        });
        for (int i = 0; i < NS; ++i) lcount[i + 1] += lcount[i];
        std::vector<TpmqrtLink<S>> tpml((size_t)lcount[NS]);
        std::vector<std::vector<TpmBuckets>> tpm_parts(ncolors, std::vector<TpmBuckets>(NS));
        std::vector<std::array<double, 3>> tpm_acc(NS, std::array<double, 3>{0.0, 0.0, 0.0});  // flops, prof flops, bytes
        parallel_ranges(nint, 64, [&](int b, int e, int slot) {
            std::vector<int> mslot(cl.size(), -1);  // out edge of m by target
            long li = lcount[slot];
            auto& acc = tpm_acc[slot];
            for (int i = b; i < e; ++i) {
                const Cluster& C = cl[ints[i]];
                for (int e2 : C.in) {
                    const Edge& cm = ed[e2];
                    const int m = cm.n1;
                    const int link0 = (int)li;
                    int rmax = 0, rmin = INT_MAX;
                    for (int f : cl[m].out) mslot[ed[f].n2] = f;
                    for (size_t j = 1; j < C.out.size(); ++j) {
                        const Edge& cn = ed[C.out[j]];
                        const int mn = mslot[cn.n2];
                        if (mn < 0) throw std::runtime_error("missing fill-in block");
                        const int ns = link_slices<S>(cn.rows), sz = (cn.rows + ns - 1) / ns;
                        for (int sl = 0; sl < ns; ++sl) {
                            const int r0 = sl * sz, rs = std::min(sz, cn.rows - r0);
                            tpml[li++] = {Ap(C.out[j]) + r0, (const S*)cn.T + (size_t)sl * NB * C.cols, Ap(mn) + r0, rs,
                                          cn.rows, ed[mn].rows, 0};
                            rmax = std::max(rmax, rs);
                            rmin = std::min(rmin, rs);
                        }
                        const double f = 4.0 * cn.rows * C.cols * cm.cols;
                        acc[0] += f;
                        acc[1] += f;
                        acc[2] += SZ * ((double)cn.rows * C.cols + NB * C.cols + 2.0 * C.cols * cm.cols + 2.0 * cn.rows * cm.cols);
                    }
                    for (int f : cl[m].out) mslot[ed[f].n2] = -1;
                    const int nl = (int)li - link0;
                    if (nl == 0) continue;
                    for (int c0 = 0; c0 < cm.cols; c0 += TILE) {
                        const int nc = std::min(TILE, cm.cols - c0);
                        int q = -1;
                        if (small_tasks && C.cols <= 64 && nc <= 64 && rmin >= 1 && rmax <= 64)
                            q = rmax <= 16 ? 0 : (rmax <= 32 ? 1 : 2);
                        else {
                            const int r = tpmqrt_res_bucket_r<S>(nc, rmax);
                            q = r < 0 ? 7 : 3 + r;
                        }
                        tpm_parts[C.color][slot][q].push_back(
                            {Ap(e2) + (size_t)c0 * cm.rows, C.cols, cm.rows, nc, link0, nl, c0});
                    }
                }
            }
        });
        for (const auto& acc : tpm_acc) {
            st.flops += acc[0];
            prof_add(P_TPMQRT, acc[1], acc[2]);
        }
        hacc(11, ht);
        hacc(1, ht);
        {
            TpmqrtLink<S>* d_tpml = upload_vec(tpml);
            std::vector<std::array<std::pair<TpmqrtTask<S>*, int>, 8>> tb(ncolors);
            long ntpm = 0;
            for (int q = 0; q < ncolors; ++q) {
                tb[q] = upload_buckets(tpm_parts[q]);
                for (const auto& x : tb[q]) ntpm += x.second;
                if (dbg_shapes)
                    for (const auto& pb : tpm_parts[q])
                        for (int bq = 3; bq < 8; ++bq)
                            for (const auto& t : pb[bq]) {
                                int rows = 0;
                                for (int j = 0; j < t.nlinks; ++j) rows += tpml[t.link0 + j].rows;
                                shape_add("tpmqrt", t.k, rows / std::max(t.nlinks, 1), t.ncols, 4.0 * t.k * t.ncols * rows);
                            }
            }

            if (ntpm)
                timed(P_TPMQRT, ncolors, ntpm, [&] {
                    for (int q = 0; q < ncolors; ++q) {
                        const auto& u = tb[q];
                        for (int b = 0; b < 3; ++b) launch_tpmqrt_wave<S>(u[b].first, d_tpml, u[b].second, 16 << b, stream_);
                        if constexpr (sizeof(S) == 4) {
                            for (int b = 0; b < 3; ++b) launch_tpmqrt_res(u[3 + b].first, d_tpml, u[3 + b].second, 4 << b, stream_);
                            launch_tpmqrt_res2(u[6].first, d_tpml, u[6].second, stream_);
                        }
                        launch_tpmqrt<S>(u[7].first, d_tpml, u[7].second, stream_);
                    }
                });
            HIPCHK(hipGetLastError());
        }
        hacc(2, ht);
        // ---- row reassignment of a rectangular system (src/ptree.cpp:1149-1356)
        if (rect()) reassign_rows<S>(L);
        // ---- scaling descriptors (src/ptree.cpp:269-286, 1621-1633)
        std::vector<CopyTask<S>> scopy;
        std::vector<GeqrtTask<S>> sgeq;
        for (int c : itf) {
            Cluster& C = cl[c];
            if (!want_sparsify(C)) continue;
            Edge& es = E(C.self_edge);
            if (es.rows < es.cols) throw std::runtime_error("a scaled interface has fewer rows than columns");
            C.Qs = arena.alloc((size_t)es.rows * es.cols * sizeof(S));
            C.Ts = arena.alloc((size_t)NB * es.cols * sizeof(S));
            scopy.push_back({Ap(C.self_edge), (S*)C.Qs, es.rows, es.cols, es.rows, es.rows, 0, 0});
            sgeq.push_back({(S*)C.Qs, (S*)C.Ts, es.rows, es.cols, es.rows, 0});
            st.flops += flops_geqrt<S>(es.rows, es.cols);
            prof_add(P_SGEQRT, flops_geqrt<S>(es.rows, es.cols), SZ * (4.0 * es.rows * es.cols + NB * es.cols));
        }
        hacc(18, ht);
        // the scaling applies of every interface block, built over the interfaces in parallel
        struct SappPart {
            LarfbBuckets c, t;
            long n = 0;
            double flops = 0.0, pflops = 0.0, pbytes = 0.0;
        };
        std::vector<SappPart> sparts(host_threads());
        const int nsp = parallel_ranges((int)itf.size(), 64, [&](int b, int e, int slot) {
            SappPart& P = sparts[slot];
            for (int ii = b; ii < e; ++ii) {
                const int c = itf[ii];
                for (int e2 : cl[c].sout) {
                    const Edge& ee = ed[e2];
                    const Cluster& N2 = cl[ee.n2];
                    const Cluster& N1 = cl[c];
                    const bool w2 = want_sparsify(N2), w1 = want_sparsify(N1);
                    if (!w2 && !w1) continue;
                    const int ldq2 = w2 ? ed[N2.self_edge].rows : 0;
                    const int ldq1 = w1 ? ed[N1.self_edge].rows : 0;
                    file_larfb({w2 ? (const S*)N2.Qs : nullptr, w2 ? (const S*)N2.Ts : nullptr, Ap(e2),
                                w1 ? (const S*)N1.Qs : nullptr, ee.rows, ee.cols, N2.cols, ldq2, ee.rows, N1.cols, ldq1, 0},
                               P.c, P.t);
                    ++P.n;
                    if (w2) {
                        const double f = 4.0 * ee.rows * ee.cols * N2.cols - 2.0 * ee.cols * N2.cols * N2.cols;
                        P.flops += f;
                        P.pflops += f;
                        P.pbytes += SZ * ((double)ee.rows * N2.cols + NB * N2.cols + 2.0 * ee.rows * ee.cols);
                    }
                    if (w1) {
                        const double f = (double)ee.rows * ee.cols * ee.cols;
                        P.flops += f;
                        P.pflops += f;
                        P.pbytes += SZ * (0.5 * ee.cols * ee.cols + 2.0 * ee.rows * ee.cols);
                    }
                }
            }
        });
        sparts.resize(nsp);
        long nsapp = 0;
        for (const auto& P : sparts) {
            nsapp += P.n;
            st.flops += P.flops;
            prof_add(P_SAPPLY, P.pflops, P.pbytes);
        }
        hacc(19, ht);
        // ---- upload and launch the scaling; the RRQR descriptors are built while it runs
        {
            std::vector<GeqrtTask<S>> sgeq_small, sgeq_big;
            for (const auto& t : sgeq) (std::min(t.m, t.n) >= big_front ? sgeq_big : sgeq_small).push_back(t);
            const GeqrtBatch sgeq_b = prep_geqrt(sgeq_small);
            std::vector<LarfbBuckets> sapp_c(nsp), sapp_t(nsp);
            for (int i = 0; i < nsp; ++i) {
                sapp_c[i] = std::move(sparts[i].c);
                sapp_t[i] = std::move(sparts[i].t);
            }
            CopyTask<S>* d_scopy = upload_vec(scopy);
            const LarfbBatch sapp_b = prep_larfb_parts(sapp_c, false), sapp_tb = prep_larfb_parts(sapp_t, true);
            hacc(17, ht);
            if (!sgeq.empty()) {
                int nl = 0;
                timed(P_SGEQRT, 0, (long)sgeq.size(), [&] {
                    launch_copy<S>(d_scopy, (int)scopy.size(), stream_, copy_parts(scopy));
                    run_geqrt(sgeq_b);
                    nl = 1 + sgeq_b.launches() + geqrt_big<S>(sgeq_big);
                });
                if (profiling) prof[P_SGEQRT].launches += nl;
            }
            if (nsapp)
                timed(P_SAPPLY, sapp_b.launches() + sapp_tb.launches(), nsapp, [&] {
                    run_larfb(sapp_b);
                    run_larfb(sapp_tb);
                });
            HIPCHK(hipGetLastError());
        }
        // ---- the merge's structure (src/ptree.cpp:1940-2123): which parents become neighbours,
        // their distance-2 lists and new edges, none of which depends on a rank.
        // Per parent: the merged neighbour and distance-2 lists (reads of the children only, so
        // the parents run in parallel; duplicates dropped through per-thread stamp arrays, stamp
        // = the parent), then the new edges in parent order with ids from prefix sums, entries
        // and out lists filled per parent in parallel, then every target's in lists appended in
        // id order (targets split over the threads by cluster number).  Sizes come later.
        const int Lp = L + 1;
        bool merged = false;
        size_t ne_merge = 0;
        auto merge_structure = [&]() {
            merged = true;
            for (int c : ints) cl[c].eliminated = true;
            ne_merge = ed.size();
            const auto& bl = bottoms[Lp];
            const int np = (int)bl.size();
            std::vector<std::vector<int>> pd2(np), pnb(np);
            // compact per-cluster lookups (the passes below touch clusters at random)
            std::vector<int> par(cl.size()), lvl(cl.size());
            std::vector<char> elim(cl.size());
            parallel_ranges((int)cl.size(), 4096, [&](int b, int e, int) {
                for (int i = b; i < e; ++i) {
                    par[i] = cl[i].parent;
                    lvl[i] = cl[i].lvl();
                    elim[i] = cl[i].eliminated;
                }
            });
            parallel_ranges(np, 64, [&](int b, int e, int) {
                std::vector<int> seen_n(cl.size(), -1), seen_d(cl.size(), -1);
                for (int i = b; i < e; ++i) {
                    const int p = bl[i];
                    std::vector<int>& d2 = pd2[i];
                    std::vector<int>& nbrs = pnb[i];
                    for (int d : cl[p].dist2)
                        if (seen_d[d] != p) { seen_d[d] = p; d2.push_back(d); }
                    for (int ch : cl[p].children) {
                        for (int e2 : cl[ch].out) {
                            const int n2 = ed[e2].n2;
                            if (elim[n2]) continue;
                            const int q = par[n2];
                            if (seen_n[q] != p) { seen_n[q] = p; nbrs.push_back(q); }
                        }
                        for (int d : cl[ch].dist2) {
                            if (lvl[d] < Lp) continue;
                            const int q = par[d];
                            if (seen_d[q] != p) { seen_d[q] = p; d2.push_back(q); }
                        }
                    }
                    std::sort(d2.begin(), d2.end());
                    std::sort(nbrs.begin(), nbrs.end());
                }
            });
            hacc(14, ht);
            std::vector<size_t> eoff(np + 1, 0);
            for (int i = 0; i < np; ++i) eoff[i + 1] = eoff[i] + pnb[i].size();
            const size_t base = ed.size(), nnew = eoff[np];
            ed.resize(base + nnew);
            parallel_ranges(np, 64, [&](int b, int e, int) {
                for (int i = b; i < e; ++i) {
                    const int p = bl[i];
                    Cluster& P = cl[p];
                    P.dist2 = std::move(pd2[i]);
                    P.out.reserve(P.out.size() + pnb[i].size());
                    for (size_t j = 0; j < pnb[i].size(); ++j) {
                        const int nn = pnb[i][j], id = (int)(base + eoff[i] + j);
                        Edge& E = ed[id];
                        E.n1 = p;
                        E.n2 = nn;
                        E.spars = lvl[p] > Lp && lvl[nn] > Lp && p != nn;
                        P.out.push_back(id);
                        if (E.spars) P.sout.push_back(id);
                        if (p == nn) P.self_edge = id;
                    }
                    sort_out(P);
                }
            });
            const int NT = host_threads();
            std::vector<std::vector<int>> ib(NT);
            for (size_t id = base; id < base + nnew; ++id)
                if (ed[id].n1 != ed[id].n2) ib[ed[id].n2 % NT].push_back((int)id);
            parallel_ranges(NT, nnew < 16384 ? NT : 1, [&](int b, int e, int) {  // small levels: one thread
                for (int t = b; t < e; ++t)
                    for (int id : ib[t]) {
                        Cluster& C2 = cl[ed[id].n2];
                        C2.in.push_back(id);
                        if (ed[id].spars) C2.sin.push_back(id);
                    }
            });
            hacc(3, ht);
        };
        // ---- RRQR descriptors (src/ptree.cpp:373-439)
        std::vector<CopyTask<S>> gat;
        std::vector<int> gat_begin;  // first gather copy of each RRQR task
        std::vector<Geqp3Task<S>> gq;
        std::vector<int> gq_cluster;
        std::vector<char> used_sketch;
        int* d_ranks = nullptr;
        int* d_rrqr_err = nullptr;
        std::vector<CopyTask<S>> gatx;   // rectangular systems: the extra rows' RRQR
        std::vector<Geqp3Task<S>> gqx;
        int* d_xranks = nullptr;         // per sparsified interface (order of gq), 0 when no task
        // the level guard of src/ptree.cpp:1566,1635; want_sparsify applies each system's own
        // top-two-levels rule
        const bool spars_level = (L >= skip && tol != 0);
        if (spars_level) {
            std::vector<int> who;
            for (int c : itf)
                if (want_sparsify(cl[c])) who.push_back(c);
            if (!who.empty()) d_ranks = (int*)arena.alloc(sizeof(int) * who.size());
            // Per interface: C gathered from its blocks (k x w), Q, tau, T, two norm vectors and
            // the pivots, in one arena block at per-interface offsets (sizes first, then the
            // tasks filled in parallel; each piece 256-byte aligned as an arena allocation is).
            const int nw = (int)who.size();
            auto al = [](size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; };
            std::vector<size_t> boff(nw + 1, 0);
            std::vector<int> goff(nw + 1, 0), wq(nw);
            parallel_ranges(nw, 256, [&](int b, int e, int) {
                for (int q = b; q < e; ++q) {
                    const Cluster& C = cl[who[q]];
                    const int k = C.cols;
                    int w = 0;
                    for (int e2 : C.sout) w += ed[e2].rows;
                    for (int e2 : C.sin) w += ed[e2].cols;
                    const size_t kmax = (size_t)std::max(1, std::min(k, w)), w1 = (size_t)std::max(w, 1);
                    wq[q] = w;
                    boff[q + 1] = al((size_t)k * w1 * sizeof(S)) + al((size_t)k * kmax * sizeof(S)) + al(kmax * sizeof(S)) +
                                  al((size_t)NB * kmax * sizeof(S)) + al(2 * w1 * sizeof(S)) + al(w1 * sizeof(int));
                    goff[q + 1] = (int)(C.sout.size() + C.sin.size());
                }
            });
            for (int q = 0; q < nw; ++q) {
                boff[q + 1] += boff[q];
                goff[q + 1] += goff[q];
            }
            char* blk = nw ? (char*)arena.alloc(boff[nw]) : nullptr;
            gat.resize((size_t)goff[nw]);
            gat_begin.assign(goff.begin(), goff.end() - 1);
            gq.resize(nw);
            gq_cluster = who;
            std::vector<std::array<double, 2>> pacc(host_threads(), std::array<double, 2>{0.0, 0.0});
            parallel_ranges(nw, 256, [&](int b, int e, int slot) {
                for (int q = b; q < e; ++q) {
                    Cluster& C = cl[who[q]];
                    const int k = C.cols, w = wq[q];
                    const size_t kmax = (size_t)std::max(1, std::min(k, w)), w1 = (size_t)std::max(w, 1);
                    char* p = blk + boff[q];
                    auto take = [&](size_t bytes) { char* r = p; p += al(bytes); return r; };
                    S* Cb = (S*)take((size_t)k * w1 * sizeof(S));
                    C.Qsp = take((size_t)k * kmax * sizeof(S));
                    C.tausp = take(kmax * sizeof(S));
                    C.Tsp = take((size_t)NB * kmax * sizeof(S));
                    S* vn = (S*)take(2 * w1 * sizeof(S));
                    int* jp = (int*)take(w1 * sizeof(int));
                    int is = 0, gi = goff[q];
                    for (int e2 : C.sout) {
                        gat[gi++] = {Ap(e2), Cb + (size_t)is * k, k, ed[e2].rows, ed[e2].rows, k, 1, 0};
                        is += ed[e2].rows;
                    }
                    for (int e2 : C.sin) {
                        gat[gi++] = {Ap(e2), Cb + (size_t)is * k, k, ed[e2].cols, ed[e2].rows, k, 0, 0};
                        is += ed[e2].cols;
                    }
                    gq[q] = {Cb, (S*)C.Qsp, (S*)C.tausp, (S*)C.Tsp, vn, vn + w1, jp, d_ranks + q, k, w, 0, 0};
                    pacc[slot][0] += SZ * 2.0 * k * w;
                    pacc[slot][1] += SZ * 3.0 * k * w;
                }
            });
            for (const auto& a : pacc) {
                prof_add(P_GATHER, 0.0, a[0]);
                if (rank_mode == 0) prof_add(P_GEQP3, 0.0, a[1]);
            }
            // sparsify_extra (src/ptree.cpp:441-493): the same RRQR on each interface's extra
            // rows, C_ex = [A(c, m)(k:rows, :) ...] over its row blocks, always by geqp3
            if (rect() && !who.empty()) {
                d_xranks = (int*)arena.alloc(sizeof(int) * who.size());
                launch_fill<int>(d_xranks, who.size(), 0, stream_);
                for (size_t q = 0; q < who.size(); ++q) {
                    Cluster& C = cl[who[q]];
                    const int k = C.cols, ex = C.rows - C.cols;
                    int w = 0;
                    for (int e : C.sin) w += ed[e].cols;
                    C.nex = ex;
                    if (ex <= 0 || w <= 0) continue;
                    const int kmax = std::min(ex, w);
                    S* Cb = (S*)arena.alloc((size_t)ex * w * sizeof(S));
                    int js = 0;
                    for (int e : C.sin) {
                        gatx.push_back({Ap(e) + k, Cb + (size_t)js * ex, ex, ed[e].cols, ed[e].rows, ex, 0, 0});
                        js += ed[e].cols;
                    }
                    C.Qex = arena.alloc((size_t)ex * kmax * sizeof(S));
                    C.tauex = arena.alloc((size_t)kmax * sizeof(S));
                    C.Tex = arena.alloc((size_t)NB * kmax * sizeof(S));
                    S* vn = (S*)arena.alloc((size_t)2 * w * sizeof(S));
                    int* jp = (int*)arena.alloc((size_t)w * sizeof(int));
                    gqx.push_back({Cb, (S*)C.Qex, (S*)C.tauex, (S*)C.Tex, vn, vn + w, jp, d_xranks + q, ex, w, 0, 0});
                    prof_add(P_GATHER, 0.0, SZ * 2.0 * ex * w);
                    prof_add(P_GEQP3, 0.0, SZ * 3.0 * ex * w);
                }
            }
        }
        hacc(16, ht);
        {
            CopyTask<S>* d_gat = upload_vec(gat);
            if (!gat.empty()) timed(P_GATHER, 1, (long)gat.size(), [&] { launch_copy<S>(d_gat, (int)gat.size(), stream_, copy_parts(gat)); });
            if (!gq.empty()) {
                RrqrAlloc mem{[&](size_t b) { return arena.alloc(b); },
                              [&](const void* h, size_t b) { return upload(h, b); }};
                mem.graph = capturing;
                const int zero = 0;
                d_rrqr_err = (int*)upload(&zero, sizeof(int));
                launch_fill<int>(d_rrqr_err, 1, 0, stream_);  // also on every graph replay
                gat_begin.push_back((int)gat.size());
                used_sketch.assign(gq.size(), 0);
                if (rank_mode == 0 && !rank_log) {
                    int nd = 0;
                    timed(P_GEQP3, 0, (long)gq.size(),
                          [&] { geqp3_grouped<S>(gq, tol, kPivotTieRel, stream_, mem, d_rrqr_err, &nd); });
                    if (profiling) prof[P_GEQP3].launches += nd;
                } else {
                    st.flops += rrqr_level<S>(gq, gat, gat_begin, gq_cluster, L, d_ranks, d_rrqr_err, used_sketch);
                }
                if (!gqx.empty()) {
                    CopyTask<S>* d_gatx = upload_vec(gatx);
                    timed(P_GATHER, 1, (long)gatx.size(), [&] { launch_copy<S>(d_gatx, (int)gatx.size(), stream_, copy_parts(gatx)); });
                    int nd = 0;
                    timed(P_GEQP3, 0, (long)gqx.size(),
                          [&] { geqp3_grouped<S>(gqx, tol, kPivotTieRel, stream_, mem, d_rrqr_err, &nd); });
                    if (profiling) prof[P_GEQP3].launches += nd;
                }
            }
            HIPCHK(hipGetLastError());
        }
        if (!gq.empty()) {
            seg_end(d_ranks, (int)gq.size(), d_rrqr_err);
            solve_sched_levels<S>(L);  // the finished levels' solve schedule, while the device works
            if (L < nlevels - 1) {
                hacc(2, ht);
                merge_structure();  // needs no rank: runs while the device works too
            }
            std::vector<int> ranks(gq.size()), xranks(d_xranks ? gq.size() : 0, 0);
            int rrqr_err = 0;
            HIPCHK(hipMemcpyAsync(ranks.data(), d_ranks, sizeof(int) * ranks.size(), hipMemcpyDeviceToHost, stream_));
            if (d_xranks)
                HIPCHK(hipMemcpyAsync(xranks.data(), d_xranks, sizeof(int) * xranks.size(), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipMemcpyAsync(&rrqr_err, d_rrqr_err, sizeof(int), hipMemcpyDeviceToHost, stream_));
            sync_stream();
            if (rrqr_err) throw std::runtime_error("cooperative geqp3: a barrier wait exceeded its spin limit");
            check_levels(L);
            for (size_t q = 0; q < gq.size(); ++q) {
                Cluster& C = cl[gq_cluster[q]];
                const int k = C.cols;
                const int w = gq[q].w;
                st.cols_before += k;
                st.ranks_after += ranks[q];
                int r = ranks[q];
                if (!used_sketch[q])
                    for (int j = 0; j < r; ++j) {
                        st.flops += 4.0 * (k - j) * (w - j);
                        prof_add(P_GEQP3, 4.0 * (k - j) * (w - j), 0.0);
                    }
                C.sparsified = true;
                C.rank = r;
                // reset_size (src/ptree.cpp:435-438); a rectangular system keeps the extra rows' rank
                // below the pivot block's [I_r; 0] (sparsify_resize_tf :1835-1850)
                C.row_rank = d_xranks ? xranks[q] : 0;
                C.cols = r;
                C.rows = r + C.row_rank;
            }
            st.n_sparsified = (int)gq.size();
            if (graph_rec) {
                gplan.rb_ranks.back() = ranks;
                seg_begin();
            }
        }
        hacc(2, ht);
        if (dbg_host)
            fprintf(stderr, "[level %d] interiors %d interfaces %d max front %d  host %.1f ms so far, GPU+sync %.1f ms\n", L,
                    st.n_interiors, st.n_interfaces, st.max_front, hms[0] + hms[1] + hms[3] + hms[4], hms[2]);
        stats.push_back(st);
        if (L == nlevels - 1) break;

        // ---- merge (src/ptree.cpp:1940-2123); its structure was built above when this level
        // read ranks back
        if (!merged) merge_structure();
        // The sparsify apply's right transform (A <- A Q_n, in place on the interface blocks, row
        // chunks) needs nothing from the merge: it is queued first and runs while the host merges.
        // (FP32 row chunks of 128: one chunk per workgroup of the C-resident kernel)
        const int kSparsRows = tuned(T_SPARS_ROWS, sizeof(S) == 4 ? 128 : 256), kSparsCols = tuned(T_SPARS_COLS, 128);
        {
            std::vector<SparsBuckets> spa1(host_threads());
            const int np1 = parallel_ranges((int)itf.size(), 256, [&](int b, int e, int slot) {
                SparsBuckets& out = spa1[slot];
                for (int ii = b; ii < e; ++ii) {
                    const int so = itf[ii];
                    const Cluster& SO = cl[so];
                    if (!SO.sparsified) continue;
                    for (int e2 : SO.out) {
                        const Edge& eo = ed[e2];
                        const int no = eo.n2;
                        if (no == so || cl[no].eliminated) continue;
                        SparsApplyTask<S> u{};
                        u.Vn = (const S*)SO.Qsp;
                        u.Tn = (const S*)SO.Tsp;
                        u.ldvn = eo.cols;
                        u.nc = eo.cols;
                        u.lda = eo.rows;
                        u.kn = eo.cols;
                        u.kc = eo.rows;
                        u.rn = SO.cols;
                        u.pad = 1;
                        for (int r0 = 0; r0 < eo.rows; r0 += kSparsRows) {
                            u.A = Ap(e2) + r0;
                            u.mc = std::min(kSparsRows, eo.rows - r0);
                            out[spars_bucket(u)].push_back(u);
                        }
                    }
                }
            });
            spa1.resize(np1);
            size_t n1 = 0;
            for (const auto& v : spa1)
                for (const auto& w : v) n1 += w.size();
            if (n1) {
                const SparsBatch sb1 = prep_spars_parts(spa1);
                timed(P_SPARS, sb1.launches(), 0, [&] { run_spars(sb1); });
            }
        }
        hacc(12, ht);
        // the parents' sizes from their children's (ranks now known), then the new edges' sizes
        current_bottom = Lp;
        for (int p : bottoms[Lp]) {
            int rs = 0, cs = 0;
            for (int ch : cl[p].children) {
                cl[ch].rpos = rs;
                cl[ch].cpos = cs;
                rs += cl[ch].rows;
                cs += cl[ch].cols;
            }
            cl[p].rows = cl[p].rows_org = rs;
            cl[p].cols = cl[p].cols_org = cs;
            cl[p].rank = cs;
            cl[p].row_rank = rs - cs;
        }
        hacc(13, ht);
        parallel_ranges((int)(ed.size() - ne_merge), 4096, [&](int b, int e, int) {
            for (int i = b; i < e; ++i) {
                Edge& E = ed[ne_merge + i];
                E.rows = cl[E.n2].rows;
                E.cols = cl[E.n1].cols;
            }
        });
        hacc(3, ht);
        {
            size_t tot = 0;
            for (size_t i = ne_merge; i < ed.size(); ++i) tot += (size_t)ed[i].rows * ed[i].cols;
            S* blk = zalloc(tot);
            size_t off = 0;
            for (size_t i = ne_merge; i < ed.size(); ++i) {
                ed[i].A = blk + off;
                off += (size_t)ed[i].rows * ed[i].cols;
            }
        }
        hacc(7, ht);
        // Sparsify-apply tasks of every interface block, built in parallel over the interfaces
        // (per-thread lists, concatenated in interface order).  The left transform Q_c^T A and
        // the scatter act on each column: column chunks, so that large blocks spread over many
        // workgroups (the right transform was queued before the merge).
        struct SparsPart {
            SparsBuckets tasks;
            long blocks = 0;
            double flops = 0.0, pflops = 0.0, pbytes = 0.0;
        };
        std::vector<SparsPart> parts(host_threads());
        const int nparts = parallel_ranges((int)itf.size(), 256, [&](int b, int e, int slot) {
            SparsPart& P = parts[slot];
            for (int ii = b; ii < e; ++ii) {
                const int so = itf[ii];
                const Cluster& SO = cl[so];
                const std::vector<int>& pout = cl[SO.parent].out;  // sorted by target
                for (int e2 : SO.out) {
                    const Edge& eo = ed[e2];
                    const int no = eo.n2;
                    if (cl[no].eliminated) continue;
                    const Cluster& NO = cl[no];
                    // the parent's out edge to NO's parent
                    auto it = std::lower_bound(pout.begin(), pout.end(), NO.parent,
                                               [&](int x, int target) { return ed[x].n2 < target; });
                    if (it == pout.end() || ed[*it].n2 != NO.parent) throw std::runtime_error("merge: parent edge missing");
                    const int f = *it;
                    SparsApplyTask<S> t{};
                    t.A = Ap(e2);
                    t.dst = Ap(f) + NO.rpos + (size_t)SO.cpos * ed[f].rows;
                    t.ldd = ed[f].rows;
                    t.mc = eo.rows;
                    t.nc = eo.cols;
                    t.lda = eo.rows;
                    t.kn = eo.cols;
                    t.kc = NO.sparsified ? NO.cols_org : eo.rows;
                    t.rn = SO.cols;
                    t.rc = NO.sparsified ? NO.cols : NO.rows;
                    ++P.blocks;
                    if (no == so) {
                        t.ident = SO.sparsified ? 1 : 0;
                        if (t.ident) t.rc = SO.cols;  // [I_r; 0]: the rows below are zero already
                        P.tasks[spars_bucket(t)].push_back(t);
                        continue;
                    }
                    if (SO.sparsified) { t.Vn = (const S*)SO.Qsp; t.Tn = (const S*)SO.Tsp; t.ldvn = eo.cols; }
                    if (NO.sparsified) { t.Vc = (const S*)NO.Qsp; t.Tc = (const S*)NO.Tsp; t.ldvc = NO.cols_org; }
                    // the extra rows of a sparsified rectangular interface: Q_ex^T A(c, n)(k:rows, :),
                    // kept below its rank rows (src/ptree.cpp:1891-1894)
                    if (NO.sparsified && NO.row_rank > 0) {
                        SparsApplyTask<S> x = t;
                        x.A = t.A + NO.cols_org;
                        x.mc = NO.nex;
                        x.kc = NO.nex;
                        x.Vc = (const S*)NO.Qex;
                        x.Tc = (const S*)NO.Tex;
                        x.ldvc = NO.nex;
                        x.rc = NO.row_rank;
                        x.dst = t.dst + NO.cols;
                        x.Vn = nullptr;
                        x.Tn = nullptr;
                        for (int c0 = 0; c0 < x.rn; c0 += kSparsCols) {
                            SparsApplyTask<S> u = x;
                            u.A = x.A + (size_t)c0 * x.lda;
                            u.dst = x.dst + (size_t)c0 * x.ldd;
                            u.rn = std::min(kSparsCols, x.rn - c0);
                            u.pad = 2;
                            P.tasks[spars_bucket(u)].push_back(u);
                        }
                    }
                    const double fl = (SO.sparsified ? 4.0 * eo.rows * eo.cols * SO.cols : 0.0) +
                                      (NO.sparsified ? 4.0 * eo.rows * SO.cols * NO.cols : 0.0);
                    P.flops += fl;
                    P.pflops += fl;
                    P.pbytes += SZ * (2.0 * eo.rows * eo.cols + (double)NO.rows * SO.cols);
                    for (int c0 = 0; c0 < t.rn; c0 += kSparsCols) {
                        SparsApplyTask<S> u = t;
                        u.A = t.A + (size_t)c0 * t.lda;
                        u.dst = t.dst + (size_t)c0 * t.ldd;
                        u.rn = std::min(kSparsCols, t.rn - c0);
                        u.pad = 2;
                        P.tasks[spars_bucket(u)].push_back(u);
                    }
                }
            }
        });
        long nspa = 0;
        std::vector<SparsBuckets> spa2(nparts);
        for (int q = 0; q < nparts; ++q) {
            nspa += parts[q].blocks;
            stats.back().flops += parts[q].flops;
            prof_add(P_SPARS, parts[q].pflops, parts[q].pbytes);
            spa2[q] = std::move(parts[q].tasks);
        }
        hacc(8, ht);
        const SparsBatch sb2 = prep_spars_parts(spa2);
        hacc(9, ht);
        if (nspa) timed(P_SPARS, sb2.launches(), nspa, [&] { run_spars(sb2); });
        HIPCHK(hipGetLastError());
        hacc(4, ht);
        if (dbg_host && L < 3) {
            static double prev[24];
            if (L == 0) std::fill(prev, prev + 24, 0.0);
            fprintf(stderr, "[host L%d] fillin %.2f S1 %.2f colour %.2f S2 %.2f launch %.2f scaling %.2f sapp %.2f rrqr %.2f "
                            "prep %.2f spa1 %.2f parents %.2f nbrs %.2f edges %.2f blocks %.2f spars %.2f upload %.2f ms\n",
                    L, hms[0] - prev[0], hms[1] - prev[1], hms[10] - prev[10], hms[11] - prev[11], hms[2] - prev[2],
                    hms[18] - prev[18], hms[19] - prev[19], hms[16] - prev[16], hms[17] - prev[17], hms[12] - prev[12],
                    hms[13] - prev[13], hms[14] - prev[14], hms[3] - prev[3], hms[7] - prev[7], hms[8] - prev[8],
                    hms[9] - prev[9]);
            std::copy(hms, hms + 24, prev);
        }
    }
    auto ht2 = hclock();
    seg_end(nullptr, 0, nullptr);
    gplan.chain_err = d_chain_err;
    {
        int chain_err = 0;
        HIPCHK(hipMemcpyAsync(&chain_err, d_chain_err, sizeof(int), hipMemcpyDeviceToHost, stream_));
        sync_stream();
        if (chain_err) throw std::runtime_error("cooperative tpqrt: a wait exceeded its spin limit");
        check_levels(nlevels - 1);
    }
    hacc(5, ht2);
    factorized = true;
    solve_sched_levels<S>(nlevels);
    const int rc = 0;
    sync_stream();
    hacc(6, ht2);
    if (dbg_host)
        fprintf(stderr,
                "[host] tables %.1f  fillin %.1f  descr %.1f (colour %.1f tpmqrt %.1f)  launch+ranks %.1f  merge: spa1 %.1f "
                "parents %.1f nbrs %.1f edges %.1f (+blocks %.1f)  spars build %.1f split+upload %.1f launch %.1f  tail-sync "
                "%.1f  solve-sched %.1f ms  [scaling %.1f sapp %.1f rrqr %.1f  prep %.1f]\n",
                hms[15], hms[0], hms[1], hms[10], hms[11], hms[2], hms[12], hms[13], hms[14], hms[3], hms[7], hms[8], hms[9],
                hms[4], hms[5], hms[6], hms[18], hms[19], hms[16], hms[17]);
    if (dbg_shapes) {
        double tot = 0;
        for (auto& e : shapes) tot += e.second[1];
        for (auto& e : shapes)
            if (e.second[1] > 0.005 * tot)
                fprintf(stderr, "[shape] %s tasks %8.0f  GF %8.2f  (%4.1f%%)\n", e.first.c_str(), e.second[0],
                        e.second[1] * 1e-9, 100.0 * e.second[1] / tot);
    }
    t_factor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

// ---------------------------------------------------------------------------- row reassignment
// A rectangular system's level step after the elimination (src/ptree.cpp:1149-1356 with
// Cluster::update_extra_rnorm, src/cluster.cpp:151-160).  Rows cols(c) .. rows(c) of every
// eliminated interior c's row blocks Q_c^T A(c, :) touch no eliminated column; each goes to the
// out-neighbour m of c that is also an in-neighbour (is_cnbr) whose block holds the largest norm
// of that row -- a strictly larger norm replaces, so on a tie the first neighbour in cluster order
// keeps the row, and all-zero rows stay with c and are dropped.  row_perm stable-sorts the extra
// rows by target.  Each interface then appends the rows it received, in interior order, to its
// diagonal block and to the blocks of its in-edges from non-eliminated clusters (shift_tf +
// Edge::merge, include/edge.h:49-57); a block whose source has no edge into c gets zero rows.
// The row norms are the one device read-back; the grown blocks are one gather launch.
template <typename S>
void Tree::reassign_rows(int L) {
    const auto& ints = interiors[L];
    struct Cand { int c, m, e; size_t off; };
    std::vector<Cand> cands;
    std::vector<RowNormTask<S>> tasks;
    size_t tot = 0;
    for (int c : ints) {
        Cluster& C = cl[c];
        const int ex = C.rows - C.cols;
        C.r2c.assign(std::max(ex, 0), c);
        C.row_perm.clear();
        if (ex <= 0) continue;
        const size_t first = cands.size();
        for (int e : C.in) {
            const int mm = ed[e].n1;
            if (find_out(C, mm) >= 0) cands.push_back({c, mm, e, 0});
        }
        std::sort(cands.begin() + first, cands.end(), [](const Cand& a, const Cand& b) { return a.m < b.m; });
        for (size_t q = first; q < cands.size(); ++q) {
            cands[q].off = tot;
            const Edge& E = ed[cands[q].e];
            tasks.push_back({(const S*)E.A + C.cols, nullptr, E.rows, ex, E.cols, 0});
            tot += (size_t)ex;
        }
    }
    std::vector<double> nrm(tot);
    if (tot) {
        double* d_norm = (double*)arena.alloc(tot * sizeof(double));
        for (size_t q = 0; q < tasks.size(); ++q) tasks[q].out = d_norm + cands[q].off;
        launch_row_norms<S>(upload_vec(tasks), (int)tasks.size(), stream_);
        HIPCHK(hipMemcpyAsync(nrm.data(), d_norm, tot * sizeof(double), hipMemcpyDeviceToHost, stream_));
        sync_stream();
    }
    // targets and their order, then the runs per out-neighbour
    std::vector<double> best;
    for (size_t q0 = 0; q0 < cands.size();) {
        const int c = cands[q0].c;
        Cluster& C = cl[c];
        const int ex = C.rows - C.cols;
        best.assign(ex, 0.0);
        size_t q1 = q0;
        for (; q1 < cands.size() && cands[q1].c == c; ++q1)
            for (int i = 0; i < ex; ++i) {
                const double v = nrm[cands[q1].off + i];
                if (v > best[i]) { best[i] = v; C.r2c[i] = cands[q1].m; }
            }
        q0 = q1;
    }
    for (int c : ints) {
        Cluster& C = cl[c];
        const int ex = (int)C.r2c.size();
        if (ex == 0) continue;
        C.row_perm.resize(ex);
        for (int i = 0; i < ex; ++i) C.row_perm[i] = i;
        std::stable_sort(C.row_perm.begin(), C.row_perm.end(), [&](int a, int b) { return C.r2c[a] < C.r2c[b]; });
        for (size_t j = 1; j < C.out.size(); ++j) {
            const int nn = ed[C.out[j]].n2;
            int i0 = 0;
            while (i0 < ex && C.r2c[C.row_perm[i0]] != nn) ++i0;
            int i1 = i0;
            while (i1 < ex && C.r2c[C.row_perm[i1]] == nn) ++i1;
            if (i1 > i0) {
                cl[nn].extra_from.push_back({c, i0, i1 - i0});
                cl[nn].tot_extra += i1 - i0;
            }
        }
    }
    // the grown blocks
    std::vector<RowGatherTask<S>> gt;
    std::vector<RowSrc<S>> src;
    for (int nn : interfaces[L]) {
        Cluster& N = cl[nn];
        if (N.tot_extra == 0) continue;
        std::sort(N.extra_from.begin(), N.extra_from.end());
        std::vector<int> grow{N.self_edge};
        for (int e : N.in)
            if (cl[ed[e].n1].lvl() != L) grow.push_back(e);
        for (int e : grow) {
            Edge& E = ed[e];
            const Cluster& N1 = cl[E.n1];
            const int rn = E.rows + N.tot_extra;
            S* dst = (S*)arena.alloc((size_t)rn * std::max(E.cols, 1) * sizeof(S));
            gt.push_back({dst, (const S*)E.A, rn, E.rows, E.rows, E.cols, (int)src.size(), N.tot_extra});
            for (const auto& t : N.extra_from) {
                const Cluster& C = cl[t[0]];
                const int fe = find_out(N1, t[0]);  // the block A(c, n1)
                for (int i = 0; i < t[2]; ++i)
                    src.push_back(fe >= 0 ? RowSrc<S>{(const S*)ed[fe].A + C.cols + C.row_perm[t[1] + i], ed[fe].rows, 0}
                                          : RowSrc<S>{nullptr, 0, 0});
            }
            E.A = dst;
            E.rows = rn;
        }
        N.rows += N.tot_extra;
        N.rank = N.cols;
        N.row_rank = N.rows - N.cols;
    }
    launch_row_gather<S>(upload_vec(gt), upload_vec(src), (int)gt.size(), stream_);
    HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------------------- large fronts
// A front of k >= kBigFront columns would keep one workgroup busy for tens of milliseconds
// (its panels are sequential).  Its panel loop runs on the host instead, one launch per
// 32-column panel with look-ahead: for every large front one workgroup applies the previous
// panel to this panel's columns and factors them, while other workgroups apply the previous
// panel to the remaining trailing columns in 128-column chunks.  A last launch applies the
// final panel of the wide fronts to their columns past min(m, n).  The arithmetic per
// column is that of the one-workgroup kernels.

template <typename S>
int Tree::geqrt_big(const std::vector<GeqrtTask<S>>& big) {
    if (big.empty()) return 0;
    int kmax = 0;
    for (const auto& t : big) kmax = std::max(kmax, std::min(t.m, t.n));
    // every step's descriptors are built and uploaded at once; step s launches its slice
    std::vector<PanelStepTask<S>> ps;
    std::vector<int> off;
    const int nsteps = (kmax + NB - 1) / NB + 1;
    const int chunk = tuned(T_PANEL_COLS, 128);
    for (int s = 0; s < nsteps; ++s) {
        off.push_back((int)ps.size());
        const int p0 = s * NB, pp = p0 - NB;  // this panel, the previous one
        std::vector<PanelStepTask<S>> rest;
        for (const auto& t : big) {
            const int kmin = std::min(t.m, t.n);
            const bool prev = s > 0 && pp < kmin, cur = p0 < kmin;
            PanelApplyTask<S> a{};
            if (prev) {
                const int nbq = std::min(NB, kmin - pp);
                a = {t.A + pp + (size_t)pp * t.lda, t.T + (size_t)pp * NB, nullptr, nullptr, t.m - pp, nbq, t.lda, 0, t.lda, 0};
            }
            int c0 = p0;  // first column the previous panel still has to reach
            if (cur) {
                const int nbp = std::min(NB, kmin - p0);
                PanelStepTask<S> u{};
                u.apply = prev;
                if (prev) {
                    u.a = a;
                    u.a.C = t.A + pp + (size_t)p0 * t.lda;
                    u.a.ncols = nbp;
                }
                u.qr = 1;
                S* Ap = t.A + p0 + (size_t)p0 * t.lda;
                u.q = {Ap, Ap, t.T + (size_t)p0 * NB, t.m - p0, nbp, t.lda, t.lda};
                ps.push_back(u);
                c0 = p0 + nbp;
            }
            if (prev)
                for (; c0 < t.n; c0 += chunk) {
                    PanelStepTask<S> u{};
                    u.apply = 1;
                    u.a = a;
                    u.a.C = t.A + pp + (size_t)c0 * t.lda;
                    u.a.ncols = std::min(chunk, t.n - c0);
                    rest.push_back(u);
                }
        }
        ps.insert(ps.end(), rest.begin(), rest.end());  // the panels first: they start the next step
    }
    off.push_back((int)ps.size());
    auto* d = upload_vec(ps);
    int launches = 0;
    for (size_t st = 0; st + 1 < off.size(); ++st)
        if (off[st + 1] > off[st]) {
            launch_panel_step<S>(d + off[st], off[st + 1] - off[st], stream_);
            ++launches;
        }
    return launches;
}
template int Tree::geqrt_big<float>(const std::vector<GeqrtTask<float>>&);
template int Tree::geqrt_big<double>(const std::vector<GeqrtTask<double>>&);

template <typename S>
int Tree::tpqrt_big(const std::vector<TpqrtTask<S>>& big, const std::vector<TpqrtLink<S>>& links) {
    if (big.empty()) return 0;
    int maxl = 0;
    for (const auto& t : big) maxl = std::max(maxl, t.nlinks);
    std::vector<PanelTask<S>> pq;
    std::vector<PanelApplyTask<S>> pa;
    std::vector<int> qoff, aoff;
    for (int li = 0; li < maxl; ++li) {
        int kmax = 0;
        for (const auto& t : big)
            if (t.nlinks > li) kmax = std::max(kmax, t.k);
        for (int p0 = 0; p0 < kmax; p0 += NB) {
            qoff.push_back((int)pq.size());
            aoff.push_back((int)pa.size());
            for (const auto& t : big) {
                if (t.nlinks <= li || p0 >= t.k) continue;
                const TpqrtLink<S>& L = links[t.link0 + li];
                const int nbp = std::min(NB, t.k - p0);
                S* Rp = t.R + p0 + (size_t)p0 * t.ldr;
                S* Bp = L.B + (size_t)p0 * L.ldb;
                pq.push_back({Rp, Bp, L.T + (size_t)p0 * NB, L.rows, nbp, t.ldr, L.ldb});
                for (int c0 = p0 + nbp; c0 < t.k; c0 += 128)
                    pa.push_back({Bp, L.T + (size_t)p0 * NB, t.R + p0 + (size_t)c0 * t.ldr, L.B + (size_t)c0 * L.ldb,
                                  L.rows, nbp, L.ldb, t.ldr, L.ldb, std::min(128, t.k - c0)});
            }
        }
    }
    qoff.push_back((int)pq.size());
    aoff.push_back((int)pa.size());
    auto* dq = upload_vec(pq);
    auto* da = upload_vec(pa);
    int launches = 0;
    for (size_t st = 0; st + 1 < qoff.size(); ++st) {
        launch_panel_qr<S>(dq + qoff[st], qoff[st + 1] - qoff[st], true, stream_);
        launch_panel_apply<S>(da + aoff[st], aoff[st + 1] - aoff[st], true, stream_);
        launches += 1 + (aoff[st + 1] > aoff[st]);
    }
    return launches;
}

// Cooperative tpqrt chains, packed into launches of at most the co-resident number of
// workgroups (a workgroup waits only on lower-numbered workgroups of its own chain).
template <typename S>
int Tree::tpqrt_coop(const std::vector<TpqrtTask<S>>& chains, const TpqrtLink<S>* d_links) {
    if (chains.empty()) return 0;
    static const int maxb = tpqrt_coop_max_blocks<S>();
    if (maxb <= 0) throw std::runtime_error("cooperative tpqrt: no occupancy");
    int panels = 0;
    for (const auto& t : chains) panels += (t.k + NB - 1) / NB;
    int* cnt = (int*)arena.alloc(sizeof(int) * panels);
    launch_fill<int>(cnt, (size_t)panels, 0, stream_);
    int launches = 0, coff = 0;
    for (size_t i = 0; i < chains.size();) {
        std::vector<TpqrtCoopTask<S>> tk;
        std::vector<int> b2t;
        int nb = 0;
        while (i < chains.size()) {
            const TpqrtTask<S>& c = chains[i];
            const int P = (c.k + NB - 1) / NB;
            if (P > maxb) throw std::runtime_error("cooperative tpqrt: chain wider than the co-resident limit");
            if (nb + P > maxb) break;
            tk.push_back({c.R, cnt + coff, c.k, c.ldr, c.link0, c.nlinks, nb, 0});
            for (int q = 0; q < P; ++q) b2t.push_back((int)tk.size() - 1);
            nb += P;
            coff += P;
            ++i;
        }
        launch_tpqrt_coop<S>(upload_vec(tk), upload_vec(b2t), nb, d_links, d_chain_err, stream_);
        ++launches;
    }
    return launches;
}

int Tree::factorize() {
    try {
        if (ed0.empty()) throw std::runtime_error("assemble() first");
        if (prec != prec_asm)
            throw std::runtime_error("the factor precision changed after assemble(): assemble() again before factorize()");
        if (graphs) {
            if (rank_mode != 0 || rank_log)
                throw std::runtime_error("graph mode needs rank mode 0 (geqp3 ranks) and no rank log");
            if (rect())
                throw std::runtime_error("graph mode is kept for square systems (the row reassignment reads back mid-level)");
            if (gplan.valid && gplan.sig == plan_signature()) {
                const int r = prec_fact == 32 ? replay_t<float>() : replay_t<double>();
                if (r == 0) return 0;
            }
            drop_graphs();
            if (graph_warm_sig != plan_signature()) {
                // one factorization outside any capture grows the arenas; the recorded one then
                // makes the same allocations from the retained chunks and calls no hipMalloc
                factorized = false;
                prec_fact = prec;
                const int rc0 = prec == 32 ? factorize_t<float>() : factorize_t<double>();
                if (rc0) return rc0;
                graph_warm_sig = plan_signature();
            }
            graph_rec = true;
        }
        factorized = false;
        prec_fact = prec;
        const int rc = prec == 32 ? factorize_t<float>() : factorize_t<double>();
        if (graph_rec) {
            graph_rec = false;
            gplan.valid = rc == 0;
            gplan.sig = plan_signature();
        }
        return rc;
    } catch (NumericalFailure& e) {
        last_error = e.what();
        graph_rec = false;
        drop_graphs();
        factorized = false;
        return 1;
    } catch (std::exception& e) {
        last_error = e.what();
        graph_rec = false;
        drop_graphs();
        return -1;
    }
}

// ---------------------------------------------------------------------------- solve
// The solve schedule is built level by level while the factorization runs: level L once
// its factors and ranks are final, at the next level's rank read-back, where the host would
// otherwise wait for the device (the remaining levels at the end).  Workspace: level-0
// clusters at their cstart, higher levels appended level by level; every level's clusters hold
// at most n rows, so n x nlevels entries always suffice and the task pointers never move.
template <typename S>
void Tree::solve_sched_begin() {
    delete sched;
    sched = new SolveSchedule();
    sched->lv.resize(nlevels);
    for (int o : bottoms[0]) cl[o].xoff = (size_t)cl[o].cstart;
    sched->off = (size_t)n;
    // a rectangular system's level-0 segments are placed once its row reassignment is known; a
    // level's clusters then hold at most 2m rows (the extra rows are counted by the interior and
    // by the interface that receives them)
    if (rect()) sched->off = 0;
    const size_t cap = std::max<size_t>(rect() ? 2 * (size_t)m * (size_t)nlevels : (size_t)n * (size_t)nlevels, 1) * sizeof(S);
    // kept across factorizations (of either precision) and only grown: a free would wait for
    // the device
    if (!xw || xw_bytes < cap) {
        if (xw) HIPCHK(hipFree(xw));
        xw = nullptr;
        xw_bytes = 0;
        HIPCHK(hipMalloc(&xw, cap));
        xw_bytes = cap;
    }
    if (rect()) {
        const size_t ns = 2 * (size_t)m + n;
        sched->scratch = arena.alloc(ns * sizeof(S));
        std::vector<long long> io(ns);
        for (size_t i = 0; i < ns; ++i) io[i] = (long long)i;
        sched->iota = upload_vec(io);
    }
}

template <typename S>
void Tree::solve_sched_levels(int upto) {
    S* X0 = (S*)xw;
    for (; sched->built < upto; ++sched->built) {
        const int L = sched->built;
        if (L > 0 || rect())
            for (int o : bottoms[L]) {
                cl[o].xoff = sched->off;
                sched->off += (size_t)cl[o].rows_org + (size_t)cl[o].tot_extra;
            }
        if (sched->off * sizeof(S) > xw_bytes) throw std::runtime_error("solve workspace bound exceeded");
        auto& lv = sched->lv[L];
        if (L > 0) {
            std::vector<VecCopyTask<S>> mf, mb;
            for (int p : bottoms[L]) {
                int k = 0, kc = 0;
                for (int ch : cl[p].children) {
                    mf.push_back({X0 + cl[ch].xoff, X0 + cl[p].xoff + k, cl[ch].rows, 0});
                    mb.push_back({X0 + cl[p].xoff + kc, X0 + cl[ch].xoff, cl[ch].cols, 0});
                    k += cl[ch].rows;
                    kc += cl[ch].cols;
                }
            }
            lv.merge_fwd = upload_vec(mf);
            lv.n_merge_fwd = (int)mf.size();
            lv.merge_bwd = upload_vec(mb);
            lv.n_merge_bwd = (int)mb.size();
        }
        if (rect() && L == 0) {  // loads and stores through the scratch region
            std::vector<long long> a, b, c, d, e, f;
            for (int o : bottoms[0]) {
                const Cluster& C = cl[o];
                for (int i = 0; i < C.rows_org; ++i) { a.push_back(C.rstart + i); b.push_back((long long)C.xoff + i); }
                for (int i = 0; i < C.cols_org; ++i) {
                    c.push_back(C.cstart + i);
                    d.push_back((long long)C.xoff + i);
                    e.push_back((long long)C.xoff + i);
                    f.push_back(C.cstart + i);
                }
            }
            sched->ld_si = upload_vec(a); sched->ld_di = upload_vec(b); sched->n_ld = (int)a.size();
            sched->lc_si = upload_vec(c); sched->lc_di = upload_vec(d); sched->n_lc = (int)c.size();
            sched->st_si = upload_vec(e); sched->st_di = upload_vec(f); sched->n_st = (int)e.size();
        }
        if (rect()) {
            std::vector<long long> rs, rd, sa, sb, ta, tb;
            std::vector<VecQTask<S>> qx;
            for (int nn : interfaces[L]) {
                const Cluster& N = cl[nn];
                long long pos = (long long)N.xoff + N.rows_org;
                for (const auto& t : N.extra_from) {
                    const Cluster& C = cl[t[0]];
                    for (int i = 0; i < t[2]; ++i) {
                        rs.push_back((long long)C.xoff + C.cols + C.row_perm[t[1] + i]);
                        rd.push_back(pos++);
                    }
                }
                if (!N.sparsified) continue;
                if (N.row_rank > 0)
                    qx.push_back({(const S*)N.Qex, (const S*)N.Tex, X0 + N.xoff + N.cols_org, N.nex, N.row_rank, N.nex, 1});
                const long long b0 = (long long)N.xoff;
                const int size = N.rows_org + N.tot_extra, x2 = size - N.cols_org, x1 = N.cols_org - N.rank;
                for (int i = 0; i < x2; ++i) { sa.push_back(b0 + N.cols_org + i); sb.push_back(b0 + N.rank + i); }
                for (int i = 0; i < x1; ++i) { sa.push_back(b0 + N.rank + i); sb.push_back(b0 + N.rank + x2 + i); }
                for (int i = 0; i < x1; ++i) { ta.push_back(b0 + N.rank + i); tb.push_back(b0 + N.rank + x2 + i); }
                for (int i = 0; i < x2; ++i) { ta.push_back(-1); tb.push_back(b0 + N.rank + i); }
            }
            lv.rg_si = upload_vec(rs); lv.rg_di = upload_vec(rd); lv.n_rg = (int)rs.size();
            lv.qex = upload_vec(qx); lv.n_qex = (int)qx.size();
            lv.sp_a = upload_vec(sa); lv.sp_b = upload_vec(sb); lv.n_sp = (int)sa.size();
            lv.ts_a = upload_vec(ta); lv.ts_b = upload_vec(tb); lv.n_ts = (int)ta.size();
        }
        std::vector<VecQTask<S>> q;
        int ncol = 0;
        for (int c : interiors[L]) ncol = std::max(ncol, cl[c].color + 1);
        // Every chain runs on the workgroup chain kernel (row-per-thread when its links fit);
        // the one-wave (k <= SPAQR_WAVE_CHAINS <= 64) and pipelined (k <= SPAQR_PIPE_CHAINS <=
        // 512) chain kernels are opt-in: on the 64^3 solve they are slower.  Read at every plan,
        // so a process can switch kernels between trees.
        const int wave_chain_k = getenv("SPAQR_WAVE_CHAINS") ? std::min(64, atoi(getenv("SPAQR_WAVE_CHAINS"))) : 0;
        const int pipe_chain_k = getenv("SPAQR_PIPE_CHAINS") ? std::min(512, atoi(getenv("SPAQR_PIPE_CHAINS"))) : 0;
        std::vector<std::array<std::vector<VecChainTask<S>>, 4>> ch(ncol);
        std::vector<VecChainLink<S>> links;
        std::vector<VecTrsvTask<S>> tr;
        std::vector<VecGemvLink<S>> tl;
        for (int c : interiors[L]) {
            const Cluster& C = cl[c];
            const Edge& es = ed[C.self_edge];
            q.push_back({(const S*)es.A, (const S*)C.T, X0 + C.xoff, es.rows, es.cols, es.rows, 1});
            VecChainTask<S> t{X0 + C.xoff, C.cols, (int)links.size(), 0, 0};
            int rmin = INT_MAX, rmax = 0;
            for (size_t j = 1; j < C.out.size(); ++j) {
                const Edge& eo = ed[C.out[j]];
                const int ns = link_slices<S>(eo.rows), sz = (eo.rows + ns - 1) / ns;
                for (int sl = 0; sl < ns; ++sl) {
                    const int r0 = sl * sz, rs = std::min(sz, eo.rows - r0);
                    links.push_back({(const S*)eo.A + r0, (const S*)eo.T + (size_t)sl * NB * C.cols, X0 + cl[eo.n2].xoff + r0,
                                     rs, eo.rows});
                    rmin = std::min(rmin, rs);
                    rmax = std::max(rmax, rs);
                }
            }
            t.nlinks = (int)links.size() - t.link0;
            t.rmax = rmax;
            int kind = 3;
            if (C.cols >= 1 && rmin >= 1) {
                if (C.cols <= wave_chain_k && rmax <= 64) kind = 0;
                else if (C.cols <= pipe_chain_k && rmax <= 256) kind = 1;
                else if (C.cols <= pipe_chain_k && rmax <= 512) kind = 2;
            }
            if (t.nlinks > 0) ch[C.color][kind].push_back(t);
            VecTrsvTask<S> tt{(const S*)es.A, X0 + C.xoff, C.cols, es.rows, (int)tl.size(), 0};
            for (int e : C.in) {
                const Edge& ei = ed[e];
                tl.push_back({(const S*)ei.A, X0 + cl[ei.n1].xoff, ei.rows, ei.cols});
            }
            tt.nlinks = (int)tl.size() - tt.link0;
            tr.push_back(tt);
        }
        std::vector<VecSparsTask<S>> sp;
        for (int c : interfaces[L]) {
            const Cluster& C = cl[c];
            if (!C.sparsified) continue;
            sp.push_back({(const S*)C.Qs, (const S*)C.Ts, (const S*)C.Qsp, (const S*)C.Tsp, X0 + C.xoff, C.cols_org,
                          ed[C.self_edge].rows, C.rank, 0});
        }
        if (getenv("SPAQR_DEBUG_SOLVE") && q.size() <= 64) {
            fprintf(stderr, "[solve shapes L%d]", L);
            for (auto& t : q) fprintf(stderr, " q(%d,%d)", t.m, t.k);
            for (int k = 0; k < ncol; ++k)
                for (int c = 0; c < 4; ++c)
                    for (auto& t : ch[k][c]) {
                        fprintf(stderr, " chain%d(k%d:", c, t.k);
                        for (int j = 0; j < t.nlinks; ++j) fprintf(stderr, "%s%d", j ? "," : "", links[t.link0 + j].rows);
                        fprintf(stderr, ")");
                    }
            for (auto& t : sp) fprintf(stderr, " spars(m%d,k%d,r%d)", t.ldq, t.k, t.r);
            for (auto& t : tr) {
                long s = 0;
                for (int j = 0; j < t.nlinks; ++j) s += tl[t.link0 + j].s;
                fprintf(stderr, " trsv(k%d,nl%d,s%ld)", t.k, t.nlinks, s);
            }
            fprintf(stderr, "\n");
        }
        {
            // this level's share of one application: Q^T on the interiors, the chains (V, T of
            // every link and the neighbour's vector segment), the sparsified interfaces forward
            // and backward (Qs, Ts, Qsp, Tsp, and the R part of Qs), and the back substitution
            // (R and the off-diagonal blocks)
            const double SZ = sizeof(S);
            double fl = 0.0, by = 0.0;
            for (const auto& t : q) { fl += 4.0 * t.m * t.k; by += SZ * ((double)t.m * t.k + NB * t.k + 2.0 * t.m); }
            for (const auto& l : links) by += SZ * 2.0 * l.rows;
            for (int c : interiors[L]) {
                const Cluster& C = cl[c];
                for (size_t j = 1; j < C.out.size(); ++j) {
                    const Edge& eo = ed[C.out[j]];
                    fl += 4.0 * eo.rows * C.cols;
                    by += SZ * ((double)eo.rows * C.cols + NB * C.cols);
                }
                for (int e : C.in) {
                    fl += 2.0 * C.cols * ed[e].cols;
                    by += SZ * ((double)C.cols * ed[e].cols + ed[e].cols);
                }
            }
            for (const auto& t : tr) { fl += (double)t.k * t.k; by += SZ * (0.5 * t.k * t.k + 2.0 * t.k); }
            for (const auto& t : sp) {  // forward: Qs^T, Qsp^T; backward: Qsp, R_s^{-1} (upper part of Qs)
                fl += 4.0 * t.ldq * t.k + 8.0 * t.k * t.r + (double)t.k * t.k;
                by += SZ * ((double)t.ldq * t.k + NB * t.k + 2.0 * ((double)t.k * t.r + NB * t.r) + 0.5 * t.k * t.k +
                            4.0 * t.k);
            }
            sched->flops += fl;
            sched->bytes += by;
        }
        lv.q = upload_vec(q);
        lv.n_q = (int)q.size();
        if (getenv("SPAQR_DEBUG_HOST")) {
            int cnt[4] = {0, 0, 0, 0}, kmax[4] = {0, 0, 0, 0};
            for (int q = 0; q < ncol; ++q)
                for (int c = 0; c < 4; ++c)
                    for (auto& t : ch[q][c]) { ++cnt[c]; kmax[c] = std::max(kmax[c], t.k); }
            fprintf(stderr, "[solve L%d] %d colours; chains (count/max k): wave %d/%d pipe1 %d/%d pipe2 %d/%d plain %d/%d\n",
                    L, ncol, cnt[0], kmax[0], cnt[1], kmax[1], cnt[2], kmax[2], cnt[3], kmax[3]);
            for (int q = 0; q < ncol; ++q) {  // per colour: the longest chain in items (links x panels)
                int items = 0, nl = 0, np = 0, rm = 0;
                for (auto& t : ch[q][3]) {
                    const int p = (t.k + NB - 1) / NB;
                    if (t.nlinks * p > items) { items = t.nlinks * p; nl = t.nlinks; np = p; rm = t.rmax; }
                }
                fprintf(stderr, "[solve L%d c%d] longest chain: %d links x %d panels, rmax %d\n", L, q, nl, np, rm);
            }
        }
        lv.chain_links = upload_vec(links);
        for (int k = 0; k < ncol; ++k) {
            lv.chain.emplace_back();
            for (int c = 0; c < 4; ++c) lv.chain.back()[c] = {upload_vec(ch[k][c]), (int)ch[k][c].size()};
        }
        lv.spars = upload_vec(sp);
        lv.n_spars = (int)sp.size();
        lv.trsv = upload_vec(tr);
        lv.n_trsv = (int)tr.size();
        lv.trsv_links = upload_vec(tl);
    }
}

// The transposed backward pass's tasks (solve_nrml only), built on its first use after a
// factorization from the finished structure: merge^T and the transposed couplings per level.
template <typename S>
void Tree::solve_sched_tr() {
    if (sched->tr_built) return;
    S* X0 = (S*)xw;
    for (int L = 0; L < nlevels; ++L) {
        auto& lv = sched->lv[L];
        if (L > 0) {
            std::vector<VecCopyTask<S>> mt;
            for (int p : bottoms[L]) {
                int kc = 0;
                for (int ch : cl[p].children) {
                    mt.push_back({X0 + cl[ch].xoff, X0 + cl[p].xoff + kc, cl[ch].cols, 0});
                    kc += cl[ch].cols;
                }
            }
            lv.merge_tr = upload_vec(mt);
            lv.n_merge_tr = (int)mt.size();
        }
        {  // transposed couplings: per source cluster of an interior's in-edge, links in interior order
            std::map<int, std::pair<int, std::vector<VecGemvtLink<S>>>> by_src;  // source -> (its columns, links)
            for (int c : interiors[L]) {
                const Cluster& C = cl[c];
                for (int e : C.in) {
                    const Edge& ei = ed[e];
                    auto& v = by_src[ei.n1];
                    v.first = ei.cols;
                    v.second.push_back({(const S*)ei.A, X0 + C.xoff, ei.rows, C.cols});
                }
            }
            std::vector<VecGemvtTask<S>> gt;
            std::vector<VecGemvtLink<S>> gl;
            for (auto& kv : by_src) {
                gt.push_back({X0 + cl[kv.first].xoff, kv.second.first, (int)gl.size(), (int)kv.second.second.size(), 0});
                gl.insert(gl.end(), kv.second.second.begin(), kv.second.second.end());
            }
            lv.gemvt = upload_vec(gt);
            lv.n_gemvt = (int)gt.size();
            lv.gemvt_links = upload_vec(gl);
        }
    }
    sched->tr_built = true;
}

// One phase of one level of the preconditioner apply on the solve workspace (ParTree::solve,
// src/ptree.cpp:2798-3203): 0 merge (merge_fwd :2244-2252 / merge_bwd :3105-3123), 1 interior
// QR (fwd_geqrt :2846-2851 and fwd_tsqrt :2160-2186 / bwd_trsv :2188-2192 with the reduction of
// the off-diagonal blocks), 2 sparsification (fwd_spars :2957-2963 / bwd_spars :3146-3174).
template <typename S>
void Tree::solve_phase_t(int L, int phase, int dir) {
    auto& lv = sched->lv[L];
    if (phase == 0) {
        if (L > 0 && dir) launch_vec_copy<S>((const VecCopyTask<S>*)lv.merge_fwd, lv.n_merge_fwd, stream_);
        if (L > 0 && !dir) launch_vec_copy<S>((const VecCopyTask<S>*)lv.merge_bwd, lv.n_merge_bwd, stream_);
    } else if (phase == 1) {
        if (dir) {
            launch_vec_q<S>((const VecQTask<S>*)lv.q, lv.n_q, stream_);
            const auto links = (const VecChainLink<S>*)lv.chain_links;
            for (auto& c : lv.chain) {
                launch_vec_chain_wave<S>((const VecChainTask<S>*)c[0].first, links, c[0].second, stream_);
                launch_vec_chain_pipe<S>((const VecChainTask<S>*)c[1].first, links, c[1].second, 1, stream_);
                launch_vec_chain_pipe<S>((const VecChainTask<S>*)c[2].first, links, c[2].second, 2, stream_);
                launch_vec_chain<S>((const VecChainTask<S>*)c[3].first, links, c[3].second, stream_);
            }
        } else {
            launch_vec_trsv<S>((const VecTrsvTask<S>*)lv.trsv, (const VecGemvLink<S>*)lv.trsv_links, lv.n_trsv, stream_);
        }
    } else if (dir) {
        move_t<S>(lv.rg_si, lv.rg_di, lv.n_rg, false);  // fwd_reassign + merge_x (src/ptree.cpp:2395-2431)
        launch_vec_spars<S>((const VecSparsTask<S>*)lv.spars, lv.n_spars, 1, stream_);
        launch_vec_q<S>((const VecQTask<S>*)lv.qex, lv.n_qex, stream_);  // orthogonalD_extra_fwd :2206-2211
        move_t<S>(lv.sp_a, lv.sp_b, lv.n_sp, true);  // splitD_fwd :2223-2232
    } else {
        move_t<S>(lv.sp_b, lv.sp_a, lv.n_sp, true);  // splitD_bwd :2234-2242
        launch_vec_spars<S>((const VecSparsTask<S>*)lv.spars, lv.n_spars, 0, stream_);
    }
}

// Workspace moves X[di[q]] = X[si[q]] (si < 0: zero), q < n; through the scratch region when
// the positions of one call overlap.
template <typename S>
void Tree::move_t(const void* si, const void* di, int cnt, bool staged) {
    if (cnt <= 0) return;
    S* X0 = (S*)xw;
    const auto a = (const long long*)si;
    const auto b = (const long long*)di;
    if (!staged) {
        launch_vec_gather<S>(X0, X0, a, b, cnt, stream_);
        return;
    }
    S* T = (S*)sched->scratch;
    const auto io = (const long long*)sched->iota;
    launch_vec_gather<S>(X0, T, a, io, cnt, stream_);
    launch_vec_gather<S>(T, X0, io, b, cnt, stream_);
}

// One level of the transposed backward pass, the first half of solve_nrml (the transposes of
// merge_bwd, bwd_trsv with its couplings, and bwd_spars, in reverse order; src/toperations.cpp
// tMerge::fwd :88-96, tQR::fwd :12-27, tScaleD::fwd :71-74, tOrthogonalD::fwd :79-81, tSplitD).
template <typename S>
void Tree::solve_phase_tr(int L) {
    auto& lv = sched->lv[L];
    if (L > 0) launch_vec_copy<S>((const VecCopyTask<S>*)lv.merge_tr, lv.n_merge_tr, stream_);
    launch_vec_trsv_t<S>((const VecTrsvTask<S>*)lv.trsv, lv.n_trsv, stream_);
    launch_vec_gemvt<S>((const VecGemvtTask<S>*)lv.gemvt, (const VecGemvtLink<S>*)lv.gemvt_links, lv.n_gemvt, stream_);
    launch_vec_spars_t<S>((const VecSparsTask<S>*)lv.spars, lv.n_spars, stream_);
    move_t<S>(lv.ts_a, lv.ts_b, lv.n_ts, true);
}

// b (m entries, permuted rows) into the level-0 segments, and the segments' first cols_org
// entries out to x (n, permuted columns); a square system's level-0 segments are x itself.
template <typename S>
void Tree::load_rows_t(const double* b) {
    S* X0 = (S*)xw;
    if (!rect()) {
        launch_convert_d2s<S>(n, b, X0, stream_);
        return;
    }
    S* T = (S*)sched->scratch;
    launch_convert_d2s<S>(m, b, T, stream_);
    launch_vec_gather<S>(T, X0, (const long long*)sched->ld_si, (const long long*)sched->ld_di, sched->n_ld, stream_);
}
template <typename S>
void Tree::load_cols_t(const double* y) {
    S* X0 = (S*)xw;
    if (!rect()) {
        launch_convert_d2s<S>(n, y, X0, stream_);
        return;
    }
    S* T = (S*)sched->scratch;
    launch_convert_d2s<S>(n, y, T, stream_);
    launch_vec_gather<S>(T, X0, (const long long*)sched->lc_si, (const long long*)sched->lc_di, sched->n_lc, stream_);
}
template <typename S>
void Tree::store_cols_t(double* x) {
    S* X0 = (S*)xw;
    if (!rect()) {
        launch_convert_s2d<S>(n, X0, x, stream_);
        return;
    }
    S* T = (S*)sched->scratch;
    launch_vec_gather<S>(X0, T, (const long long*)sched->st_si, (const long long*)sched->st_di, sched->n_st, stream_);
    launch_convert_s2d<S>(n, T, x, stream_);
}

template <typename S>
int Tree::solve_ls_t(const double* b, double* x) {
    load_rows_t<S>(b);
    for (int L = 0; L < nlevels; ++L)
        for (int ph = 0; ph < 3; ++ph) solve_phase_t<S>(L, ph, 1);
    for (int L = nlevels - 1; L >= 0; --L)
        for (int ph = 2; ph >= 0; --ph) solve_phase_t<S>(L, ph, 0);
    store_cols_t<S>(x);
    HIPCHK(hipGetLastError());
    return 0;
}

template <typename S>
int Tree::solve_nrml_t(const double* y, double* x) {
    solve_sched_tr<S>();
    load_cols_t<S>(y);
    for (int L = 0; L < nlevels; ++L) solve_phase_tr<S>(L);
    for (int L = nlevels - 1; L >= 0; --L)
        for (int ph = 2; ph >= 0; --ph) solve_phase_t<S>(L, ph, 0);
    store_cols_t<S>(x);
    HIPCHK(hipGetLastError());
    return 0;
}

int Tree::solve_ls_dev(const double* b, double* x) {
    try {
        if (!factorized || !sched) throw std::runtime_error("factorize() first");
        int rc = 0;
        timed(P_SOLVE, 1, 1, [&] { rc = (prec_fact == 32) ? solve_ls_t<float>(b, x) : solve_ls_t<double>(b, x); });
        prof_add(P_SOLVE, sched->flops, sched->bytes + 8.0 * (n + m));
        return rc;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

int Tree::solve_nrml_dev(const double* y, double* x) {
    try {
        if (!factorized || !sched) throw std::runtime_error("factorize() first");
        int rc = 0;
        timed(P_SOLVE, 1, 1, [&] { rc = (prec_fact == 32) ? solve_nrml_t<float>(y, x) : solve_nrml_t<double>(y, x); });
        prof_add(P_SOLVE, sched->flops, sched->bytes + 16.0 * n);
        return rc;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

// cgls (src/is.cpp:274-345) with x0 = 0 on device vectors: CG on the normal equations with
// solve_nrml as the preconditioner, stopping when ||A^T (b - A x)||^2 < tol^2 ||A^T b||^2.  The
// scalars (five dot products per iteration) come back to the host, as the reference's do.
int Tree::cgls_dev(const double* b, double* x, int iters, double tol_, int* iters_out, double* rel) {
    try {
        if (!factorized) throw std::runtime_error("factorize() first");
        if (!d_tptr) throw std::runtime_error("cgls needs the adjoint product (a rectangular system)");
        double* W = krylov_ws(sizeof(double) * (2 * (size_t)m + 3 * (size_t)n + 4));
        double* r = W;
        double* q = r + m;
        double* sv = q + m;
        double* p = sv + n;
        double* z = p + n;
        double* part = d_scratch;
        double* hdev = d_scratch + dot_blocks;
        auto dot = [&](int len, const double* a, const double* c) {
            launch_dot(len, a, c, part, dot_blocks, hdev, 0, stream_);
            double h = 0.0;
            HIPCHK(hipMemcpyAsync(&h, hdev, sizeof(double), hipMemcpyDeviceToHost, stream_));
            sync_stream();
            return h;
        };
        launch_fill<double>(x, (size_t)n, 0.0, stream_);
        HIPCHK(hipMemcpyAsync(r, b, sizeof(double) * m, hipMemcpyDeviceToDevice, stream_));
        if (spmv_t_dev(b, sv) != 0) throw std::runtime_error(last_error);
        const double rhs2 = dot(n, sv, sv);
        *iters_out = 0;
        if (rhs2 == 0.0) { *rel = 0.0; return 0; }
        const double threshold = tol_ * tol_ * rhs2;
        double res2 = rhs2;
        if (res2 < threshold) { *rel = std::sqrt(res2 / rhs2); return 0; }
        if (solve_nrml_dev(sv, p) != 0) throw std::runtime_error(last_error);
        double abs_new = dot(n, sv, p);
        int i = 0;
        while (i < iters) {
            if (spmv_dev(p, q) != 0) throw std::runtime_error(last_error);
            const double qq = dot(m, q, q);
            if (!std::isfinite(qq) || qq == 0.0) throw std::runtime_error("cgls: A p vanished or is not finite");
            const double alpha = abs_new / qq;
            launch_axpby(n, alpha, p, 1.0, x, stream_);
            launch_axpby(m, -alpha, q, 1.0, r, stream_);
            if (spmv_t_dev(r, sv) != 0) throw std::runtime_error(last_error);
            res2 = dot(n, sv, sv);
            if (res2 < threshold) break;
            if (solve_nrml_dev(sv, z) != 0) throw std::runtime_error(last_error);
            const double abs_old = abs_new;
            abs_new = dot(n, sv, z);
            const double beta = abs_new / abs_old;
            launch_axpby(n, 1.0, z, beta, p, stream_);
            ++i;
        }
        HIPCHK(hipGetLastError());
        sync_stream();
        *iters_out = i;
        *rel = std::sqrt(res2 / rhs2);
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

template <typename S>
int Tree::solve_t(double* x) {
    S* X0 = (S*)xw;
    launch_convert_d2s<S>(n, x, X0, stream_);
    for (int L = 0; L < nlevels; ++L)
        for (int ph = 0; ph < 3; ++ph) solve_phase_t<S>(L, ph, 1);
    for (int L = nlevels - 1; L >= 0; --L)
        for (int ph = 2; ph >= 0; --ph) solve_phase_t<S>(L, ph, 0);
    launch_convert_s2d<S>(n, X0, x, stream_);
    HIPCHK(hipGetLastError());
    return 0;
}

int Tree::solve_load(const double* x) {
    try {
        if (!factorized || !sched) throw std::runtime_error("factorize() first");
        if (rect()) throw std::runtime_error("the phase-by-phase solve is kept for square systems");
        if (prec_fact == 32) launch_convert_d2s<float>(n, x, (float*)xw, stream_);
        else launch_convert_d2s<double>(n, x, (double*)xw, stream_);
        HIPCHK(hipGetLastError());
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

int Tree::solve_phase(int level, int phase, int dir) {
    try {
        if (!factorized || !sched) throw std::runtime_error("factorize() first");
        if (level < 0 || level >= nlevels || phase < 0 || phase > 2) throw std::runtime_error("no such solve phase");
        if (prec_fact == 32) solve_phase_t<float>(level, phase, dir);
        else solve_phase_t<double>(level, phase, dir);
        HIPCHK(hipGetLastError());
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

int Tree::solve_store(double* x) {
    try {
        if (!factorized || !sched) throw std::runtime_error("factorize() first");
        if (prec_fact == 32) launch_convert_s2d<float>(n, (const float*)xw, x, stream_);
        else launch_convert_s2d<double>(n, (const double*)xw, x, stream_);
        HIPCHK(hipGetLastError());
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

int Tree::solve_dev(double* x) {
    try {
        if (!factorized || !sched) throw std::runtime_error("factorize() first");
        if (rect()) throw std::runtime_error("a rectangular system's solve maps m rows to n columns: use solve_ls");
        int rc = 0;
        timed(P_SOLVE, 1, 1, [&] { rc = (prec_fact == 32) ? solve_t<float>(x) : solve_t<double>(x); });
        // plus the FP64 vector in and out
        prof_add(P_SOLVE, sched->flops, sched->bytes + 16.0 * n);
        return rc;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

int Tree::spmv_dev(const double* x, double* y) {
    try {
        timed(P_SPMV, 1, 1, [&] { launch_spmv(m, d_rowptr, d_col, d_val, x, y, stream_); });
        prof_add(P_SPMV, 2.0 * d_nnz, 12.0 * d_nnz + 4.0 * (m + 1) + 8.0 * (n + m));
        HIPCHK(hipGetLastError());
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

int Tree::spmv_t_dev(const double* r, double* y) {
    try {
        if (!d_tptr) throw std::runtime_error("the adjoint product is kept for rectangular systems");
        timed(P_SPMV, 1, 1, [&] { launch_spmv(n, d_tptr, d_tcol, d_tval, r, y, stream_); });
        prof_add(P_SPMV, 2.0 * d_nnz, 12.0 * d_nnz + 4.0 * (n + 1) + 8.0 * (n + m));
        HIPCHK(hipGetLastError());
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

int Tree::diag_R(int order, double* out, int maxn) {
    try {
        if (!factorized) throw std::runtime_error("factorize() first");
        const Cluster& C = cl.at(order);
        if (C.self_edge < 0 || C.lvl() != C.merge_level) return -1;
        const Edge& e = ed[C.self_edge];
        const int k = e.cols;
        if ((long long)k * k > maxn) return -2;
        std::vector<double> h((size_t)e.rows * e.cols);
        if (prec_fact == 32) {
            std::vector<float> f(h.size());
            HIPCHK(hipMemcpy(f.data(), e.A, f.size() * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < f.size(); ++i) h[i] = f[i];
        } else {
            HIPCHK(hipMemcpy(h.data(), e.A, h.size() * 8, hipMemcpyDeviceToHost));
        }
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < k; ++i) out[i + (size_t)j * k] = (i <= j) ? h[i + (size_t)j * e.rows] : 0.0;
        return k;
    } catch (std::exception& e) {
        last_error = e.what();
        return -3;
    }
}

// ---------------------------------------------------------------------------- GMRES
double* Tree::krylov_ws(size_t bytes) {
    if (!d_krylov || krylov_bytes < bytes) {
        sync_stream();  // queued work may still read the old basis
        if (d_krylov) HIPCHK(hipFree(d_krylov));
        d_krylov = nullptr;
        krylov_bytes = 0;
        HIPCHK(hipMalloc(&d_krylov, bytes));
        krylov_bytes = bytes;
    }
    return d_krylov;
}

// dist_gmres (src/is.cpp:447-527) on device vectors.  Krylov mode 0 is the reference's
// scheme: MGS Arnoldi with one dot + one axpy launch per basis vector (scalars stay on the
// device), Givens rotations on the host from one small readback per iteration.  Mode 1 (the
// default): classical Gram-Schmidt applied twice in fused passes (eight launches per
// iteration whatever the basis size), Givens rotations, the triangular solve and the update
// of x on the device; the host reads back one residual estimate per iteration.  restart > 0 restarts from the current iterate (an
// extension; the reference runs one cycle) until ||b - A x|| / ||b|| <= true_tol.
int Tree::gmres_dev(const double* b, double* x, int iters, double tol_, int restart, double true_tol,
                    GmresResult* res, double* hist, int hist_cap, int* hist_len, bool flexible) {
    // Left-preconditioned (flexible == false): dist_gmres, src/is.cpp:447-527 -- Arnoldi on
    // M^{-1}A with modified Gram-Schmidt and Givens rotations, x0 = 0, stop when the
    // preconditioned residual estimate |beta_{k+1}| drops to tol * ||M^{-1} b||.
    // Flexible (right-preconditioned) variant: z_k = M^{-1} q_k is kept, w = A z_k is
    // orthogonalised, x = x0 + Z y, and |beta_{k+1}| estimates the TRUE residual ||b - A x||, so a
    // preconditioner whose rounding makes it slightly non-linear (FP32 factors) does not stall
    // the stopping test; it stops at |beta_{k+1}| <= tol * ||b||.
    // Restarts (restart > 0) continue from the current iterate with the same absolute target.
    try {
        if (!factorized) throw std::runtime_error("factorize() first");
        if (rect()) throw std::runtime_error("GMRES needs a square system; a least-squares system is solved by cgls");
        const int m = (restart > 0) ? std::min(restart, iters) : iters;
        if (m + 2 > 4096) throw std::runtime_error("restart length too large (at most 4094 steps per cycle)");
        const size_t nv = (size_t)(m + 1) + (flexible ? (size_t)m : 0) + 1;  // basis, Z, work vector
        // fused mode: H, cs, sn, beta, hcol, c, y, est, the multi-dot partials and norm chunks
        const bool fused = krylov_mode == 1;
        const int nb = mdot_blocks(n);
        const size_t ldh = (size_t)m + 1;
        // H (ldh x m), cs, sn (m), beta (ldh), hcol, c (ldh + 1 each), y (ldh), est (1), partials
        // (ldh x nb), norm chunks (nb)
        const size_t small = fused ? ldh * m + 2 * (size_t)m + 4 * ldh + 3 + ldh * nb + nb : 0;
        double* Q = krylov_ws(sizeof(double) * ((size_t)n * nv + small));
        double* d_H = Q + (size_t)n * nv;
        double* d_cs = d_H + ldh * m;
        double* d_sn = d_cs + m;
        double* d_beta = d_sn + m;
        double* d_hcol = d_beta + ldh;
        double* d_c = d_hcol + ldh + 1;
        double* d_y = d_c + ldh + 1;
        double* d_est = d_y + ldh;
        double* d_part = d_est + 1;
        double* d_normpart = d_part + ldh * nb;
        double* Z = flexible ? Q + (size_t)n * (m + 1) : nullptr;
        double* w = Q + (size_t)n * (nv - 1);
        double* part = d_scratch;
        double* hdev = d_scratch + dot_blocks;  // up to 4096 scalars
        auto dot = [&](const double* a, const double* bb, double* out, int sq) {
            launch_dot(n, a, bb, part, dot_blocks, out, sq, stream_);
        };
        int nh = 0;
        launch_fill<double>(x, (size_t)n, 0.0, stream_);
        double bnorm = 0.0;
        dot(b, b, hdev, 1);
        HIPCHK(hipMemcpyAsync(&bnorm, hdev, sizeof(double), hipMemcpyDeviceToHost, stream_));
        sync_stream();
        int total = 0, cycles = 0;
        double last_err = 1.0, true_rel = 1.0, ref_norm = -1.0;
        bool first = true;
        while (true) {
            cycles++;
            // r = b - A x (left: r = M^{-1}(b - A x))
            if (first) {
                HIPCHK(hipMemcpyAsync(Q, b, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
            } else {
                spmv_dev(x, w);
                HIPCHK(hipMemcpyAsync(Q, b, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
                launch_axpby(n, -1.0, w, 1.0, Q, stream_);
            }
            if (!flexible && solve_dev(Q) != 0) throw std::runtime_error(last_error);
            dot(Q, Q, hdev, 1);
            double r_norm = 0.0;
            HIPCHK(hipMemcpyAsync(&r_norm, hdev, sizeof(double), hipMemcpyDeviceToHost, stream_));
            sync_stream();
            if (!(r_norm == r_norm)) throw std::runtime_error("GMRES: NaN residual");
            if (r_norm == 0.0) { true_rel = 0.0; break; }
            if (ref_norm < 0) ref_norm = flexible ? bnorm : r_norm;
            launch_scal_dev(n, hdev, Q, Q, stream_);
            if (fused) HIPCHK(hipMemcpyAsync(d_beta, hdev, sizeof(double), hipMemcpyDeviceToDevice, stream_));
            std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), beta(m + 1, 0.0);
            auto h = [&](int i, int j) -> double& { return H[i + (size_t)j * (m + 1)]; };
            beta[0] = r_norm;
            int k = 0;
            std::vector<double> hcol(m + 2);
            for (; k < m; ++k) {
                double* qk = Q + (size_t)k * n;
                double* qn = Q + (size_t)(k + 1) * n;
                if (flexible) {
                    double* zk = Z + (size_t)k * n;
                    HIPCHK(hipMemcpyAsync(zk, qk, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
                    if (solve_dev(zk) != 0) throw std::runtime_error(last_error);
                    spmv_dev(zk, qn);
                } else {
                    spmv_dev(qk, qn);
                    if (solve_dev(qn) != 0) throw std::runtime_error(last_error);
                }
                if (fused) {
                    // CGS2 in fused passes, the Givens step on the device; one scalar comes back
                    timed(P_KRYLOV, 8, 1, [&] {
                        launch_cgs2(n, Q, k + 1, qn, d_part, d_c, d_hcol, d_normpart, stream_);
                        launch_givens_step(k, nb, d_normpart, d_hcol, d_H, (int)ldh, d_cs, d_sn, d_beta, ref_norm, d_est,
                                           stream_);
                        launch_scal_dev(n, d_hcol + k + 1, qn, qn, stream_);
                    });
                    prof_add(P_KRYLOV, 8.0 * n * (k + 1) + 3.0 * n, 8.0 * n * (4.0 * (k + 1) + 6.0));
                    HIPCHK(hipMemcpyAsync(&last_err, d_est, sizeof(double), hipMemcpyDeviceToHost, stream_));
                    sync_stream();
                    if (!(last_err == last_err)) throw std::runtime_error("GMRES: NaN residual");
                    if (hist && nh < hist_cap) hist[nh++] = last_err;
                    if (last_err <= tol_) break;
                    continue;
                }
                timed(P_KRYLOV, 2L * (k + 1) * 2 + 3, 1, [&] {
                    for (int i = 0; i <= k; ++i) {
                        dot(qn, Q + (size_t)i * n, hdev + i, 0);
                        launch_axpy_dev(n, hdev + i, Q + (size_t)i * n, qn, stream_);
                    }
                    dot(qn, qn, hdev + k + 1, 1);
                    launch_scal_dev(n, hdev + k + 1, qn, qn, stream_);
                });
                prof_add(P_KRYLOV, 4.0 * n * (k + 1) + 3.0 * n, 8.0 * n * (2.0 * (k + 1) + 3.0) + 8.0 * n * (k + 1));
                HIPCHK(hipMemcpyAsync(hcol.data(), hdev, sizeof(double) * (k + 2), hipMemcpyDeviceToHost, stream_));
                sync_stream();
                for (int i = 0; i <= k + 1; ++i) h(i, k) = hcol[i];
                for (int i = 0; i < k; ++i) {
                    double tmp = cs[i] * h(i, k) + sn[i] * h(i + 1, k);
                    h(i + 1, k) = -sn[i] * h(i, k) + cs[i] * h(i + 1, k);
                    h(i, k) = tmp;
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
                last_err = std::fabs(beta[k + 1]) / ref_norm;
                if (hist && nh < hist_cap) hist[nh++] = last_err;
                if (last_err <= tol_) break;
            }
            k = std::min(k, m - 1);
            const double* basis = flexible ? Z : Q;
            if (fused) {
                launch_lsq_update(n, k + 1, d_H, (int)ldh, d_beta, d_y, basis, x, stream_);
            } else {
                std::vector<double> y(k + 1, 0.0);
                for (int i = k; i >= 0; --i) {
                    y[i] = beta[i];
                    for (int j = i + 1; j <= k; ++j) y[i] -= h(i, j) * y[j];
                    y[i] /= h(i, i);
                }
                for (int j = 0; j <= k; ++j) launch_axpby(n, y[j], basis + (size_t)j * n, 1.0, x, stream_);
            }
            total += k + 1;
            first = false;
            // true residual
            spmv_dev(x, w);
            launch_axpby(n, 1.0, b, -1.0, w, stream_);
            dot(w, w, hdev, 1);
            double rn = 0.0;
            HIPCHK(hipMemcpyAsync(&rn, hdev, sizeof(double), hipMemcpyDeviceToHost, stream_));
            sync_stream();
            true_rel = rn / bnorm;
            if (restart <= 0) break;
            if (true_rel <= true_tol || total >= iters) break;
        }
        sync_stream();
        if (res) {
            res->iters = total - 1;
            res->cycles = cycles;
            res->prec_res = last_err;
            res->true_res = true_rel;
        }
        if (hist_len) *hist_len = nh;
        return 0;
    } catch (std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

}  // namespace spaqr
