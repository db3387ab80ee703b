// Level planner for the MI355X spaQR path: the host side of ParTree (include/ptree.h:56-161
// in the reference) restated for one GPU.  The host keeps the symbolic structure (ordering,
// cluster hierarchy, fill-in, merge) in flat arrays; every numeric operation of a tree
// level is issued as a few grouped batched launches over device descriptor tables, and the
// factors stay resident in HBM for the preconditioner solve and GMRES.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <array>
#include <vector>

#include "kernels.hpp"

#include <memory>

namespace spaqr {

class SymbolicDevice;

struct SepID {
    int lvl = -1, sep = 0;
    SepID() {}
    SepID(int l, int s) : lvl(l), sep(s) {}
    bool operator==(const SepID& o) const { return lvl == o.lvl && sep == o.sep; }
    bool operator!=(const SepID& o) const { return !(*this == o); }
    bool operator<(const SepID& o) const { return lvl < o.lvl || (lvl == o.lvl && sep < o.sep); }
};

struct ClusterID {
    SepID self, l, r;
    int part = 2, section = 0;
    bool operator==(const ClusterID& o) const {
        return self == o.self && l == o.l && r == o.r && part == o.part && section == o.section;
    }
    bool operator!=(const ClusterID& o) const { return !(*this == o); }
    bool operator<(const ClusterID& o) const;
};

struct HostCSC {
    int nr = 0, nc = 0;
    std::vector<int> colptr, rowind;
    std::vector<double> val;
};

struct Cluster {
    int order = -1, merge_level = 0;
    int inst = 0;                         // system of a batch this cluster belongs to
    ClusterID id;
    int cstart = 0, rstart = 0;
    int rows = 0, cols = 0, rows_org = 0, cols_org = 0, rank = 0;
    // rectangular systems: rows kept beyond the columns (rows - cols; after the extra-row
    // sparsification its rank), an interior's extra-row targets and their order (row_perm), an
    // interface's received rows {interior, start in its row_perm, count} in interior order,
    // and the extra-row sparsification's reflectors (nex rows)
    int row_rank = 0, tot_extra = 0, nex = 0;
    std::vector<int> r2c, row_perm;
    std::vector<std::array<int, 3>> extra_from;
    void* Qex = nullptr;
    void* Tex = nullptr;
    void* tauex = nullptr;
    int parent = -1;
    std::vector<int> children;
    int rpos = -1, cpos = -1;
    bool eliminated = false;
    bool sparsified = false;
    std::vector<int> out, in, sout, sin;  // edge ids
    int self_edge = -1;
    std::vector<int> dist2;               // sorted cluster orders
    // device factors (type-erased; element type is the tree's precision)
    void* T = nullptr;                    // NB x cols, diagonal geqrt (interiors)
    void* Qs = nullptr;                   // rows x cols, scaling QR (sparsified interfaces)
    void* Ts = nullptr;                   // NB x cols
    void* Qsp = nullptr;                  // cols x rank, sparsification reflectors
    void* Tsp = nullptr;                  // NB x cols
    void* tausp = nullptr;
    int color = 0;
    size_t xoff = 0;                      // offset of this cluster's segment in the solve workspace
    int lvl() const { return id.self.lvl; }
};

struct Edge {
    int n1, n2;        // block A(n2, n1): rows(n2) x cols(n1)
    int rows, cols;
    void* A = nullptr; // device, column-major, ld = rows
    void* T = nullptr; // tpqrt T (NB x cols(n1)) when n1 is an interior and n2 != n1
    bool spars = false;
    bool alive = true;
};

struct LevelStats {
    int level = 0;
    int n_interiors = 0, n_interfaces = 0, n_sparsified = 0;
    int max_front = 0;
    int colors = 0;
    long long cols_before = 0, ranks_after = 0;
    double flops = 0.0;
};

// Kernel classes timed with HIP events on the tree's stream when profiling is on.
enum ProfCat {
    P_GEQRT = 0, P_LARFB, P_TPQRT, P_TPMQRT, P_SGEQRT, P_SAPPLY, P_GATHER, P_GEQP3, P_SPARS, P_FILL,
    P_SOLVE, P_SPMV, P_KRYLOV, P_RSKETCH, P_NCAT
};
struct ProfEntry {
    double ms = 0, flops = 0, bytes = 0;
    long launches = 0, tasks = 0;
};

// Planner tunables: tile, chunk and panel-path widths picked per workload by the autotuner.
enum TuneParam {
    T_BIG_FRONT = 0,    // min(m, n) from which a front's panels run as one launch per panel (96)
    T_TPMQRT_TILE,      // column tile of a trailing-update chain task (64)
    T_LARFB_COLS,       // column chunk of a Q^T C task (128; at most 64 for FP32 blocks over 512 rows)
    T_TRSM_ROWS,        // row chunk of a C R^{-1} task (128 FP32, 256 FP64)
    T_SPARS_ROWS,       // row chunk of the sparsify right transform (128 FP32, 256 FP64)
    T_SPARS_COLS,       // column chunk of the sparsify left transform and scatter (128)
    T_RANK_COLS,        // column chunk of the rank path's column-norm and tail kernels (128)
    T_SKETCH_UNPIVOTED, // sketch width from which the sketch QR runs unpivoted (96)
    T_TPQRT_COOP,       // 1: multi-panel coupling chains on the cooperative kernel, 2: off (1)
    T_SMALL_TASKS,      // 1: blocks of <= 64 x 64 on the one-wave kernels, 2: off (1)
    T_SKETCH_BF16,      // FP32 fronts: 1 sketches C Omega in bf16 with FP32 accumulation, 2 in FP32 (1)
    T_SKETCH_TILE,      // output-tile columns of the bf16 sketch: 64, 128, 256 (128)
    T_SKETCH_GROW,      // learned sketch width = grow% x predicted rank + pad: 100 .. 200 (125)
    T_SKETCH_PAD,       // the pad + 1 (0 restores the built-in 8): 1 .. 65
    T_PANEL_COLS,       // column chunk of the trailing update in a big front's panel steps (128)
    T_NPARAM
};

struct GmresResult {
    int iters = 0;        // 0-based k of the last cycle plus earlier cycles' steps (as dist_gmres counts)
    int cycles = 0;
    double prec_res = 0;  // last preconditioned relative residual estimate
    double true_res = 0;  // ||b - A x|| / ||b|| in FP64
};

class Tree {
public:
    Tree(int nlevels, int skip);
    ~Tree();

    // configuration (include/tree.h:137-143; include/ptree.h:123-125)
    double tol = 0.1;
    int nlevels, skip;
    int scale = 1, square = 1, verbose = 0;  // square: 1 for m == n, 0 for a least-squares A (m > n)
    float order = 1.0f;
    int prec = 64;  // 64: FP64 factors; 32: FP32 factors (GMRES stays FP64)
    std::vector<double> X;  // coordinates, d x n
    int xdim = 0;
    std::vector<int> user_parts;  // leaf parts given in place of coordinates (the METIS path)
    // The per-matrix symbolic phase (ordering and assembly) on the device: square systems on the
    // geometric path (coordinates).  The same permutation, clusters, blocks and SpMV storage as
    // the host phase; other systems use the host phase.
    bool device_symbolic = false;
    double t_partition_ms = 0.0, t_assemble_ms = 0.0;  // the last partition() / assemble()
    std::vector<int> leaf_parts;  // the leaf parts of the last partition, natural order

    // A may have more rows than columns (m > n, a least-squares system): it is then ordered
    // without coordinates (the caller's leaf parts or the graph bisection) and its rows are
    // filed under the clusters by form_rmap.
    int partition(const HostCSC& A);
    int n_rows() const { return m; }
    bool rect() const { return m != n; }
    const std::vector<int>& rperm() const { return rperm_; }
    // A batch of independent square systems factored together: A is their block-diagonal
    // matrix (system s owns unknowns [off_s, off_s + n_s), in order), X the coordinates of all
    // unknowns; system s gets its own nested dissection with levels[s] levels.  Every tree
    // level's phases then run as one set of grouped launches over all systems.
    int partition_batch(const HostCSC& A, const std::vector<int>& sizes, const std::vector<int>& levels);
    int n_systems() const { return (int)inst_n.size(); }
    int system_offset(int s) const { return inst_off[s]; }
    int system_size(int s) const { return inst_n[s]; }
    // GMRES on every system of the batch at once (one Krylov space per system, shared SpMV and
    // preconditioner launches, per-system reductions), each with the stopping rules of
    // gmres_dev; res gets one entry per system.
    int gmres_batch_dev(const double* b, double* x, int iters, double tol, int restart, double true_tol, bool flexible,
                        std::vector<GmresResult>& res);
    int assemble();
    int factorize();
    int solve_dev(double* x);                       // preconditioner apply, in place, device vector
    // The preconditioner apply in pieces: load x into the solve workspace, run one phase of one
    // level forward (dir 1) or backward (dir 0) -- phase 0 merge, 1 interior QR (forward: Q^T
    // and the coupling chains; backward: the triangular solve), 2 sparsification -- and store
    // the workspace back.  Forward over the levels and phases in order, then backward in
    // reverse, is solve_dev.
    int solve_load(const double* x);
    int solve_phase(int level, int phase, int dir);
    int solve_store(double* x);
    int spmv_dev(const double* x, double* y);       // y = A_perm x, device vectors (y: m entries)
    int spmv_t_dev(const double* r, double* y);     // y = A_perm^T r (r: m entries, y: n)
    // ParTree::solve(b, x) (src/ptree.cpp:2267): b (m, permuted rows) -> x (n, permuted columns)
    int solve_ls_dev(const double* b, double* x);
    // x = M^{-1} M^{-T} y, M^{-1} the backward half of the solve: the preconditioner of cgls
    int solve_nrml_dev(const double* y, double* x);
    // cgls (src/is.cpp:274-345) on App x = b in the least-squares sense, x0 = 0; returns the
    // iteration count as the reference counts it; *rel = ||A^T r|| / ||A^T b||.
    int cgls_dev(const double* b, double* x, int iters, double tol, int* iters_out, double* rel);
    int gmres_dev(const double* b, double* x, int iters, double tol, int restart, double true_tol,
                  GmresResult* res, double* hist, int hist_cap, int* hist_len, bool flexible = false);
    int n_unknowns() const { return n; }
    const std::vector<int>& cperm() const { return cperm_; }
    int num_clusters() const { return (int)cl.size(); }
    const Cluster& cluster(int o) const { return cl[o]; }
    const std::vector<LevelStats>& level_stats() const { return stats; }
    double factor_flops() const;
    int diag_R(int order, double* out, int maxn);  // upper triangle of A(c,c) after factorize
    hipStream_t stream() const { return stream_; }
    // Sparsification rank path.  0: geqp3 on the gathered C (the reference method); 1: pivoted
    // QR of a Gaussian sketch whose width comes from the rank predictor, judged by the tail
    // test, geqp3 fallback; 2: the same with the sketch as wide as min(k, w) (no predictor).
    int rank_mode = 0;
    bool rank_log = false;     // record (features, rank) of every sparsified interface
    int set_rank_model(const float* w);  // kMlpParams predictor weights (host)
    struct RankLog {
        std::vector<float> x;  // kRankFeatures per entry
        std::vector<int> rank, k, w, level, mode, order;  // order: the interface's cluster
    } rlog;
    // Sketch rounds per interface in rank modes 1 and 2: the first sketch plus up to
    // rank_rounds - 1 retries with a doubled width before the geqp3 fallback (1..3).
    int rank_rounds = 3;
    // Pivot check of every interior's R after its elimination: non-finite diagonal entries and
    // pivots at or below pivot_rel * max |R_ii| (0: exactly zero) end the factorization with a
    // numerical failure (factorize returns 1).  get_info: flags per cluster (status.hpp).
    double pivot_rel = 0.0;
    int get_info(std::vector<int>& out);
    // Planner tunables (TuneParam, spaqr_tree_set_tuning); 0 keeps the built-in value.
    int tune[16] = {0};
    int krylov_mode = 1;  // 0: MGS + host Givens (dist_gmres's scheme), 1: fused CGS2 + device Givens
    // hipGraph mode: the first factorization records every stretch of launches between two host
    // synchronisations (the rank read-backs) as a graph and launches it; a later factorization
    // of the same assembled matrix with the same settings replays the graphs without planning,
    // reading the ranks back at the same points, and plans afresh if any rank differs.  Needs
    // rank mode 0 (geqp3 ranks).
    int set_graphs(bool on);
    bool graphs_on() const { return graphs; }
    int graph_segments() const { return (int)gplan.exec.size(); }
    long graph_replays = 0;
    int tuned(int p, int dflt) const { return tune[p] > 0 ? tune[p] : dflt; }
    struct RankStats {
        long tasks = 0, sketched = 0, fallback = 0, direct = 0, retried = 0;
        double sketch_cols = 0, full_cols = 0, rank_sum = 0;
    } rstats;
    bool profiling = false;
    ProfEntry prof[P_NCAT];
    void prof_reset();
    int sync();  // waits for the stream and folds finished event timings into prof
    std::string last_error;
    double t_factor_ms = 0.0;

private:
    int n = 0, m = 0;
    HostCSC A0, App;
    std::vector<int> cperm_, rperm_;
    std::vector<int> c2r_count;  // rows filed under each permuted column (all 1 for square systems)
    void form_rmap();
    std::unique_ptr<SymbolicDevice> symdev;  // the device symbolic phase's matrix, from partition to assemble
    size_t n_blocks0 = 0;                    // entries of the assembled level-0 blocks
    void order_device(const HostCSC& A);
    void assemble_device();
    std::vector<int> inst_n, inst_lv, inst_off;  // per system: unknowns, levels, first position
    int* d_seg = nullptr;                        // device copy of inst_off plus n (batch GMRES)
    void make_hierarchy();
    std::vector<Cluster> cl_part;             // the hierarchy as partition() left it (assemble() starts there)
    std::vector<Cluster> cl0;                 // post-assembly structure, restored by factorize()
    std::vector<Edge> ed0;
    void* d_blocks0 = nullptr;                // pristine assembled blocks (factor precision)
    std::vector<ClusterID> cpermed;
    std::vector<Cluster> cl;                  // indexed by order
    std::vector<Edge> ed;
    std::vector<std::vector<int>> bottoms, interiors, interfaces;
    int current_bottom = 0;
    bool factorized = false;
    std::vector<LevelStats> stats;
    hipStream_t stream_ = nullptr;
    // graph mode
    bool graphs = false;
    bool capturing = false;
    hipStream_t up_stream_ = nullptr;  // descriptor uploads while the main stream is captured
    long asm_gen = 0;                  // bumped by assemble(): a new matrix invalidates the plan
    struct GraphPlan {
        bool valid = false;
        std::vector<hipGraphExec_t> exec;  // segments in order
        std::vector<int*> rb_ranks_dev;    // rank read-back after the segment (nullptr: none)
        std::vector<int*> rb_err_dev;
        std::vector<std::vector<int>> rb_ranks;
        int* chain_err = nullptr;
        std::vector<long long> sig;
    } gplan;
    std::vector<long long> plan_signature() const;
    void ensure_stream();
    int* d_cinfo = nullptr;   // per cluster (arena)
    int* d_lflags = nullptr;  // per level: OR of the flags, and the first cluster that raised them
    int* d_lfirst = nullptr;
    std::vector<int> h_info;  // host-side flags (predictor misses)
    void check_levels(int upto);
    void seg_begin();
    void seg_end(int* ranks_dev, int nranks, int* err_dev);
    void drop_graphs();
    template <typename S> int replay_t();
    void h2d(void* d, const void* h, size_t bytes);
    bool graph_rec = false;  // this factorization records the plan
    std::vector<long long> graph_warm_sig;  // settings whose arenas are already grown

    // device memory
    // Bump allocator over retained 256 MiB (or larger) hipMalloc chunks: reset() rewinds it
    // for the next factorization without returning memory, release() frees the chunks.
    struct Arena {
        std::vector<std::pair<void*, size_t>> chunks;
        size_t next = 0;
        char* cur = nullptr;
        size_t left = 0, total = 0;
        void* alloc(size_t bytes);
        void reset();
        void release();
    } arena, desc_arena;
    std::vector<double> h_blocks0;  // assembled level-0 blocks (FP64, host)
    std::vector<size_t> h_block_off;
    int* d_rowptr = nullptr;
    int* d_col = nullptr;
    double* d_val = nullptr;
    int* d_tptr = nullptr;  // CSR of App^T (CSC of App): the adjoint SpMV of cgls
    int* d_tcol = nullptr;
    double* d_tval = nullptr;
    long long d_nnz = 0;
    double* d_scratch = nullptr;  // dot partials + scalars
    int dot_blocks = 512;

    // rank path state
    float* d_mlp_w = nullptr;
    bool has_model = false;
    void* d_omega = nullptr;
    int om_rows = 0, om_cols = 0, om_elem = 0;
    void* d_omega_bf = nullptr;  // the same Gaussian rounded to bf16, ld ob_rows (a multiple of 32)
    int ob_rows = 0, ob_cols = 0;
    const void* omega_bf16(int rows, int cols);
    template <typename S> bool sketch_bf16() const { return sizeof(S) == 4 && tuned(T_SKETCH_BF16, 1) == 1; }
    // the sketch launch: bf16 when omb is set (FP32 fronts), else the FP32 / FP64 kernel on om
    template <typename S, typename Task, typename Tile>
    void run_sketch(const Task* d, const Tile* tiles, int nt, const S* om, const void* omb, int tcols);
    template <typename S>
    double rrqr_level(std::vector<Geqp3Task<S>>& gq, const std::vector<CopyTask<S>>& gat,
                    const std::vector<int>& gat_begin, const std::vector<int>& gq_cluster, int L, int* d_ranks,
                    int* d_err, std::vector<char>& used_sketch);
    template <typename S> const S* omega(int rows, int cols);
    template <typename S> int geqrt_big(const std::vector<GeqrtTask<S>>& big);
    template <typename S>
    int tpqrt_big(const std::vector<TpqrtTask<S>>& big, const std::vector<TpqrtLink<S>>& links);
    template <typename S> int tpqrt_coop(const std::vector<TpqrtTask<S>>& chains, const TpqrtLink<S>* d_links);
    int* d_chain_err = nullptr;  // raised by a cooperative chain whose wait exceeded its spin limit

    // solve schedule
    struct SolveSchedule;
    SolveSchedule* sched = nullptr;
    void* xw = nullptr;  // solve workspace (precision of the factors)
    size_t xw_bytes = 0;  // its capacity in bytes (grown, never shrunk)
    int prec_asm = 0;     // factor precision d_blocks0 was assembled in
    int prec_fact = 0;    // factor precision of the current factorization
    // GMRES workspace (FP64): Krylov basis, Z of the flexible variant and one work vector,
    // kept across calls and only grown
    double* d_krylov = nullptr;
    size_t krylov_bytes = 0;
    double* krylov_ws(size_t bytes);

    bool want_sparsify(const Cluster& c) const;
    int new_edge(int c1, int c2);
    int colour_interiors(const std::vector<int>& ints);
    std::vector<std::pair<int, int>> colour_pairs(const std::vector<int>& ints) const;
    static int dsatur_colours(std::vector<std::pair<int, int>> pairs, int V, std::vector<int>& col);
    void sort_out(Cluster& c);
    int find_out(const Cluster& c, int n2) const;
    bool fillin_dist2();
    void fillin_edges(bool independent);
    template <typename S> int factorize_t();
    template <typename S> void solve_sched_begin();
    template <typename S> void solve_sched_levels(int upto);
    template <typename S> int solve_t(double* x);
    template <typename S> void reassign_rows(int L);
    template <typename S> int solve_ls_t(const double* b, double* x);
    template <typename S> int solve_nrml_t(const double* y, double* x);
    template <typename S> void solve_phase_tr(int level);  // one level of the transposed backward pass
    template <typename S> void solve_sched_tr();
    template <typename S> void move_t(const void* si, const void* di, int cnt, bool staged);
    template <typename S> void load_rows_t(const double* b);
    template <typename S> void load_cols_t(const double* y);
    template <typename S> void store_cols_t(double* x);
    template <typename S> void solve_phase_t(int level, int phase, int dir);
    void* upload(const void* h, size_t bytes);
    // pinned staging for descriptor uploads: one host copy, then an asynchronous DMA; the
    // chunks are reused after every stream synchronisation
    struct PinnedArena {
        std::vector<std::pair<char*, size_t>> chunks;
        size_t next = 0;
        char* cur = nullptr;
        size_t left = 0;
        char* alloc(size_t bytes);
        void reset() { next = 0; cur = nullptr; left = 0; }
        void release();
    } staging;
    void sync_stream();
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_pending;
    template <class F> void timed(int cat, long launches, long tasks, F&& f);
    void prof_add(int cat, double flops, double bytes) {
        prof[cat].flops += flops;
        prof[cat].bytes += bytes;
    }
    // Tasks filed by builder slot and launch bucket (parts[i][q]): each bucket's slots copied in
    // slot order into one staging block (a single threaded pass) and uploaded once.
    template <typename T, size_t NBK>
    std::array<std::pair<T*, int>, NBK> upload_buckets(const std::vector<std::array<std::vector<T>, NBK>>& parts);
    template <typename T> T* upload_vec(const std::vector<T>& v) {
        return v.empty() ? nullptr : (T*)upload(v.data(), v.size() * sizeof(T));
    }
};

}  // namespace spaqr
