// C ABI (include/spaqr_gpu.h) over the level planner and the batched kernels.
#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spaqr_gpu.h"
#include "errors.hpp"
#include "kernels.hpp"
#include "launch.hpp"
#include "learned.hpp"
#include "rrqr.hpp"
#include "tree.hpp"
#include "tree_impl.hpp"

using namespace spaqr;

struct spaqr_tree {
    Tree* t;
};

static_assert(sizeof(spaqr_geqrt_desc) == sizeof(GeqrtTask<double>), "geqrt desc layout");
static_assert(sizeof(spaqr_larfb_desc) == sizeof(LarfbTask<double>), "larfb desc layout");
static_assert(sizeof(spaqr_tpqrt_desc) == sizeof(TpqrtTask<double>), "tpqrt desc layout");
static_assert(sizeof(spaqr_tpqrt_link) == sizeof(TpqrtLink<double>), "tpqrt link layout");
static_assert(sizeof(spaqr_tpmqrt_desc) == sizeof(TpmqrtTask<double>), "tpmqrt desc layout");
static_assert(sizeof(spaqr_tpmqrt_link) == sizeof(TpmqrtLink<double>), "tpmqrt link layout");
static_assert(sizeof(spaqr_geqp3_desc) == sizeof(Geqp3Task<double>), "geqp3 desc layout");
static_assert(sizeof(spaqr_spars_desc) == sizeof(SparsApplyTask<double>), "spars desc layout");


// batched-kernel call: upload descriptors, launch, synchronize, free
template <typename F>
static int run_batched(const void* host, size_t bytes, const void* host2, size_t bytes2, void* stream, F launch) {
    hipStream_t s = (hipStream_t)stream;
    void* d = nullptr;
    void* d2 = nullptr;
    if (hipMalloc(&d, bytes ? bytes : 8) != hipSuccess) { api_error() = "hipMalloc failed"; return -100; }
    if (bytes2 && hipMalloc(&d2, bytes2) != hipSuccess) { (void)hipFree(d); api_error() = "hipMalloc failed"; return -100; }
    hipError_t e = hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && bytes2) e = hipMemcpyAsync(d2, host2, bytes2, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        launch(d, d2, s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (d2) (void)hipFree(d2);
    if (e != hipSuccess) { api_error() = hipGetErrorString(e); return -100; }
    return 0;
}


// Device buffers of one batched call: descriptor copies (kept with a host copy until the
// stream syncs) and zeroed scratch, all freed when the call returns.
struct CallScratch {
    hipStream_t s;
    std::vector<void*> owned;
    std::vector<std::vector<char>> keep;
    explicit CallScratch(hipStream_t st) : s(st) {}
    ~CallScratch() {
        for (void* p : owned) (void)hipFree(p);
    }
    void* alloc(size_t b) {
        void* p = nullptr;
        if (hipMalloc(&p, b ? b : 8) != hipSuccess) throw std::runtime_error("hipMalloc failed");
        owned.push_back(p);
        return p;
    }
    void* upload(const void* h, size_t b) {
        void* p = alloc(b);
        keep.emplace_back((const char*)h, (const char*)h + b);
        if (b && hipMemcpyAsync(p, keep.back().data(), b, hipMemcpyHostToDevice, s) != hipSuccess)
            throw std::runtime_error("descriptor upload failed");
        return p;
    }
    template <typename T> T* upload_vec(const std::vector<T>& v) {
        return v.empty() ? nullptr : (T*)upload(v.data(), v.size() * sizeof(T));
    }
    int finish() {
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { api_error() = hipGetErrorString(e); return -100; }
        return 0;
    }
};

extern "C" {

int spaqr_shutdown(void) {
    HostPool::shutdown();
    return 0;
}

spaqr_tree* spaqr_tree_create(int nlevels, int skip) {
    if (nlevels < 1 || skip < 0) { api_error() = "nlevels must be >= 1 and skip >= 0"; return nullptr; }
    try {
        spaqr_tree* h = new spaqr_tree;
        h->t = new Tree(nlevels, skip);
        return h;
    } catch (std::exception& e) {
        api_error() = e.what();
        return nullptr;
    }
}

void spaqr_tree_destroy(spaqr_tree* h) {
    if (!h) return;
    delete h->t;
    delete h;
}

#define TREE_OR_FAIL(h)                 \
    if (!h || !h->t) {                  \
        api_error() = "null tree handle";     \
        return -1;                      \
    }

int spaqr_tree_set_tol(spaqr_tree* h, double tol) { TREE_OR_FAIL(h); if (!(tol >= 0)) return -2; h->t->tol = tol; return 0; }
int spaqr_tree_set_skip(spaqr_tree* h, int s) { TREE_OR_FAIL(h); if (s < 0) return -2; h->t->skip = s; return 0; }
int spaqr_tree_set_scale(spaqr_tree* h, int s) {
    TREE_OR_FAIL(h);
    if (s != 1) { h->t->last_error = "ParTree::factorize requires scale == 1 (src/ptree.cpp:875)"; return -2; }
    h->t->scale = s;
    return 0;
}
int spaqr_tree_set_square(spaqr_tree* h, int s) {
    TREE_OR_FAIL(h);
    if (s != 0 && s != 1) { h->t->last_error = "set_square takes 0 (least squares, m > n) or 1"; return -2; }
    h->t->square = s;
    return 0;
}
int spaqr_tree_set_order(spaqr_tree* h, float o) {
    TREE_OR_FAIL(h);
    if (o != 1.0f) { h->t->last_error = "ParTree supports order 1 only (pspaQR.cpp:174-180)"; return -2; }
    h->t->order = o;
    return 0;
}
int spaqr_tree_set_verbose(spaqr_tree* h, int v) { TREE_OR_FAIL(h); h->t->verbose = v; return 0; }
int spaqr_tree_set_precision(spaqr_tree* h, int bits) {
    TREE_OR_FAIL(h);
    if (bits != 64 && bits != 32) { h->t->last_error = "precision must be 64 or 32"; return -2; }
    h->t->prec = bits;
    return 0;
}
int spaqr_tree_set_Xcoo(spaqr_tree* h, const double* X, int d, int n) {
    TREE_OR_FAIL(h);
    if (!X || d <= 0 || n <= 0) return -2;
    h->t->X.assign(X, X + (size_t)d * n);
    h->t->xdim = d;
    return 0;
}
int spaqr_tree_set_parts(spaqr_tree* h, const int* parts, int n) {
    TREE_OR_FAIL(h);
    if (n < 0 || (n > 0 && !parts)) return -2;
    h->t->user_parts.assign(parts, parts + n);
    return 0;
}
int spaqr_tree_get_parts(const spaqr_tree* h, int* out, int cap) {
    if (!h || !h->t) return -1;
    const auto& p = h->t->leaf_parts;
    if (out) std::copy(p.begin(), p.begin() + std::min<size_t>(p.size(), (size_t)std::max(cap, 0)), out);
    return (int)p.size();
}
int spaqr_tree_partition(spaqr_tree* h, int nrows, int ncols, const int* colptr, const int* rowind, const double* val) {
    TREE_OR_FAIL(h);
    if (nrows <= 0 || ncols <= 0 || !colptr || !rowind || !val) return -2;
    HostCSC A;
    A.nr = nrows;
    A.nc = ncols;
    A.colptr.assign(colptr, colptr + ncols + 1);
    A.rowind.assign(rowind, rowind + colptr[ncols]);
    A.val.assign(val, val + colptr[ncols]);
    return h->t->partition(A);
}
int spaqr_tree_partition_batch(spaqr_tree* h, int nsys, const int* sizes, const int* levels, int n, const int* colptr,
                               const int* rowind, const double* val) {
    TREE_OR_FAIL(h);
    if (nsys < 1 || !sizes || !levels) return -2;
    if (n <= 0 || !colptr || !rowind || !val) return -5;
    HostCSC A;
    A.nr = A.nc = n;
    A.colptr.assign(colptr, colptr + n + 1);
    A.rowind.assign(rowind, rowind + colptr[n]);
    A.val.assign(val, val + colptr[n]);
    return h->t->partition_batch(A, std::vector<int>(sizes, sizes + nsys), std::vector<int>(levels, levels + nsys));
}
int spaqr_tree_set_tuning(spaqr_tree* h, int param, int value) {
    TREE_OR_FAIL(h);
    if (param < 0 || param >= T_NPARAM) { h->t->last_error = "unknown tuning parameter"; return -2; }
    auto in = [&](std::initializer_list<int> ok) {
        for (int v : ok)
            if (v == value) return true;
        return false;
    };
    bool valid = value == 0;
    switch (param) {
        case T_BIG_FRONT: case T_SKETCH_UNPIVOTED: valid = valid || (value >= 16 && value <= (1 << 30)); break;
        case T_TPMQRT_TILE: valid = valid || in({16, 32, 64}); break;
        case T_LARFB_COLS: valid = valid || in({16, 32, 64, 128}); break;
        case T_TRSM_ROWS: case T_SPARS_ROWS: case T_RANK_COLS: valid = valid || in({32, 64, 128, 256}); break;
        case T_SPARS_COLS: valid = valid || in({32, 64, 128}); break;
        case T_TPQRT_COOP: case T_SMALL_TASKS: case T_SKETCH_BF16: valid = valid || in({1, 2}); break;
        case T_SKETCH_TILE: valid = valid || in({64, 128, 256}); break;
        case T_SKETCH_GROW: valid = valid || (value >= 100 && value <= 200); break;
        case T_SKETCH_PAD: valid = valid || (value >= 1 && value <= 65); break;
        case T_PANEL_COLS: valid = valid || in({32, 64, 128, 256}); break;
        default: break;
    }
    if (!valid) { h->t->last_error = "tuning value out of range for this parameter"; return -3; }
    h->t->tune[param] = value;
    return 0;
}
int spaqr_tree_get_tuning(const spaqr_tree* h, int param) {
    if (!h || !h->t || param < 0 || param >= T_NPARAM) return -1;
    return h->t->tune[param];
}
int spaqr_tree_set_pivot_check(spaqr_tree* h, double rel) {
    TREE_OR_FAIL(h);
    if (!(rel >= 0.0 && rel < 1.0)) { h->t->last_error = "the pivot threshold must be in [0, 1)"; return -2; }
    h->t->pivot_rel = rel;
    return 0;
}
int spaqr_tree_set_krylov_mode(spaqr_tree* h, int mode) {
    TREE_OR_FAIL(h);
    if (mode != 0 && mode != 1) { h->t->last_error = "Krylov mode must be 0 or 1"; return -2; }
    h->t->krylov_mode = mode;
    return 0;
}
int spaqr_tree_get_info(spaqr_tree* h, int* out, int cap) {
    TREE_OR_FAIL(h);
    try {
        std::vector<int> v;
        const int n = h->t->get_info(v);
        if (out)
            for (int i = 0; i < std::min(n, std::max(cap, 0)); ++i) out[i] = v[i];
        return n;
    } catch (std::exception& e) {
        h->t->last_error = e.what();
        return -100;
    }
}
int spaqr_tree_set_graphs(spaqr_tree* h, int on) {
    TREE_OR_FAIL(h);
    return h->t->set_graphs(on != 0);
}
int spaqr_tree_graph_info(const spaqr_tree* h, long long* out) {
    TREE_OR_FAIL(h);
    if (!out) return -2;
    out[0] = h->t->graphs_on();
    out[1] = h->t->graph_segments();
    out[2] = h->t->graph_replays;
    return 0;
}
int spaqr_tree_num_systems(const spaqr_tree* h) { return (h && h->t) ? h->t->n_systems() : -1; }
int spaqr_tree_get_systems(const spaqr_tree* h, int* out) {
    TREE_OR_FAIL(h);
    if (!out) return -2;
    for (int s = 0; s < h->t->n_systems(); ++s) {
        out[2 * s] = h->t->system_offset(s);
        out[2 * s + 1] = h->t->system_size(s);
    }
    return 0;
}
int spaqr_tree_get_cluster_systems(const spaqr_tree* h, int* out) {
    TREE_OR_FAIL(h);
    if (!out) return -2;
    for (int o = 0; o < h->t->num_clusters(); ++o) out[o] = h->t->cluster(o).inst;
    return 0;
}
int spaqr_gmres_batch(spaqr_tree* h, const double* b, double* x, int iters, double tol, int restart, double true_tol,
                      int flexible, spaqr_gmres_result* out) {
    TREE_OR_FAIL(h);
    if (!b || !x || iters <= 0) return -2;
    std::vector<GmresResult> r;
    const int rc = h->t->gmres_batch_dev(b, x, iters, tol, restart, true_tol, flexible != 0, r);
    if (rc == 0 && out)
        for (size_t s = 0; s < r.size(); ++s) {
            out[s].iters = r[s].iters;
            out[s].cycles = r[s].cycles;
            out[s].prec_res = r[s].prec_res;
            out[s].true_res = r[s].true_res;
        }
    return rc;
}
int spaqr_tree_assemble(spaqr_tree* h) { TREE_OR_FAIL(h); return h->t->assemble(); }
int spaqr_tree_factorize(spaqr_tree* h) { TREE_OR_FAIL(h); return h->t->factorize(); }
int spaqr_tree_solve(spaqr_tree* h, double* x) { TREE_OR_FAIL(h); if (!x) return -2; return h->t->solve_dev(x); }
int spaqr_tree_solve_load(spaqr_tree* h, const double* x) {
    TREE_OR_FAIL(h);
    if (!x) return -2;
    return h->t->solve_load(x);
}
int spaqr_tree_solve_phase(spaqr_tree* h, int level, int phase, int dir) {
    TREE_OR_FAIL(h);
    return h->t->solve_phase(level, phase, dir);
}
int spaqr_tree_solve_store(spaqr_tree* h, double* x) {
    TREE_OR_FAIL(h);
    if (!x) return -2;
    return h->t->solve_store(x);
}
int spaqr_tree_permute(const spaqr_tree* h, int to_natural, const double* in, double* out) {
    TREE_OR_FAIL(h);
    if (!in || !out || in == out) return -3;
    const auto& p = h->t->cperm();
    if (to_natural)
        for (size_t k = 0; k < p.size(); ++k) out[p[k]] = in[k];
    else
        for (size_t k = 0; k < p.size(); ++k) out[k] = in[p[k]];
    return 0;
}
int spaqr_tree_spmv(spaqr_tree* h, const double* x, double* y) {
    TREE_OR_FAIL(h);
    if (!x || !y) return -2;
    return h->t->spmv_dev(x, y);
}
int spaqr_gmres(spaqr_tree* h, const double* b, double* x, int iters, double tol, int restart, double true_tol,
                spaqr_gmres_result* out, double* hist, int hist_cap, int* hist_len) {
    TREE_OR_FAIL(h);
    if (!b || !x || iters <= 0) return -2;
    GmresResult r;
    int rc = h->t->gmres_dev(b, x, iters, tol, restart, true_tol, &r, hist, hist_cap, hist_len, false);
    if (out) { out->iters = r.iters; out->cycles = r.cycles; out->prec_res = r.prec_res; out->true_res = r.true_res; }
    return rc;
}
int spaqr_fgmres(spaqr_tree* h, const double* b, double* x, int iters, double tol, int restart, double true_tol,
                 spaqr_gmres_result* out, double* hist, int hist_cap, int* hist_len) {
    TREE_OR_FAIL(h);
    if (!b || !x || iters <= 0) return -2;
    GmresResult r;
    int rc = h->t->gmres_dev(b, x, iters, tol, restart, true_tol, &r, hist, hist_cap, hist_len, true);
    if (out) { out->iters = r.iters; out->cycles = r.cycles; out->prec_res = r.prec_res; out->true_res = r.true_res; }
    return rc;
}
int spaqr_tree_n(const spaqr_tree* h) { return (h && h->t) ? h->t->n_unknowns() : -1; }
int spaqr_tree_m(const spaqr_tree* h) { return (h && h->t) ? h->t->n_rows() : -1; }
int spaqr_tree_set_device_symbolic(spaqr_tree* h, int on) {
    TREE_OR_FAIL(h);
    h->t->device_symbolic = on != 0;
    return 0;
}
int spaqr_tree_symbolic_ms(const spaqr_tree* h, double* out) {
    TREE_OR_FAIL(h);
    if (!out) return -2;
    out[0] = h->t->t_partition_ms;
    out[1] = h->t->t_assemble_ms;
    return 0;
}
int spaqr_tree_get_rperm(const spaqr_tree* h, int* out) {
    TREE_OR_FAIL(h);
    if (!out) return -2;
    const auto& p = h->t->rperm();
    std::memcpy(out, p.data(), sizeof(int) * p.size());
    return 0;
}
int spaqr_tree_permute_rows(const spaqr_tree* h, int to_natural, const double* in, double* out) {
    TREE_OR_FAIL(h);
    if (!in || !out || in == out) return -3;
    const auto& p = h->t->rperm();
    if (to_natural)
        for (size_t k = 0; k < p.size(); ++k) out[p[k]] = in[k];
    else
        for (size_t k = 0; k < p.size(); ++k) out[k] = in[p[k]];
    return 0;
}
int spaqr_tree_get_rows(const spaqr_tree* h, int* out) {
    TREE_OR_FAIL(h);
    if (!out) return -2;
    for (int o = 0; o < h->t->num_clusters(); ++o) {
        const Cluster& c = h->t->cluster(o);
        int* r = out + 5 * o;
        int moved = 0;
        for (int v : c.r2c) moved += (v != o);
        r[0] = c.rstart; r[1] = c.rows_org; r[2] = c.tot_extra; r[3] = c.row_rank; r[4] = moved;
    }
    return 0;
}
int spaqr_tree_solve_ls(spaqr_tree* h, const double* b, double* x) {
    TREE_OR_FAIL(h);
    if (!b || !x) return -2;
    return h->t->solve_ls_dev(b, x);
}
int spaqr_tree_solve_nrml(spaqr_tree* h, const double* y, double* x) {
    TREE_OR_FAIL(h);
    if (!y || !x) return -2;
    return h->t->solve_nrml_dev(y, x);
}
int spaqr_tree_spmv_t(spaqr_tree* h, const double* r, double* y) {
    TREE_OR_FAIL(h);
    if (!r || !y) return -2;
    return h->t->spmv_t_dev(r, y);
}
int spaqr_cgls(spaqr_tree* h, const double* b, double* x, int iters, double tol, int* iters_out, double* rel) {
    TREE_OR_FAIL(h);
    if (!b || !x || iters < 0 || !iters_out || !rel) return -2;
    return h->t->cgls_dev(b, x, iters, tol, iters_out, rel);
}
int spaqr_tree_get_cperm(const spaqr_tree* h, int* out) {
    TREE_OR_FAIL(h);
    const auto& p = h->t->cperm();
    std::memcpy(out, p.data(), sizeof(int) * p.size());
    return 0;
}
int spaqr_tree_num_clusters(const spaqr_tree* h) { return (h && h->t) ? h->t->num_clusters() : -1; }
int spaqr_tree_get_clusters(const spaqr_tree* h, int* out) {
    TREE_OR_FAIL(h);
    for (int o = 0; o < h->t->num_clusters(); ++o) {
        const Cluster& c = h->t->cluster(o);
        int* r = out + 13 * o;
        r[0] = c.merge_level; r[1] = c.id.self.lvl; r[2] = c.id.self.sep;
        r[3] = c.id.l.lvl; r[4] = c.id.l.sep; r[5] = c.id.r.lvl; r[6] = c.id.r.sep;
        r[7] = c.id.section; r[8] = c.id.part; r[9] = c.cstart; r[10] = c.cols_org;
        r[11] = c.parent; r[12] = c.sparsified ? c.rank : -1;
    }
    return 0;
}
int spaqr_tree_level_stats(const spaqr_tree* h, double* out, int max_levels) {
    TREE_OR_FAIL(h);
    const auto& s = h->t->level_stats();
    int nl = std::min((int)s.size(), max_levels);
    for (int i = 0; i < nl; ++i) {
        double* r = out + 8 * i;
        r[0] = s[i].n_interiors; r[1] = s[i].n_interfaces; r[2] = s[i].n_sparsified; r[3] = s[i].max_front;
        r[4] = s[i].colors; r[5] = (double)s[i].cols_before; r[6] = (double)s[i].ranks_after; r[7] = s[i].flops;
    }
    return nl;
}
int spaqr_tree_diag_R(spaqr_tree* h, int order, double* out, int maxn) { TREE_OR_FAIL(h); return h->t->diag_R(order, out, maxn); }
double spaqr_tree_factor_ms(const spaqr_tree* h) { return (h && h->t) ? h->t->t_factor_ms : -1.0; }
void* spaqr_tree_stream(spaqr_tree* h) { return (h && h->t) ? (void*)h->t->stream() : nullptr; }
const char* spaqr_last_error(const spaqr_tree* h) {
    if (h && h->t && !h->t->last_error.empty()) return h->t->last_error.c_str();
    return api_error().c_str();
}

int spaqr_tree_sync(spaqr_tree* h) {
    TREE_OR_FAIL(h);
    return h->t->sync();
}
int spaqr_tree_set_profiling(spaqr_tree* h, int on) {
    TREE_OR_FAIL(h);
    h->t->profiling = on != 0;
    return 0;
}
int spaqr_tree_profile(spaqr_tree* h, double* out, int max_classes) {
    TREE_OR_FAIL(h);
    int rc = h->t->sync();
    if (rc) return rc;
    int nc = std::min((int)P_NCAT, max_classes);
    for (int i = 0; i < nc; ++i) {
        const ProfEntry& p = h->t->prof[i];
        out[5 * i] = p.ms; out[5 * i + 1] = (double)p.launches; out[5 * i + 2] = (double)p.tasks;
        out[5 * i + 3] = p.flops; out[5 * i + 4] = p.bytes;
    }
    return nc;
}
int spaqr_tree_profile_reset(spaqr_tree* h) {
    TREE_OR_FAIL(h);
    int rc = h->t->sync();
    h->t->prof_reset();
    return rc;
}

// ---------------------------------------------------------------- rank path
int spaqr_tree_set_rank_mode(spaqr_tree* h, int mode) {
    TREE_OR_FAIL(h);
    if (mode < 0 || mode > 2) { api_error() = "rank mode must be 0, 1 or 2"; return -2; }
    h->t->rank_mode = mode;
    return 0;
}
int spaqr_tree_set_rank_model(spaqr_tree* h, const float* w, int nparams) {
    TREE_OR_FAIL(h);
    if (!w) return -2;
    if (nparams != kMlpParams) { api_error() = "rank model: wrong parameter count"; return -3; }
    try {
        return h->t->set_rank_model(w);
    } catch (std::exception& e) {
        h->t->last_error = e.what();
        return -100;
    }
}
int spaqr_tree_set_rank_log(spaqr_tree* h, int on) {
    TREE_OR_FAIL(h);
    h->t->rank_log = on != 0;
    return 0;
}
int spaqr_tree_rank_log(const spaqr_tree* h, float* x, int* meta, int cap) {
    if (!h || !h->t) return -1;
    const auto& L = h->t->rlog;
    const int n = (int)L.rank.size();
    const int m = std::min(n, std::max(cap, 0));
    for (int i = 0; i < m; ++i) {
        if (x) std::memcpy(x + (size_t)i * kRankFeatures, L.x.data() + (size_t)i * kRankFeatures, sizeof(float) * kRankFeatures);
        if (meta) {
            meta[5 * i] = L.rank[i];
            meta[5 * i + 1] = L.k[i];
            meta[5 * i + 2] = L.w[i];
            meta[5 * i + 3] = L.level[i];
            meta[5 * i + 4] = L.mode[i];
        }
    }
    return n;
}
int spaqr_tree_rank_log_clear(spaqr_tree* h) {
    TREE_OR_FAIL(h);
    h->t->rlog = Tree::RankLog();
    return 0;
}
int spaqr_tree_rank_stats(const spaqr_tree* h, double* out) {
    if (!h || !h->t) return -1;
    if (!out) return -2;
    const auto& r = h->t->rstats;
    out[0] = (double)r.tasks;
    out[1] = (double)r.sketched;
    out[2] = (double)r.fallback;
    out[3] = (double)r.direct;
    out[4] = r.sketch_cols;
    out[5] = r.full_cols;
    out[6] = (double)r.retried;
    return 0;
}
int spaqr_tree_rank_stats_reset(spaqr_tree* h) {
    TREE_OR_FAIL(h);
    h->t->rstats = Tree::RankStats();
    return 0;
}
int spaqr_mlp_params(void) { return kMlpParams; }
int spaqr_mlp_features(void) { return kRankFeatures; }
int spaqr_mlp_predict(const float* X, int n, const float* W, float* out) {
    if (n < 0 || (n > 0 && (!X || !out)) || !W) return -2;
    if (n == 0) return 0;
    try {
        return mlp_predict_host(X, n, W, out);
    } catch (std::exception& e) {
        api_error() = e.what();
        return -100;
    }
}
int spaqr_mlp_grad(const float* X, const float* y, int n, const float* W, float* grad, float* loss) {
    if (n <= 0 || !X || !y || !W || !grad || !loss) return -2;
    try {
        return mlp_grad_host(X, y, n, W, grad, loss);
    } catch (std::exception& e) {
        api_error() = e.what();
        return -100;
    }
}
int spaqr_mlp_train(const float* X, const float* y, int n, float* W, int epochs, float lr, float* loss_hist) {
    if (n <= 0 || !X || !y || !W || epochs < 0 || !(lr > 0.f)) return -2;
    try {
        return mlp_train_host(X, y, n, W, epochs, lr, loss_hist);
    } catch (std::exception& e) {
        api_error() = e.what();
        return -100;
    }
}

void* spaqr_device_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) { api_error() = "hipMalloc failed"; return nullptr; }
    return p;
}
int spaqr_device_free(void* p) { return hipFree(p) == hipSuccess ? 0 : -100; }
int spaqr_memcpy_h2d(void* d, const void* h, size_t bytes) {
    return hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -100;
}
int spaqr_memcpy_d2h(void* h, const void* d, size_t bytes) {
    return hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -100;
}
int spaqr_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : -100; }

// ---------------------------------------------------------------- batched kernels
int spaqr_geqrt_batched(int prec, int count, const spaqr_geqrt_desc* d, void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    for (int i = 0; i < count; ++i)
        if (d[i].m < 0 || d[i].n < 0 || d[i].lda < std::max(1, d[i].m)) return -3;
    if (count == 0) return 0;
    return run_batched(d, sizeof(*d) * count, nullptr, 0, stream, [&](void* dd, void*, hipStream_t s) {
        if (prec == 64) launch_geqrt<double>((const GeqrtTask<double>*)dd, count, s);
        else launch_geqrt<float>((const GeqrtTask<float>*)dd, count, s);
    });
}

int spaqr_larfb_batched(int prec, int count, const spaqr_larfb_desc* d, void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    for (int i = 0; i < count; ++i)
        if (d[i].m < 0 || d[i].n < 0 || d[i].k < 0 || d[i].ldc < std::max(1, d[i].m) || (d[i].R && d[i].n != d[i].k2))
            return -3;
    if (count == 0) return 0;
    if (prec == 64)
        return run_batched(d, sizeof(*d) * count, nullptr, 0, stream, [&](void* dd, void*, hipStream_t s) {
            launch_larfb<double>((const LarfbTask<double>*)dd, count, s);
        });
    // FP32: blocks the C-resident kernels take (Q^T C only: m <= 512, n <= 128; C R^{-1} only:
    // m <= 128, k2 <= 512) go to them
    // (buckets 0..3: Q^T C kernels by rows; 4..6: C R^{-1} by k2)
    std::vector<LarfbTask<float>> res[7], rest;
    for (int i = 0; i < count; ++i) {
        LarfbTask<float> t;
        static_assert(sizeof(t) == sizeof(d[i]), "larfb desc layout");
        memcpy(&t, &d[i], sizeof(t));
        int q = larfb_res_bucket(t);
        if (q < 0 && trsm_res_bucket(t) >= 0) q = 4 + trsm_res_bucket(t);
        (q < 0 ? rest : res[q]).push_back(t);
    }
    for (int q = 0; q < 7; ++q) {
        if (res[q].empty()) continue;
        const int n = (int)res[q].size();
        const int rc = run_batched(res[q].data(), sizeof(LarfbTask<float>) * n, nullptr, 0, stream,
                                   [&](void* dd, void*, hipStream_t s) {
                                       const LarfbTask<float>* td = (const LarfbTask<float>*)dd;
                                       if (q < 3) launch_larfb_res(td, n, 4 << q, s);
                                       else if (q == 3) launch_larfb_res2(td, n, s);
                                       else launch_trsm_res(td, n, 4 << (q - 4), s);
                                   });
        if (rc) return rc;
    }
    if (rest.empty()) return 0;
    const int n = (int)rest.size();
    return run_batched(rest.data(), sizeof(LarfbTask<float>) * n, nullptr, 0, stream,
                       [&](void* dd, void*, hipStream_t s) { launch_larfb<float>((const LarfbTask<float>*)dd, n, s); });
}

}  // extern "C"

// Chains are dispatched as the level planner dispatches them: k <= 64 with links of at most
// 64 rows one per wave; several 32-column panels (k > 32) with more than one link on the
// cooperative kernel (workgroup p owns panel p, links pipelined through progress counters);
// the rest one workgroup per chain.
template <typename S>
static int tpqrt_dispatch(int count, const spaqr_tpqrt_desc* d, int nlinks, const spaqr_tpqrt_link* links,
                          hipStream_t s) {
    static_assert(sizeof(TpqrtTask<S>) == sizeof(spaqr_tpqrt_desc), "tpqrt desc layout");
    static_assert(sizeof(TpqrtLink<S>) == sizeof(spaqr_tpqrt_link), "tpqrt link layout");
    const TpqrtTask<S>* td = (const TpqrtTask<S>*)d;
    const TpqrtLink<S>* tl = (const TpqrtLink<S>*)links;
    std::vector<TpqrtTask<S>> wave, plain, coop;
    for (int i = 0; i < count; ++i) {
        const TpqrtTask<S>& t = td[i];
        bool small = t.k <= 64;
        for (int j = 0; j < t.nlinks && small; ++j) small = tl[t.link0 + j].rows <= 64;
        if (small) wave.push_back(t);
        else if (t.k > NB && t.nlinks > 1) coop.push_back(t);
        else plain.push_back(t);
    }
    CallScratch cs(s);
    try {
        auto* d_links = (const TpqrtLink<S>*)cs.upload(links, sizeof(*links) * std::max(nlinks, 1));
        launch_tpqrt_wave<S>(cs.upload_vec(wave), d_links, (int)wave.size(), s);
        launch_tpqrt<S>(cs.upload_vec(plain), d_links, (int)plain.size(), s);
        if (!coop.empty()) {
            const int maxb = tpqrt_coop_max_blocks<S>();
            int panels = 0;
            for (const auto& t : coop) panels += (t.k + NB - 1) / NB;
            std::vector<int> zeros(panels + 1, 0);
            int* cnt = (int*)cs.upload(zeros.data(), sizeof(int) * zeros.size());
            int* err = cnt + panels;
            int coff = 0;
            for (size_t i = 0; i < coop.size();) {
                std::vector<TpqrtCoopTask<S>> tk;
                std::vector<int> b2t;
                int nb = 0;
                while (i < coop.size()) {
                    const TpqrtTask<S>& c = coop[i];
                    const int P = (c.k + NB - 1) / NB;
                    if (P > maxb) throw std::runtime_error("cooperative tpqrt: chain wider than the co-resident limit");
                    if (nb + P > maxb) break;
                    tk.push_back({c.R, cnt + coff, c.k, c.ldr, c.link0, c.nlinks, nb, 0});
                    for (int q = 0; q < P; ++q) b2t.push_back((int)tk.size() - 1);
                    nb += P;
                    coff += P;
                    ++i;
                }
                launch_tpqrt_coop<S>(cs.upload_vec(tk), cs.upload_vec(b2t), nb, d_links, err, s);
            }
            int herr = 0;
            if (hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                throw std::runtime_error("cooperative tpqrt: synchronization failed");
            if (herr) throw std::runtime_error("cooperative tpqrt: a wait exceeded its spin limit");
        }
        return cs.finish();
    } catch (std::exception& e) {
        api_error() = e.what();
        (void)hipStreamSynchronize(s);
        return -100;
    }
}

extern "C" {

int spaqr_tpqrt_batched(int prec, int count, const spaqr_tpqrt_desc* d, int nlinks, const spaqr_tpqrt_link* links,
                        void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    if (nlinks < 0 || (nlinks > 0 && !links)) return -4;
    for (int i = 0; i < count; ++i)
        if (d[i].k < 0 || d[i].link0 < 0 || d[i].nlinks < 0 || d[i].link0 + d[i].nlinks > nlinks ||
            d[i].ldr < std::max(1, d[i].k))
            return -3;
    for (int j = 0; j < nlinks; ++j)
        if (links[j].rows < 0 || links[j].ldb < std::max(1, links[j].rows)) return -5;
    if (count == 0) return 0;
    return prec == 64 ? tpqrt_dispatch<double>(count, d, nlinks, links, (hipStream_t)stream)
                      : tpqrt_dispatch<float>(count, d, nlinks, links, (hipStream_t)stream);
}

int spaqr_tpmqrt_batched(int prec, int count, const spaqr_tpmqrt_desc* d, int nlinks, const spaqr_tpmqrt_link* links,
                         void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    for (int i = 0; i < count; ++i)
        if (d[i].link0 < 0 || d[i].nlinks < 0 || d[i].link0 + d[i].nlinks > nlinks) return -3;
    if (count == 0) return 0;
    if (prec == 64)
        return run_batched(d, sizeof(*d) * count, links, sizeof(*links) * nlinks, stream,
                           [&](void* dd, void* dl, hipStream_t s) {
                               launch_tpmqrt<double>((const TpmqrtTask<double>*)dd, (const TpmqrtLink<double>*)dl, count, s);
                           });
    // FP32: chains the C-resident kernel takes (tiles <= 64 columns, links <= 512 rows) go to it
    std::vector<TpmqrtLink<float>> hl(nlinks);
    static_assert(sizeof(TpmqrtLink<float>) == sizeof(spaqr_tpmqrt_link), "tpmqrt link layout");
    if (nlinks) memcpy(hl.data(), links, sizeof(TpmqrtLink<float>) * nlinks);
    std::vector<TpmqrtTask<float>> res[4], rest;
    for (int i = 0; i < count; ++i) {
        TpmqrtTask<float> t;
        static_assert(sizeof(t) == sizeof(d[i]), "tpmqrt desc layout");
        memcpy(&t, &d[i], sizeof(t));
        const int q = tpmqrt_res_bucket(t, hl.data());
        (q < 0 ? rest : res[q]).push_back(t);
    }
    for (int q = 0; q < 5; ++q) {
        std::vector<TpmqrtTask<float>>& v = q < 4 ? res[q] : rest;
        if (v.empty()) continue;
        const int n = (int)v.size();
        const int rc = run_batched(v.data(), sizeof(TpmqrtTask<float>) * n, links, sizeof(*links) * nlinks, stream,
                                   [&](void* dd, void* dl, hipStream_t s) {
                                       const TpmqrtTask<float>* td = (const TpmqrtTask<float>*)dd;
                                       const TpmqrtLink<float>* tl = (const TpmqrtLink<float>*)dl;
                                       if (q < 3) launch_tpmqrt_res(td, tl, n, 4 << q, s);
                                       else if (q == 3) launch_tpmqrt_res2(td, tl, n, s);
                                       else launch_tpmqrt<float>(td, tl, n, s);
                                   });
        if (rc) return rc;
    }
    return 0;
}

int spaqr_geqp3_batched(int prec, int count, const spaqr_geqp3_desc* d, double tol, double tie_rel, void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    if (!(tol >= 0)) return -4;
    if (!(tie_rel >= 0 && tie_rel < 1)) return -5;
    if (count == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    std::vector<void*> owned;
    std::vector<std::vector<char>> keep;
    RrqrAlloc mem{[&](size_t b) {
                      void* p = nullptr;
                      if (hipMalloc(&p, b ? b : 8) != hipSuccess) throw std::runtime_error("hipMalloc failed");
                      owned.push_back(p);
                      return p;
                  },
                  [&](const void* h, size_t b) {
                      void* p = nullptr;
                      if (hipMalloc(&p, b ? b : 8) != hipSuccess) throw std::runtime_error("hipMalloc failed");
                      owned.push_back(p);
                      keep.emplace_back((const char*)h, (const char*)h + b);
                      (void)hipMemcpyAsync(p, keep.back().data(), b, hipMemcpyHostToDevice, s);
                      return p;
                  }};
    int rc = 0;
    try {
        const int zero = 0;
        int* d_err = (int*)mem.upload(&zero, sizeof(int));
        if (prec == 64) {
            std::vector<Geqp3Task<double>> v((const Geqp3Task<double>*)d, (const Geqp3Task<double>*)d + count);
            geqp3_grouped<double>(v, tol, tie_rel, s, mem, d_err);
        } else {
            std::vector<Geqp3Task<float>> v((const Geqp3Task<float>*)d, (const Geqp3Task<float>*)d + count);
            geqp3_grouped<float>(v, tol, tie_rel, s, mem, d_err);
        }
        int herr = 0;
        hipError_t e = hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { api_error() = hipGetErrorString(e); rc = -100; }
        else if (herr) { api_error() = "cooperative geqp3: barrier wait exceeded its spin limit"; rc = -100; }
    } catch (std::exception& ex) {
        api_error() = ex.what();
        (void)hipStreamSynchronize(s);
        rc = -100;
    }
    for (void* p : owned) (void)hipFree(p);
    return rc;
}

int spaqr_spars_apply_batched(int prec, int count, const spaqr_spars_desc* d, void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    if (count == 0) return 0;
    if (prec == 64)
        return run_batched(d, sizeof(*d) * count, nullptr, 0, stream, [&](void* dd, void*, hipStream_t s) {
            launch_spars_apply<double>((const SparsApplyTask<double>*)dd, count, s);
        });
    // FP32: right-transform (pad 1) and left-transform (pad 2) tasks the C-resident kernels take
    std::vector<SparsApplyTask<float>> res[6], rest;
    for (int i = 0; i < count; ++i) {
        SparsApplyTask<float> t;
        static_assert(sizeof(t) == sizeof(d[i]), "spars desc layout");
        memcpy(&t, &d[i], sizeof(t));
        const int q = spars_res_bucket(t);
        (q < 0 ? rest : res[q]).push_back(t);
    }
    for (int q = 0; q < 7; ++q) {
        std::vector<SparsApplyTask<float>>& v = q < 6 ? res[q] : rest;
        if (v.empty()) continue;
        const int n = (int)v.size();
        const int rc = run_batched(v.data(), sizeof(SparsApplyTask<float>) * n, nullptr, 0, stream,
                                   [&](void* dd, void*, hipStream_t s) {
                                       if (q < 6) launch_spars_res((const SparsApplyTask<float>*)dd, n, 4 << (q % 3), q < 3, s);
                                       else launch_spars_apply<float>((const SparsApplyTask<float>*)dd, n, s);
                                   });
        if (rc) return rc;
    }
    return 0;
}

// ---------------------------------------------------------------- rank-path kernels
int spaqr_tree_set_rank_retries(spaqr_tree* h, int retries) {
    TREE_OR_FAIL(h);
    if (retries < 0 || retries > 2) { h->t->last_error = "rank retries must be 0, 1 or 2"; return -2; }
    h->t->rank_rounds = retries + 1;
    return 0;
}
int spaqr_tree_rank_log_clusters(const spaqr_tree* h, int* out, int cap) {
    if (!h || !h->t) return -1;
    const auto& o = h->t->rlog.order;
    const int n = (int)o.size();
    if (out)
        for (int i = 0; i < std::min(n, std::max(cap, 0)); ++i) out[i] = o[i];
    return n;
}

int spaqr_gauss_fill(int prec, void* om, int rows, int cols, int ld, uint64_t seed, void* stream) {
    if (prec != 16 && prec != 32 && prec != 64) return -1;
    if (!om && (size_t)rows * cols > 0) return -2;
    if (rows < 0 || cols < 0) return -3;
    if (ld < std::max(1, rows)) return -5;
    if ((size_t)rows * cols == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (prec == 64) launch_gauss<double>((double*)om, rows, cols, ld, seed, s);
    else if (prec == 32) launch_gauss<float>((float*)om, rows, cols, ld, seed, s);
    else launch_gauss_bf16(om, rows, cols, ld, seed, s);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { api_error() = hipGetErrorString(e); return -100; }
    return 0;
}

int spaqr_colnorm_batched(int prec, int count, const spaqr_colnorm_desc* d, void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    for (int i = 0; i < count; ++i)
        if (d[i].k < 1 || d[i].w < 1 || !d[i].C || !d[i].cn || !d[i].cmax2) return -3;
    if (count == 0) return 0;
    static_assert(sizeof(ColNormTask<double>) == sizeof(spaqr_colnorm_desc), "colnorm desc layout");
    return run_batched(d, sizeof(*d) * count, nullptr, 0, stream, [&](void* dd, void*, hipStream_t s) {
        if (prec == 64) launch_colnorm<double>((const ColNormTask<double>*)dd, count, s);
        else launch_colnorm<float>((const ColNormTask<float>*)dd, count, s);
    });
}

int spaqr_sketch_batched(int prec, int count, const spaqr_sketch_desc* d, const void* omega, int ldo, void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    if (!omega) return -4;
    int wmax = 0;
    for (int i = 0; i < count; ++i) {
        if (d[i].k < 1 || d[i].w < 1 || d[i].s < 1 || !d[i].C || !d[i].Y) return -3;
        wmax = std::max(wmax, d[i].w);
    }
    if (ldo < wmax) return -5;
    if (count == 0) return 0;
    static_assert(sizeof(SketchTask<double>) == sizeof(spaqr_sketch_desc), "sketch desc layout");
    std::vector<SketchTile> tiles;
    for (int i = 0; i < count; ++i)
        for (int r0 = 0; r0 < d[i].k; r0 += 64)
            for (int c0 = 0; c0 < d[i].s; c0 += 64) tiles.push_back({i, r0, c0, 0});
    const int nt = (int)tiles.size();
    return run_batched(d, sizeof(*d) * count, tiles.data(), sizeof(SketchTile) * tiles.size(), stream,
                       [&](void* dd, void* dt, hipStream_t s) {
                           if (prec == 64)
                               launch_sketch<double>((const SketchTask<double>*)dd, (const SketchTile*)dt, nt,
                                                     (const double*)omega, ldo, s);
                           else
                               launch_sketch<float>((const SketchTask<float>*)dd, (const SketchTile*)dt, nt,
                                                    (const float*)omega, ldo, s);
                       });
}

int spaqr_sketch_bf16_batched(int count, const spaqr_sketch_desc* d, const void* omega_bf16, int ldo, int tile_cols,
                              void* stream) {
    if (count < 0 || (count > 0 && !d)) return -2;
    if (!omega_bf16 || ((uintptr_t)omega_bf16 & 15) != 0) return -4;
    int wmax = 0;
    for (int i = 0; i < count; ++i) {
        if (d[i].k < 1 || d[i].w < 1 || d[i].s < 1 || !d[i].C || !d[i].Y) return -3;
        wmax = std::max(wmax, d[i].w);
    }
    if (ldo % 8 != 0 || ldo < (wmax + 31) / 32 * 32) return -5;
    if (tile_cols != 64 && tile_cols != 128 && tile_cols != 256) return -6;
    if (count == 0) return 0;
    std::vector<SketchTile> tiles;
    for (int i = 0; i < count; ++i)
        for (int r0 = 0; r0 < d[i].k; r0 += 64)
            for (int c0 = 0; c0 < d[i].s; c0 += tile_cols) tiles.push_back({i, r0, c0, 0});
    const int nt = (int)tiles.size();
    return run_batched(d, sizeof(*d) * count, tiles.data(), sizeof(SketchTile) * tiles.size(), stream,
                       [&](void* dd, void* dt, hipStream_t s) {
                           launch_sketch_bf16((const SketchTask<float>*)dd, (const SketchTile*)dt, nt, omega_bf16,
                                              ldo, tile_cols, s);
                       });
}

int spaqr_rank_features_batched(int prec, int count, const spaqr_rank_features_desc* d, double tol, void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    for (int i = 0; i < count; ++i)
        if (d[i].k < 1 || d[i].w < 1 || !d[i].Y16 || !d[i].cn || !d[i].cmax2 || !d[i].x) return -3;
    if (!(tol >= 0)) return -4;
    if (count == 0) return 0;
    static_assert(sizeof(FeatTask<double>) == sizeof(spaqr_rank_features_desc), "features desc layout");
    return run_batched(d, sizeof(*d) * count, nullptr, 0, stream, [&](void* dd, void*, hipStream_t s) {
        if (prec == 64) launch_rank_features<double>((const FeatTask<double>*)dd, count, tol, s);
        else launch_rank_features<float>((const FeatTask<float>*)dd, count, tol, s);
    });
}

int spaqr_tail_lead_batched(int prec, int count, const spaqr_tail_lead_desc* d, double tol, void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    for (int i = 0; i < count; ++i)
        if (d[i].r < 1 || d[i].w < 1 || !d[i].B || !d[i].cn || !d[i].cmax2 || !d[i].ry || !d[i].rank) return -3;
    if (!(tol >= 0)) return -4;
    if (count == 0) return 0;
    static_assert(sizeof(TailLeadTask<double>) == sizeof(spaqr_tail_lead_desc), "tail lead desc layout");
    return run_batched(d, sizeof(*d) * count, nullptr, 0, stream, [&](void* dd, void*, hipStream_t s) {
        if (prec == 64) launch_tail_lead<double>((const TailLeadTask<double>*)dd, count, tol, s);
        else launch_tail_lead<float>((const TailLeadTask<float>*)dd, count, tol, s);
    });
}

int spaqr_tail_gemm_batched(int count, const spaqr_tail_gemm_desc* d, void* stream) {
    if (count < 0 || (count > 0 && !d)) return -2;
    for (int i = 0; i < count; ++i)
        if (d[i].r < 1 || d[i].k < 1 || d[i].w < 1 || d[i].ldm < d[i].r || !d[i].M || !d[i].C || !d[i].B) return -3;
    if (count == 0) return 0;
    static_assert(sizeof(TailGemmTask) == sizeof(spaqr_tail_gemm_desc), "tail gemm desc layout");
    std::vector<SketchTile> tiles;
    for (int i = 0; i < count; ++i)
        for (int r0 = 0; r0 < d[i].r; r0 += 64)
            for (int c0 = 0; c0 < d[i].w; c0 += 64) tiles.push_back({i, r0, c0, 0});
    const int nt = (int)tiles.size();
    return run_batched(d, sizeof(*d) * count, tiles.data(), sizeof(SketchTile) * tiles.size(), stream,
                       [&](void* dd, void* dt, hipStream_t s) {
                           launch_tail_gemm((const TailGemmTask*)dd, (const SketchTile*)dt, nt, s);
                       });
}

int spaqr_tail_select_batched(int prec, int count, const spaqr_tail_desc* d, double tol, void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    for (int i = 0; i < count; ++i)
        if (d[i].k < 1 || d[i].w < 1 || !d[i].B || !d[i].cmax2 || !d[i].ry || !d[i].rank) return -3;
    if (!(tol >= 0)) return -4;
    if (count == 0) return 0;
    static_assert(sizeof(TailTask<double>) == sizeof(spaqr_tail_desc), "tail desc layout");
    return run_batched(d, sizeof(*d) * count, nullptr, 0, stream, [&](void* dd, void*, hipStream_t s) {
        if (prec == 64) launch_tail_select<double>((const TailTask<double>*)dd, count, tol, s);
        else launch_tail_select<float>((const TailTask<float>*)dd, count, tol, s);
    });
}

}  // extern "C"
