// Definitions shared by the translation units of the level planner (tree.cpp, tree_rank.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <thread>
#include <stdexcept>
#include <string>
#include <vector>

#include "tree.hpp"

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t _e = (x);                                                                         \
        if (_e != hipSuccess)                                                                        \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " #x); \
    } while (0)

namespace spaqr {

// geqp3 pivot candidates within this relative band of the largest norm count as tied and
// the lowest index wins (LAPACK's plain argmax lets rounding order decide exact ties).
constexpr double kPivotTieRel = 1e-10;

// Fronts with at least this many columns run their panel loop from the host (geqrt_big,
// tpqrt_big) so that the trailing updates spread over many workgroups.
constexpr int kBigFront = 96;

// Splits larfb work so that a large block is spread over many workgroups: Q^T C acts on each
// column independently, so tasks are cut into column chunks; a fused trailing C R^{-1}
// mixes columns, so it is moved to a second list of row chunks (V = nullptr: trsm only),
// to run after the first.
template <typename S>
void split_larfb(const std::vector<LarfbTask<S>>& in, std::vector<LarfbTask<S>>& cols_out,
                 std::vector<LarfbTask<S>>& trsm_out, int col_chunk = 128, int row_chunk = sizeof(S) == 4 ? 128 : 256) {
    for (const auto& t : in) {
        // (FP32 blocks of more than 512 rows: 64-column chunks, the two-half resident kernel's width)
        const int cw = (sizeof(S) == 4 && t.m > 512) ? std::min(col_chunk, 64) : col_chunk;
        if (t.V)
            for (int c0 = 0; c0 < t.n; c0 += cw) {
                LarfbTask<S> u = t;
                u.C = t.C + (size_t)c0 * t.ldc;
                u.n = std::min(cw, t.n - c0);
                u.R = nullptr;
                u.k2 = 0;
                cols_out.push_back(u);
            }
        if (t.R)
            for (int r0 = 0; r0 < t.m; r0 += row_chunk) {
                LarfbTask<S> u = t;
                u.V = nullptr;
                u.T = nullptr;
                u.C = t.C + r0;
                u.m = std::min(row_chunk, t.m - r0);
                trsm_out.push_back(u);
            }
    }
}

// Bucket of the C-resident FP32 larfb kernels for a task (0: m <= 128, 1: m <= 256, 2: m <= 512,
// all with n <= 128; 3: m <= 1024 with n <= 64), or -1 when the task stays on the workgroup
// kernel (FP64, a trsm, larger blocks).  SPAQR_LARFB_RES=0 keeps every task on the workgroup kernel.
template <typename S>
inline int larfb_res_bucket(const LarfbTask<S>& t) {
    static const bool on = getenv("SPAQR_LARFB_RES") == nullptr || atoi(getenv("SPAQR_LARFB_RES")) != 0;
    if (!on || sizeof(S) != 4 || !t.V || t.R || t.n > 128 || t.m > 1024) return -1;
    if (t.m > 512) return t.n <= 64 ? 3 : -1;
    return t.m <= 128 ? 0 : (t.m <= 256 ? 1 : 2);
}

// Bucket of the C-resident FP32 tpmqrt kernels for a chain task (0: links of <= 128 rows,
// 1: <= 256, 2: <= 512, 3: <= 1024), or -1 when it stays on the workgroup kernel (FP64, a tile
// wider than 64 columns or a longer link).  SPAQR_TPMQRT_RES=0 keeps every task on the workgroup kernel.
template <typename S>
inline int tpmqrt_res_bucket_r(int ncols, int rmax) {  // rmax: the longest link's rows
    static const bool on = getenv("SPAQR_TPMQRT_RES") == nullptr || atoi(getenv("SPAQR_TPMQRT_RES")) != 0;
    if (!on || sizeof(S) != 4 || ncols > 64 || rmax > 1024) return -1;
    return rmax <= 128 ? 0 : (rmax <= 256 ? 1 : (rmax <= 512 ? 2 : 3));
}
template <typename S>
inline int tpmqrt_res_bucket(const TpmqrtTask<S>& t, const TpmqrtLink<S>* links) {
    int rmax = 0;
    for (int j = 0; j < t.nlinks; ++j) rmax = std::max(rmax, links[t.link0 + j].rows);
    return tpmqrt_res_bucket_r<S>(t.ncols, rmax);
}

// Bucket of the C-resident FP32 sparsify kernels for a task: 0..2 (block rows <= 128 / 256 / 512)
// for the right transform (pad 1: kn rows of A^T, mc <= 128), 3..5 for the left transform with
// its scatter (pad 2 with Q_c: kc rows, rn <= 128), -1 for the workgroup kernel.
// SPAQR_SPARS_RES=0 keeps every task on the workgroup kernel.
template <typename S>
inline int spars_res_bucket(const SparsApplyTask<S>& t) {
    static const bool on = getenv("SPAQR_SPARS_RES") == nullptr || atoi(getenv("SPAQR_SPARS_RES")) != 0;
    if (!on || sizeof(S) != 4 || t.ident) return -1;
    auto b = [](int m) { return m <= 128 ? 0 : (m <= 256 ? 1 : 2); };
    if (t.pad == 1 && t.Vn && t.kn <= 512 && t.mc <= 128) return b(t.kn);
    if (t.pad == 2 && t.Vc && t.kc <= 512 && t.rn <= 128) return 3 + b(t.kc);
    return -1;
}

// Bucket of the C-resident FP32 trsm kernel for a C R^{-1}-only task (0: k2 <= 128, 1: <= 256,
// 2: <= 512; rows m <= 128), or -1 for the workgroup kernel.  SPAQR_TRSM_RES=0 turns it off.
template <typename S>
inline int trsm_res_bucket(const LarfbTask<S>& t) {
    static const bool on = getenv("SPAQR_TRSM_RES") == nullptr || atoi(getenv("SPAQR_TRSM_RES")) != 0;
    if (!on || sizeof(S) != 4 || t.V || !t.R || t.m > 128 || t.k2 > 512) return -1;
    return t.k2 <= 128 ? 0 : (t.k2 <= 256 ? 1 : 2);
}

// FP32 coupling blocks taller than kLinkRows are factored as several links, one per row slice
// of at most kLinkRows rows (balanced), each with its own T: the QR of [R; B_0; B_1; ...] by the
// same chain of tpqrt steps, with every panel inside the register-resident panel QR.  The same
// slices serve the trailing update and the solve.  FP64 keeps whole blocks, as the reference.
constexpr int kLinkRows = 1024;
template <typename S>
inline int link_slices(int rows) {
    return (sizeof(S) == 4 && rows > kLinkRows) ? (rows + kLinkRows - 1) / kLinkRows : 1;
}

// Flops of an m x n Householder QR.
template <typename S>
inline double flops_geqrt(double m, double n) { return 2 * m * n * n - 2 * n * n * n / 3; }

// Host threads for the planner loops: OMP_NUM_THREADS when set, else the hardware count,
// at most 16.
inline int host_threads() {
    static const int n = [] {
        const char* e = getenv("OMP_NUM_THREADS");
        int v = e ? atoi(e) : 0;
        if (v <= 0) v = (int)std::thread::hardware_concurrency();
        return std::max(1, std::min(v, 16));
    }();
    return n;
}

// Open-addressing set of 64-bit keys (no removal; the all-ones key is reserved), sized for
// n keys: the planner's duplicate checks without a node allocation per insert.
class FlatKeySet {
  public:
    explicit FlatKeySet(size_t n) {
        size_t c = 16;
        bits_ = 4;
        while (c < 2 * n + 16) {
            c <<= 1;
            ++bits_;
        }
        slots_.assign(c, kEmpty);
    }
    bool insert(unsigned long long k) {  // true when k was not in the set
        const size_t mask = slots_.size() - 1;
        size_t h = (size_t)((k * 0x9E3779B97F4A7C15ull) >> (64 - bits_));
        while (true) {
            if (slots_[h] == kEmpty) {
                slots_[h] = k;
                return true;
            }
            if (slots_[h] == k) return false;
            h = (h + 1) & mask;
        }
    }

  private:
    static constexpr unsigned long long kEmpty = ~0ull;
    std::vector<unsigned long long> slots_;
    int bits_;
};

// Workgroups per copy task: enough that the largest block of the launch is spread at about
// 16 tiles of 32 x 32 per workgroup.
template <typename S>
inline int copy_parts(const std::vector<CopyTask<S>>& v) {
    long most = 0;
    for (const auto& t : v) most = std::max(most, (long)((t.rows + 31) / 32) * ((t.cols + 31) / 32));
    return (int)std::max(1L, std::min(64L, (most + 15) / 16));
}

// Persistent workers for parallel_ranges: a planner level makes several short parallel passes,
// and starting threads for each cost more than the passes.  run(T, job) runs job(t) for t in
// [1, T) on the workers and job(0) on the caller, and returns when all have finished; callers
// are serialized.
// The pool is a heap object that no static destructor touches: its workers are joined by
// shutdown() (spaqr_shutdown), and a process that exits without it simply ends them.
class HostPool {
  public:
    static HostPool& get() {
        std::lock_guard<std::mutex> lk(inst_mutex());
        HostPool*& p = inst();
        if (!p) p = new HostPool(host_threads() - 1);
        return *p;
    }
    // Joins the workers; a later parallel pass starts a new pool.  No pass may be running.
    static void shutdown() {
        HostPool* p = nullptr;
        {
            std::lock_guard<std::mutex> lk(inst_mutex());
            p = inst();
            inst() = nullptr;
        }
        delete p;
    }
    void run(int T, const std::function<void(int)>& job) {
        std::lock_guard<std::mutex> serial(call_);
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = &job;
            T_ = T;
            pending_ = T - 1;
            ++gen_;
        }
        cv_.notify_all();
        job(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return pending_ == 0; });
        job_ = nullptr;
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }

  private:
    static HostPool*& inst() {
        static HostPool* p = nullptr;
        return p;
    }
    static std::mutex& inst_mutex() {
        static std::mutex* m = new std::mutex;  // never destroyed: usable in any exit order
        return *m;
    }
    explicit HostPool(int workers) {
        for (int w = 1; w <= workers; ++w) th_.emplace_back([this, w] { loop(w); });
    }
    void loop(int w) {
        unsigned long long seen = 0;
        for (;;) {
            const std::function<void(int)>* job;
            int T;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                job = job_;
                T = T_;
            }
            if (w < T) {
                (*job)(w);
                std::lock_guard<std::mutex> lk(m_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }
    std::mutex call_, m_;
    std::condition_variable cv_, done_;
    const std::function<void(int)>* job_ = nullptr;
    int T_ = 0, pending_ = 0;
    unsigned long long gen_ = 0;
    bool stop_ = false;
    std::vector<std::thread> th_;
};

// Runs f(begin, end, slot) over [0, n) cut into contiguous ranges (at least min_per items
// each) on up to host_threads() threads.  Slot t gets the t-th range, so per-slot results
// concatenated in slot order reproduce the sequential order.  Returns the number of slots;
// an exception in any range is rethrown after all have finished.
template <class F>
int parallel_ranges(int n, int min_per, F&& f) {
    const int T = std::min(host_threads(), std::max(1, n / std::max(1, min_per)));
    if (T <= 1) {
        f(0, n, 0);
        return 1;
    }
    std::vector<std::exception_ptr> err(T);
    const std::function<void(int)> run = [&](int t) {
        try {
            f((int)((long long)n * t / T), (int)((long long)n * (t + 1) / T), t);
        } catch (...) {
            err[t] = std::current_exception();
        }
    };
    HostPool::get().run(T, run);
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
    return T;
}

template <typename T, size_t NBK>
std::array<std::pair<T*, int>, NBK> Tree::upload_buckets(const std::vector<std::array<std::vector<T>, NBK>>& parts) {
    const int np = (int)parts.size();
    std::array<std::vector<size_t>, NBK> off;
    std::array<T*, NBK> hb{}, db{};
    std::array<size_t, NBK> nq{};
    for (size_t q = 0; q < NBK; ++q) {
        off[q].assign(np + 1, 0);
        for (int i = 0; i < np; ++i) off[q][i + 1] = off[q][i] + parts[i][q].size();
        nq[q] = off[q][np];
        if (nq[q]) {
            db[q] = (T*)desc_arena.alloc(nq[q] * sizeof(T));
            hb[q] = (T*)staging.alloc(nq[q] * sizeof(T));
        }
    }
    parallel_ranges(np, 1, [&](int b, int e, int) {
        for (int i = b; i < e; ++i)
            for (size_t q = 0; q < NBK; ++q)
                if (!parts[i][q].empty()) memcpy(hb[q] + off[q][i], parts[i][q].data(), parts[i][q].size() * sizeof(T));
    });
    std::array<std::pair<T*, int>, NBK> out;
    for (size_t q = 0; q < NBK; ++q) {
        if (nq[q]) h2d(db[q], hb[q], nq[q] * sizeof(T));
        out[q] = {db[q], (int)nq[q]};
    }
    return out;
}

// Splits off the tasks whose block fits one wave: dim(t) is the block's register length
// (<= 64) or -1 to keep the task on the workgroup kernels; small tasks are bucketed by the
// register length classes 16 / 32 / 64 of the one-wave kernels.
template <typename T, typename D>
void split_small(const std::vector<T>& in, D dim, std::vector<T> (&small)[3], std::vector<T>& big) {
    for (const auto& t : in) {
        const int d = dim(t);
        if (d < 0) big.push_back(t);
        else small[d <= 16 ? 0 : (d <= 32 ? 1 : 2)].push_back(t);
    }
}

template <class F>
void Tree::timed(int cat, long launches, long tasks, F&& f) {
    if (!profiling || capturing) {  // events are not recorded into graphs
        f();
        return;
    }
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, stream_));
    f();
    HIPCHK(hipEventRecord(b, stream_));
    prof_pending.push_back({cat, {a, b}});
    prof[cat].launches += launches;
    prof[cat].tasks += tasks;
}

}  // namespace spaqr
