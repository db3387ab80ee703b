// Grouped geqp3 driver (see rrqr.hpp).
#include "rrqr.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "launch.hpp"

namespace spaqr {

template <typename S>
int geqp3_grouped(const std::vector<Geqp3Task<S>>& tasks, double tol, double tie_rel, hipStream_t stream,
                  const RrqrAlloc& mem, int* d_err, int* dispatches) {
    static int max_blocks_cache[2][2] = {{-1, -1}, {-1, -1}};
    static size_t lds_cache[2] = {0, 0};
    const int pi = sizeof(S) == 8 ? 0 : 1;
    size_t& lds = lds_cache[pi];
    if (max_blocks_cache[pi][0] < 0) {
        lds = geqp3_coop_lds_budget<S>(kGeqp3CoopBlocksPerCU);
        max_blocks_cache[pi][0] = lds ? geqp3_coop_max_blocks<S>(lds, false) : 0;
        max_blocks_cache[pi][1] = lds ? geqp3_coop_max_blocks<S>(lds, true) : 0;
    }
    std::vector<Geqp3Task<S>> fallback;
    struct Pending { std::vector<Geqp3CoopTask<S>> t; std::vector<int> b2t; int blocks = 0, jmax = 0, active = 0; };
    std::vector<int> zeros;
    std::vector<Pending> launches[2];  // [0] slices in LDS, [1] slices in HBM
    launches[0].emplace_back();
    launches[1].emplace_back();
    // A launch lasts as long as its longest pivot sequence, so interfaces of similar height
    // are packed together: largest first.
    std::vector<int> order(tasks.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return std::min(tasks[a].k, tasks[a].w) > std::min(tasks[b].k, tasks[b].w);
    });
    for (int oi : order) {
        const auto& g = tasks[oi];
        const int k = g.k, w = g.w;
        const int kmax = (tol >= 1.0) ? 0 : std::min(k, w);
        // LDS layout: k x cpb slice + vn1, vn2, pos per column; the pivot column
        const size_t per_col = (size_t)k * sizeof(S) + 2 * sizeof(S) + sizeof(int);
        const size_t fixed = (size_t)k * sizeof(S) + 64;
        int cpb = (lds > fixed) ? (int)((lds - fixed) / per_col) : 0;
        if (w > 0) cpb = std::min(cpb, w);
        int P = (w > 0 && cpb > 0) ? (w + cpb - 1) / cpb : 1;
        int mode = 0;
        if (cpb <= 0 || P > max_blocks_cache[pi][0]) {
            // HBM layout: only vn1, vn2, pos per column and the pivot column in LDS
            const int maxb = max_blocks_cache[pi][1];
            const size_t per_col_h = 2 * sizeof(S) + sizeof(int);
            const int cap = (lds > fixed) ? (int)((lds - fixed) / per_col_h) : 0;
            cpb = std::max(32, (w + std::max(maxb, 1) - 1) / std::max(maxb, 1));
            P = (w + cpb - 1) / cpb;
            mode = 1;
            if (maxb <= 0 || cpb > cap || P > maxb) {
                fallback.push_back(g);
                continue;
            }
        }
        const int maxb = max_blocks_cache[pi][mode];
        if (launches[mode].back().blocks + P > maxb) launches[mode].emplace_back();
        Pending& L = launches[mode].back();
        Geqp3CoopTask<S> c{};
        c.C = g.C;
        c.Vout = g.Vout;
        c.tau = g.tau;
        c.T = g.T;
        c.rank_out = g.rank_out;
        c.jpvt_out = g.jpvt;
        c.xch = P > 1 ? (S*)mem.alloc(sizeof(S) * geqp3_coop_xch_elems(k, P)) : nullptr;
        c.bar = nullptr;
        c.k = k;
        c.w = std::max(w, 1);
        c.b0 = L.blocks;
        c.P = P;
        c.cpb = std::max(cpb, 1);
        c.kmax = kmax;
        if (w == 0) { c.w = 0; }
        for (int b = 0; b < P; ++b) L.b2t.push_back((int)L.t.size());
        L.t.push_back(c);
        L.blocks += P;
        L.jmax = std::max(L.jmax, kmax);
        if (kmax > 0) L.active++;
    }
    static const bool dbg = getenv("SPAQR_DEBUG_RRQR") != nullptr;
    if (dbg) {
        for (int mode = 0; mode < 2; ++mode) {
            int kmx = 0, wmx = 0, pmx = 0, nt = 0;
            for (auto& L : launches[mode])
                for (auto& c : L.t) { kmx = std::max(kmx, c.k); wmx = std::max(wmx, c.w); pmx = std::max(pmx, c.P); ++nt; }
            if (!nt) continue;
            fprintf(stderr, "[geqp3%s] tasks %d maxb %d lds %zu launches %zu fallback %zu  max k %d w %d P %d\n",
                    mode ? " hbm" : "", nt, max_blocks_cache[pi][mode], lds, launches[mode].size(), fallback.size(),
                    kmx, wmx, pmx);
        }
        for (auto& f : fallback) fprintf(stderr, "   fallback k %d w %d\n", f.k, f.w);
        fflush(stderr);
    }
    // all tables are uploaded before the first dispatch so that the launches run back to back
    size_t ntask = 0;
    for (int mode = 0; mode < 2; ++mode)
        for (auto& L : launches[mode]) ntask += L.t.size();
    zeros.assign(std::max<size_t>(ntask, 1), 0);
    int* d_bars = (int*)mem.upload(zeros.data(), zeros.size() * sizeof(int));
    if (mem.graph) launch_fill<int>(d_bars, zeros.size(), 0, stream);
    struct Ready { Geqp3CoopTask<S>* t; int* b2t; int blocks; bool hbm; };
    std::vector<Ready> ready;
    size_t bar_off = 0;
    for (int mode = 0; mode < 2; ++mode)
        for (auto& L : launches[mode]) {
            if (L.blocks == 0) continue;
            for (size_t i = 0; i < L.t.size(); ++i) L.t[i].bar = d_bars + bar_off + i;
            bar_off += L.t.size();
            auto* d_t = (Geqp3CoopTask<S>*)mem.upload(L.t.data(), L.t.size() * sizeof(Geqp3CoopTask<S>));
            auto* d_b = (int*)mem.upload(L.b2t.data(), L.b2t.size() * sizeof(int));
            ready.push_back({d_t, d_b, L.blocks, mode == 1});
        }
    Geqp3Task<S>* d_fb = fallback.empty()
                             ? nullptr
                             : (Geqp3Task<S>*)mem.upload(fallback.data(), fallback.size() * sizeof(Geqp3Task<S>));
    for (const auto& r : ready) {
        hipError_t e = launch_geqp3_coop<S>(r.t, r.b2t, d_err, r.blocks, tol, tie_rel, lds, r.hbm, stream, !mem.graph);
        if (e != hipSuccess)
            throw std::runtime_error(std::string("cooperative geqp3 launch: ") + hipGetErrorString(e) + " (blocks " +
                                     std::to_string(r.blocks) + ", dynamic LDS " + std::to_string(lds) + " B, " +
                                     (r.hbm ? "HBM" : "LDS") + " slices, co-resident limit " +
                                     std::to_string(max_blocks_cache[pi][r.hbm ? 1 : 0]) + ")");
    }
    if (d_fb) launch_geqp3<S>(d_fb, (int)fallback.size(), tol, tie_rel, stream);
    if (dispatches) *dispatches = (int)ready.size() + (d_fb ? 1 : 0);
    return (int)fallback.size();
}

template int geqp3_grouped<double>(const std::vector<Geqp3Task<double>>&, double, double, hipStream_t, const RrqrAlloc&, int*, int*);
template int geqp3_grouped<float>(const std::vector<Geqp3Task<float>>&, double, double, hipStream_t, const RrqrAlloc&, int*, int*);

}  // namespace spaqr
