// Sparsification rank path of the level planner: the reference's geqp3 on the gathered C
// (src/ptree.cpp:373-439), or the randomized path of learned.hpp -- 16-column sketch
// features, the predictor's sketch width, pivoted QR of the sketch, the tail test and the
// per-interface geqp3 fallback -- plus the host side of the predictor's training.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "launch.hpp"
#include "learned.hpp"
#include "rrqr.hpp"
#include "status.hpp"
#include "tree_impl.hpp"

namespace spaqr {

namespace {
constexpr double kSketchGrow = 1.25;      // sketch width = grow * predicted rank + pad
constexpr int kSketchPad = 8;
// The sketch's pivoted QR runs to its full width s: its own |R_jj| / |R_00| test is relative to
// the sketch's largest column, which can sit far above the largest column of C that the tail
// test measures against, so stopping it early only causes fallbacks.
constexpr double kSketchTolFactor = 0.0;
constexpr uint64_t kSketchSeed = 0x5eed5ca1eull;
// Sketches at least this wide are QR'd without pivoting (blocked Householder on MFMA): any
// prefix of a Gaussian sketch spans the dominant range, so the tail test finds a rank a few
// columns above the pivoted one, while the pivoted QR would pay one cross-workgroup barrier
// per column.  SPAQR_SKETCH_UNPIVOTED overrides the width.
constexpr int kSketchUnpivoted = 96;
}  // namespace

int Tree::set_rank_model(const float* w) {
    if (!d_mlp_w) HIPCHK(hipMalloc(&d_mlp_w, sizeof(float) * kMlpParams));
    HIPCHK(hipMemcpy(d_mlp_w, w, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    has_model = true;
    return 0;
}

template <typename S>
const S* Tree::omega(int rows, int cols) {
    rows = std::max(rows, 1);
    cols = std::max(cols, 1);
    if (!d_omega || rows > om_rows || cols > om_cols || om_elem != (int)sizeof(S)) {
        sync_stream();  // the previous matrix may still be read by queued work
        if (d_omega) (void)hipFree(d_omega);
        d_omega = nullptr;
        const int r2 = std::max(rows, om_elem == (int)sizeof(S) ? om_rows : 0);
        const int c2 = std::max(cols, om_elem == (int)sizeof(S) ? om_cols : 0);
        HIPCHK(hipMalloc(&d_omega, sizeof(S) * (size_t)r2 * c2));
        launch_gauss<S>((S*)d_omega, r2, c2, r2, kSketchSeed, stream_);
        om_rows = r2;
        om_cols = c2;
        om_elem = (int)sizeof(S);
    }
    return (const S*)d_omega;
}

const void* Tree::omega_bf16(int rows, int cols) {
    rows = (std::max(rows, 1) + 31) / 32 * 32;
    cols = std::max(cols, 1);
    if (!d_omega_bf || rows > ob_rows || cols > ob_cols) {
        sync_stream();  // the previous matrix may still be read by queued work
        if (d_omega_bf) (void)hipFree(d_omega_bf);
        d_omega_bf = nullptr;
        const int r2 = std::max(rows, ob_rows), c2 = std::max(cols, ob_cols);
        HIPCHK(hipMalloc(&d_omega_bf, 2 * (size_t)r2 * c2));
        launch_gauss_bf16(d_omega_bf, r2, c2, r2, kSketchSeed, stream_);
        ob_rows = r2;
        ob_cols = c2;
    }
    return d_omega_bf;
}

template <typename S, typename Task, typename Tile>
void Tree::run_sketch(const Task* d, const Tile* tiles, int nt, const S* om, const void* omb, int tcols) {
    if constexpr (sizeof(S) == 4) {
        if (omb) {
            launch_sketch_bf16(d, tiles, nt, omb, ob_rows, tcols, stream_);
            return;
        }
    }
    launch_sketch<S>(d, tiles, nt, om, om_rows, stream_);
}

template <typename S>
double Tree::rrqr_level(std::vector<Geqp3Task<S>>& gq, const std::vector<CopyTask<S>>& gat,
                        const std::vector<int>& gat_begin, const std::vector<int>& gq_cluster, int L, int* d_ranks,
                        int* d_err, std::vector<char>& used_sketch) {
    const int nt = (int)gq.size();
    const double SZ = sizeof(S);
    double flops = 0.0;
    RrqrAlloc mem{[&](size_t b) { return arena.alloc(b); }, [&](const void* h, size_t b) { return upload(h, b); }};
    const bool sketch_mode = rank_mode == 1 || rank_mode == 2;
    if (rank_mode == 1 && !has_model)
        throw std::runtime_error("rank mode 1 needs predictor weights (spaqr_tree_set_rank_model)");
    const bool feats = rank_log || rank_mode == 1;
    std::vector<int> elig;  // interfaces wide enough for a 16-column sketch
    std::vector<char> is_elig(nt, 0);
    if (tol > 0.0 && tol < 1.0)
        for (int q = 0; q < nt; ++q)
            if (std::min(gq[q].k, gq[q].w) > kSketchFeatureCols) {
                elig.push_back(q);
                is_elig[q] = 1;
            }
    const int ne = (int)elig.size();
    int wmax = 1;
    for (int q : elig) wmax = std::max(wmax, gq[q].w);
    S* d_cmax2 = (S*)arena.alloc(sizeof(S) * std::max(nt, 1));
    float* d_X = nullptr;
    const int kColChunk = tuned(T_RANK_COLS, 128);  // columns per workgroup of the column-norm and tail kernels
    // squared column norms of every eligible C, kept for the features and the tail test (the
    // pivoted sketch QRs reuse the geqp3 norm workspace)
    std::vector<S*> cnorm(nt, nullptr);
    if (ne && (feats || sketch_mode)) {
        size_t tot_w = 0;
        for (int q : elig) tot_w += (size_t)gq[q].w;
        S* cbuf = (S*)arena.alloc(sizeof(S) * std::max<size_t>(tot_w, 1));
        for (int q : elig) {
            cnorm[q] = cbuf;
            cbuf += gq[q].w;
        }
        std::vector<ColNormTask<S>> cn;
        for (int q : elig)
            for (int c0 = 0; c0 < gq[q].w; c0 += kColChunk)
                cn.push_back({gq[q].C + (size_t)c0 * gq[q].k, cnorm[q] + c0, d_cmax2 + q, gq[q].k,
                              std::min(kColChunk, gq[q].w - c0)});
        auto* d_cn = upload_vec(cn);
        timed(P_RSKETCH, 2, ne, [&] {
            launch_fill<S>(d_cmax2, (size_t)nt, S(0), stream_);
            launch_colnorm<S>(d_cn, (int)cn.size(), stream_);
        });
        for (int q : elig) prof_add(P_RSKETCH, 2.0 * gq[q].k * gq[q].w, SZ * gq[q].k * gq[q].w);
    }
    if (ne && feats) {
        const bool bf = sketch_bf16<S>();
        const S* om = bf ? nullptr : omega<S>(wmax, kSketchFeatureCols);
        const void* omb = bf ? omega_bf16(wmax, kSketchFeatureCols) : nullptr;
        std::vector<SketchTask<S>> sk;
        std::vector<SketchTile> tiles;
        std::vector<GeqrtTask<S>> qr;
        std::vector<FeatTask<S>> ft;
        d_X = (float*)arena.alloc(sizeof(float) * kRankFeatures * ne);
        for (int i = 0; i < ne; ++i) {
            const int q = elig[i], k = gq[q].k, w = gq[q].w;
            S* Y16 = (S*)arena.alloc(sizeof(S) * (size_t)k * kSketchFeatureCols);
            S* T16 = (S*)arena.alloc(sizeof(S) * NB * kSketchFeatureCols);
            for (int r0 = 0; r0 < k; r0 += 64) tiles.push_back({(int)sk.size(), r0, 0, 0});
            sk.push_back({gq[q].C, Y16, k, w, kSketchFeatureCols, 0});
            qr.push_back({Y16, T16, k, std::min(kSketchFeatureCols, k), k, 0});
            const Cluster& C = cl[gq_cluster[q]];
            ft.push_back({Y16, cnorm[q], d_cmax2 + q, d_X + (size_t)i * kRankFeatures, k, w, (int)C.sout.size(),
                          (int)C.sin.size(), L, 0});
            prof_add(P_RSKETCH, 2.0 * k * w * kSketchFeatureCols, SZ * ((double)k * w + 2.0 * k * kSketchFeatureCols));
        }
        auto* d_sk = upload_vec(sk);
        auto* d_tl = upload_vec(tiles);
        auto* d_qr = upload_vec(qr);
        auto* d_ft = upload_vec(ft);
        timed(P_RSKETCH, 3, ne, [&] {
            run_sketch(d_sk, d_tl, (int)tiles.size(), om, omb, 64);
            launch_geqrt<S>(d_qr, ne, stream_);
            launch_rank_features<S>(d_ft, ne, tol, stream_);
        });
    }
    std::vector<int> rk(nt, -1), ry(nt, 0), s_of(nt, 0);
    std::vector<char> applied(nt, 0);
    if (ne && sketch_mode) {
        if (rank_mode == 1) {
            float* d_frac = (float*)arena.alloc(sizeof(float) * ne);
            timed(P_RSKETCH, 1, ne, [&] { launch_mlp_predict(d_X, ne, d_mlp_w, d_frac, stream_); });
            std::vector<float> frac(ne);
            const double grow = tuned(T_SKETCH_GROW, (int)std::lround(100 * kSketchGrow)) / 100.0;
            const int pad = tuned(T_SKETCH_PAD, kSketchPad + 1) - 1;  // stored + 1: 0 means built-in
            HIPCHK(hipMemcpyAsync(frac.data(), d_frac, sizeof(float) * ne, hipMemcpyDeviceToHost, stream_));
            sync_stream();
            for (int i = 0; i < ne; ++i) {
                const int q = elig[i], kw = std::min(gq[q].k, gq[q].w);
                const double f = std::isfinite(frac[i]) ? std::min(1.0f, std::max(0.0f, frac[i])) : 1.0;
                s_of[q] = std::max(1, std::min(kw, (int)std::ceil(grow * f * kw) + pad));
            }
        } else {
            for (int q : elig) s_of[q] = std::min(gq[q].k, gq[q].w);
        }
        // Rounds: sketch, pivoted QR of the sketch, Q^T C and the tail test.  An interface whose
        // tail test fails is gathered again and retried with a sketch twice as wide (a sketch as
        // wide as min(k, w) spans the range of C, so the last round can only fail by rounding).
        std::vector<int> todo = elig;
        const int rounds = std::max(1, std::min(3, rank_rounds));
        for (int round = 0; round < rounds && !todo.empty(); ++round) {
            int smax = 1;
            for (int q : todo) smax = std::max(smax, s_of[q]);
            const bool bf = sketch_bf16<S>();
            const int tcols = bf ? tuned(T_SKETCH_TILE, 128) : 64;
            const S* om = bf ? nullptr : omega<S>(wmax, smax);
            const void* omb = bf ? omega_bf16(wmax, smax) : nullptr;
            if (round > 0) {  // C was overwritten by the previous round's Q^T C
                std::vector<CopyTask<S>> regather;
                for (int q : todo)
                    for (int g = gat_begin[q]; g < gat_begin[q + 1]; ++g) regather.push_back(gat[g]);
                auto* d_rg = upload_vec(regather);
                timed(P_GATHER, 1, (long)regather.size(), [&] { launch_copy<S>(d_rg, (int)regather.size(), stream_, copy_parts(regather)); });
            }
            static const int unpiv_env = getenv("SPAQR_SKETCH_UNPIVOTED") ? atoi(getenv("SPAQR_SKETCH_UNPIVOTED"))
                                                                            : kSketchUnpivoted;
            const int unpiv_min = tuned(T_SKETCH_UNPIVOTED, unpiv_env);
            std::vector<SketchTask<S>> sk;
            std::vector<SketchTile> tiles;
            std::vector<Geqp3Task<S>> gy;
            std::vector<GeqrtTask<S>> uy_small, uy_big;  // unpivoted sketch QRs, in place in Vout
            std::vector<int> ry0(nt, 0);
            for (int q : todo) {
                const int k = gq[q].k, w = gq[q].w, s = s_of[q];
                const bool unpiv = unpiv_min > 0 && s >= unpiv_min;
                S* Y = unpiv ? gq[q].Vout : (S*)arena.alloc(sizeof(S) * (size_t)k * s);
                for (int r0 = 0; r0 < k; r0 += 64)
                    for (int c0 = 0; c0 < s; c0 += tcols) tiles.push_back({(int)sk.size(), r0, c0, 0});
                sk.push_back({gq[q].C, Y, k, w, s, 0});
                if (unpiv) {
                    (std::min(k, s) >= tuned(T_BIG_FRONT, kBigFront) ? uy_big : uy_small).push_back({Y, gq[q].T, k, s, k, 0});
                    ry0[q] = std::min(k, s);
                    prof_add(P_GEQP3, flops_geqrt<S>(k, s), SZ * (2.0 * k * s + NB * s));
                } else {
                    Geqp3Task<S> g = gq[q];
                    g.C = Y;
                    g.w = s;
                    gy.push_back(g);
                    prof_add(P_GEQP3, 0.0, SZ * 3.0 * k * s);
                }
                flops += 2.0 * k * w * s;
                prof_add(P_RSKETCH, 2.0 * k * w * s, SZ * ((double)k * w + (double)k * s));
                rstats.sketch_cols += s;
                if (round == 0) rstats.full_cols += std::min(k, w);
            }
            auto* d_sk = upload_vec(sk);
            auto* d_tl = upload_vec(tiles);
            timed(P_RSKETCH, 1, (long)todo.size(),
                  [&] { run_sketch(d_sk, d_tl, (int)tiles.size(), om, omb, tcols); });
            // rank bounds of the unpivoted QRs (full width); the pivoted QRs write their own
            if (!uy_small.empty() || !uy_big.empty())
                HIPCHK(hipMemcpyAsync(d_ranks, upload(ry0.data(), sizeof(int) * nt), sizeof(int) * nt,
                                      hipMemcpyDeviceToDevice, stream_));
            int nd = 0;
            timed(P_GEQP3, 0, (long)todo.size(), [&] {
                auto* d_uy = upload_vec(uy_small);
                launch_geqrt<S>(d_uy, (int)uy_small.size(), stream_);
                nd = !uy_small.empty() + geqrt_big<S>(uy_big);
                int ndp = 0;
                if (!gy.empty()) geqp3_grouped<S>(gy, tol * kSketchTolFactor, kPivotTieRel, stream_, mem, d_err, &ndp);
                nd += ndp;
            });
            if (profiling) prof[P_GEQP3].launches += nd;
            HIPCHK(hipMemcpyAsync(ry.data(), d_ranks, sizeof(int) * nt, hipMemcpyDeviceToHost, stream_));
            sync_stream();
            std::vector<LarfbTask<S>> lb;
            std::vector<TailTask<S>> tt;
            // FP32 interfaces of more than 1024 rows (past the C-resident applies, whose
            // workgroup kernel re-reads C for every panel): the tail test needs only B's first r
            // rows, Q(:, 0:r)^T C, and the column norms of C, so Q^T is formed on the identity (k
            // columns instead of w) and its leading rows multiply C as one MFMA product.
            // SPAQR_TAIL_GEMM=0 applies Q^T to all of C for every interface.
            static const bool tail_gemm_env = getenv("SPAQR_TAIL_GEMM") == nullptr || atoi(getenv("SPAQR_TAIL_GEMM")) != 0;
            const bool tail_gemm = sizeof(S) == 4 && tail_gemm_env;
            std::vector<EyeTask<S>> eye;
            std::vector<TailGemmTask> tg;
            std::vector<SketchTile> tg_tiles;
            std::vector<TailLeadTask<S>> tlead;
            int* d_tail = (int*)arena.alloc(sizeof(int) * std::max(nt, 1));
            for (int q : todo) {
                const int k = gq[q].k, w = gq[q].w, r = ry[q];
                for (int j = 0; j < r; ++j) flops += 4.0 * (k - j) * (s_of[q] - j);
                applied[q] = 0;
                if (r <= 0) continue;
                if (tail_gemm && cnorm[q] && k > 1024) {
                    S* M = (S*)arena.alloc(sizeof(S) * (size_t)k * k);
                    S* Bl = (S*)arena.alloc(sizeof(S) * (size_t)r * w);
                    eye.push_back({M, k, k});
                    lb.push_back({gq[q].Vout, gq[q].T, M, nullptr, k, k, r, k, k, 0, 0, 0});
                    for (int r0 = 0; r0 < r; r0 += 64)
                        for (int c0 = 0; c0 < w; c0 += 64) tg_tiles.push_back({(int)tg.size(), r0, c0, 0});
                    tg.push_back({(const float*)M, (const float*)gq[q].C, (float*)Bl, r, k, w, k});
                    for (int c0 = 0; c0 < w; c0 += kColChunk)
                        tlead.push_back({Bl + (size_t)c0 * r, cnorm[q] + c0, d_cmax2 + q, d_ranks + q, d_tail + q, r,
                                         std::min(kColChunk, w - c0)});
                    flops += 4.0 * k * k * r - 2.0 * k * (double)r * r + 2.0 * r * (double)k * w;
                    prof_add(P_RSKETCH, 4.0 * k * k * r - 2.0 * k * (double)r * r + 2.0 * r * (double)k * w,
                             SZ * ((double)k * r + 2.0 * k * k + (double)k * w + (double)r * w));
                } else {
                    lb.push_back({gq[q].Vout, gq[q].T, gq[q].C, nullptr, k, w, r, k, k, 0, 0, 0});
                    for (int c0 = 0; c0 < w; c0 += kColChunk)
                        tt.push_back({gq[q].C + (size_t)c0 * k, d_cmax2 + q, d_ranks + q, d_tail + q, k,
                                      std::min(kColChunk, w - c0)});
                    flops += 4.0 * k * w * r - 2.0 * w * (double)r * r;
                    prof_add(P_RSKETCH, 4.0 * k * w * r - 2.0 * w * (double)r * r, SZ * ((double)k * r + 2.0 * k * w));
                }
                applied[q] = 1;
            }
            std::vector<int> again;
            if (!lb.empty()) {
                std::vector<LarfbTask<S>> lb_c, lb_t, lb_res[4], lb_rest;
                split_larfb(lb, lb_c, lb_t);
                for (const auto& u : lb_c) {
                    const int q = larfb_res_bucket(u);
                    (q < 0 ? lb_rest : lb_res[q]).push_back(u);
                }
                auto* d_lb = upload_vec(lb_rest);
                LarfbTask<S>* d_lbr[4];
                for (int q = 0; q < 4; ++q) d_lbr[q] = upload_vec(lb_res[q]);
                auto* d_tt = upload_vec(tt);
                auto* d_eye = upload_vec(eye);
                auto* d_tg = upload_vec(tg);
                auto* d_tgt = upload_vec(tg_tiles);
                auto* d_tlead = upload_vec(tlead);
                timed(P_RSKETCH, 3, (long)lb.size(), [&] {
                    launch_fill<int>(d_tail, (size_t)nt, 0, stream_);
                    launch_eye<S>(d_eye, (int)eye.size(), stream_);
                    if constexpr (sizeof(S) == 4) {
                        for (int q = 0; q < 3; ++q) launch_larfb_res(d_lbr[q], (int)lb_res[q].size(), 4 << q, stream_);
                        launch_larfb_res2(d_lbr[3], (int)lb_res[3].size(), stream_);
                    }
                    launch_larfb<S>(d_lb, (int)lb_rest.size(), stream_);
                    launch_tail_gemm(d_tg, d_tgt, (int)tg_tiles.size(), stream_);
                    launch_tail_select<S>(d_tt, (int)tt.size(), tol, stream_);
                    launch_tail_lead<S>(d_tlead, (int)tlead.size(), tol, stream_);
                });
                std::vector<int> got(nt, 0);
                HIPCHK(hipMemcpyAsync(got.data(), d_tail, sizeof(int) * nt, hipMemcpyDeviceToHost, stream_));
                sync_stream();
                for (int q : todo) {
                    if (!applied[q]) continue;
                    rk[q] = (got[q] == INT_MAX) ? -1 : got[q];
                    const int kw = std::min(gq[q].k, gq[q].w);
                    if (rk[q] < 0 && s_of[q] < kw && round + 1 < rounds) {
                        s_of[q] = std::min(kw, 2 * s_of[q]);
                        again.push_back(q);
                        rstats.retried++;
                    }
                }
            }
            todo.swap(again);
        }
    }
    // ---- decide: sketch result, or geqp3 on C
    std::vector<int> fb;
    std::vector<int> fixed(nt, 0);
    bool any_sketch = false;
    for (int q = 0; q < nt; ++q) {
        if (sketch_mode && is_elig[q] && applied[q] && rk[q] >= 0) {
            used_sketch[q] = 1;
            fixed[q] = std::max(rk[q], 1);  // choose_rank keeps at least one column
            any_sketch = true;
        } else {
            fb.push_back(q);
        }
    }
    if (any_sketch) {  // accepted sketch ranks; the fallback geqp3 below writes the others
        int* src = (int*)upload(fixed.data(), sizeof(int) * nt);
        HIPCHK(hipMemcpyAsync(d_ranks, src, sizeof(int) * nt, hipMemcpyDeviceToDevice, stream_));
    }
    if (!fb.empty()) {
        std::vector<CopyTask<S>> regather;
        std::vector<Geqp3Task<S>> gf;
        for (int q : fb) {
            if (applied[q])
                for (int g = gat_begin[q]; g < gat_begin[q + 1]; ++g) regather.push_back(gat[g]);
            gf.push_back(gq[q]);
            prof_add(P_GEQP3, 0.0, SZ * 3.0 * gq[q].k * gq[q].w);
        }
        if (!regather.empty()) {
            auto* d_rg = upload_vec(regather);
            timed(P_GATHER, 1, (long)regather.size(), [&] { launch_copy<S>(d_rg, (int)regather.size(), stream_, copy_parts(regather)); });
        }
        int nd = 0;
        timed(P_GEQP3, 0, (long)gf.size(),
              [&] { geqp3_grouped<S>(gf, tol, kPivotTieRel, stream_, mem, d_err, &nd); });
        if (profiling) prof[P_GEQP3].launches += nd;
    }
    rstats.tasks += nt;
    for (int q = 0; q < nt; ++q) {
        if (used_sketch[q]) rstats.sketched++;
        else if (sketch_mode && is_elig[q]) {
            rstats.fallback++;
            if ((size_t)gq_cluster[q] < h_info.size()) h_info[gq_cluster[q]] |= kInfoPredictorMiss;
        }
        else rstats.direct++;
    }
    if (rank_log && ne) {
        std::vector<float> X((size_t)ne * kRankFeatures);
        std::vector<int> fin(nt);
        HIPCHK(hipMemcpyAsync(X.data(), d_X, sizeof(float) * X.size(), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(fin.data(), d_ranks, sizeof(int) * nt, hipMemcpyDeviceToHost, stream_));
        sync_stream();
        for (int i = 0; i < ne; ++i) {
            const int q = elig[i];
            rlog.x.insert(rlog.x.end(), X.begin() + (size_t)i * kRankFeatures, X.begin() + (size_t)(i + 1) * kRankFeatures);
            rlog.rank.push_back(fin[q]);
            rlog.k.push_back(gq[q].k);
            rlog.w.push_back(gq[q].w);
            rlog.level.push_back(L);
            rlog.mode.push_back(used_sketch[q] ? rank_mode : 0);
            rlog.order.push_back(gq_cluster[q]);
        }
    }
    return flops;
}

template const float* Tree::omega<float>(int, int);
template const double* Tree::omega<double>(int, int);
template double Tree::rrqr_level<float>(std::vector<Geqp3Task<float>>&, const std::vector<CopyTask<float>>&,
                                        const std::vector<int>&, const std::vector<int>&, int, int*, int*,
                                        std::vector<char>&);
template double Tree::rrqr_level<double>(std::vector<Geqp3Task<double>>&, const std::vector<CopyTask<double>>&,
                                         const std::vector<int>&, const std::vector<int>&, int, int*, int*,
                                         std::vector<char>&);

// ---------------------------------------------------------------------------- predictor host side
namespace {
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 4))); }
    ~DevBuf() { (void)hipFree(p); }
    float* f() const { return (float*)p; }
};
}  // namespace

int mlp_predict_host(const float* X, int n, const float* W, float* out) {
    DevBuf dX(sizeof(float) * (size_t)n * kRankFeatures), dW(sizeof(float) * kMlpParams), dO(sizeof(float) * n);
    HIPCHK(hipMemcpy(dX.p, X, sizeof(float) * (size_t)n * kRankFeatures, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dW.p, W, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    launch_mlp_predict(dX.f(), n, dW.f(), dO.f(), nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, dO.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return 0;
}

int mlp_grad_host(const float* X, const float* y, int n, const float* W, float* grad, float* loss) {
    const int nb = mlp_grad_blocks(n);
    DevBuf dX(sizeof(float) * (size_t)n * kRankFeatures), dy(sizeof(float) * n), dW(sizeof(float) * kMlpParams),
        dP(sizeof(float) * (size_t)nb * (kMlpParams + 1)), dG(sizeof(float) * kMlpParams), dL(sizeof(float));
    HIPCHK(hipMemcpy(dX.p, X, sizeof(float) * (size_t)n * kRankFeatures, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dy.p, y, sizeof(float) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dW.p, W, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    launch_mlp_grad(dX.f(), dy.f(), n, dW.f(), dP.f(), nb, nullptr);
    launch_mlp_reduce(dP.f(), nb, dG.f(), dL.f(), nullptr, nullptr, nullptr, 0, 0.f, 0, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(grad, dG.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(loss, dL.p, sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int mlp_train_host(const float* X, const float* y, int n, float* W, int epochs, float lr, float* loss_hist) {
    const int nb = mlp_grad_blocks(n);
    DevBuf dX(sizeof(float) * (size_t)n * kRankFeatures), dy(sizeof(float) * n), dW(sizeof(float) * kMlpParams),
        dP(sizeof(float) * (size_t)nb * (kMlpParams + 1)), dm(sizeof(float) * kMlpParams),
        dv(sizeof(float) * kMlpParams), dL(sizeof(float) * std::max(epochs, 1));
    HIPCHK(hipMemcpy(dX.p, X, sizeof(float) * (size_t)n * kRankFeatures, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dy.p, y, sizeof(float) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dW.p, W, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(dm.p, 0, sizeof(float) * kMlpParams));
    HIPCHK(hipMemset(dv.p, 0, sizeof(float) * kMlpParams));
    for (int e = 0; e < epochs; ++e) {
        launch_mlp_grad(dX.f(), dy.f(), n, dW.f(), dP.f(), nb, nullptr);
        launch_mlp_reduce(dP.f(), nb, nullptr, dL.f() + e, dW.f(), dm.f(), dv.f(), e + 1, lr, 1, nullptr);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(W, dW.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    if (loss_hist && epochs > 0) HIPCHK(hipMemcpy(loss_hist, dL.p, sizeof(float) * epochs, hipMemcpyDeviceToHost));
    return 0;
}

}  // namespace spaqr
