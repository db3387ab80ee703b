// gfx950 kernels for the spaQR per-front hot path (see kernels.hpp for the task layouts).
//
// Workgroup-level building blocks:
//   wg_panel_qr   Householder QR of a <= 32-column panel (geqrt2 / tpqrt2 with L = 0),
//                 producing the panel's T: the panel in registers (FP32: a column per lane,
//                 up to 1024 rows; FP64: a row per thread, up to 256 rows), taller panels
//                 streamed from memory a row per thread, one barrier per column step.
//   wg_apply_lt   H^T [Atop; C] with a block reflector of <= 32 columns (larfb 'L','T' /
//                 tprfb 'L','T'): W = Atop + V^T C staged through LDS in 64-row chunks,
//                 W <- T^T W, C -= V W with V rows held in registers.
//   wg_apply_rn   C H (larfb 'R','N') for the sparsification's column transform.
//   wg_trsm_ru    C R^{-1} (trsm Right/Upper/NoTrans/NonUnit), 32-column blocked.
// Each batched kernel walks its descriptor (a front, a chain of neighbours, a block) with
// one 256-thread workgroup, so every front of a level is factored in one grouped launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <type_traits>

#include "dpp.hpp"
#include "kernels.hpp"
#include "mfma.hpp"

namespace spaqr {

template <typename S>
struct ApplySmem {
    S Vs[32 * 65];
    S Cs[32 * 65];
    S W[32 * 33];
    S Ts[32 * 33];
};
// FP32 applies run on MFMA: a 64-row x 128-column C tile and per-wave W blocks.
template <>
struct ApplySmem<float> {
    float Vs[32 * 65];
    float Cs[128 * 65];
    float W[32 * 33];
    float Ts[32 * 33];
    float Ws[4 * 32 * 33];
};


// Shared memory of the workgroup panel QR.  `mem`: wg_panel_qr_mem, wg_larft_blocked and the
// row-layout register panel (FP64); `col`: the column-layout register panel (FP32).
template <typename S>
struct MemPanelSmem {
    S Ts[32 * 33];
    S d[32];
    S t[32];
    S tau, scal;
    S red[2][4][32];  // per-step column sums of the row-layout register panel (double-buffered)
    S rowj[2][32];
};
template <typename S>
struct alignas(16) ColPanelSmem {
    static constexpr int kSegMax = 128;  // register rows per thread, at most
    S top[32 * 33];         // the top nbp x nbp block, row-major: top[r * 33 + c]
    S G[32 * 32];           // G(p, j) = v_p^T v_j at j * 32 + p
    S red[2][8][32];        // per-step partial dot products [parity][segment][column]
    S rowj[2][32];          // row j of the top block at step j
    S tau[32], scl[32];     // tau_j and the scale 1 / (alpha_j - beta_j) of reflector j
    S T[32 * 32];           // the panel's T, T(q, j) at j * 32 + q
    S xb[8][kSegMax + 4];   // column j's register rows, per segment (padded: the two
                            // segments of a wave read from different banks)
};
template <typename S>
union PanelSmemCol {
    MemPanelSmem<S> mem;
    ColPanelSmem<S> col;
};
template <typename S>
struct PanelSmemRow {
    MemPanelSmem<S> mem;
};
template <typename S>
using PanelSmem = typename std::conditional<sizeof(S) == 4, PanelSmemCol<S>, PanelSmemRow<S>>::type;
// Kernels that alternate panels and applies.  FP32: one union (every panel and apply starts
// and ends with a barrier, so the phases never overlap); FP64: side by side.
template <typename S>
union ApplyPanelSmemU {
    ApplySmem<S> ap;
    PanelSmem<S> pn;
};
template <typename S>
struct ApplyPanelSmemS {
    ApplySmem<S> ap;
    PanelSmem<S> pn;
};
template <typename S>
using ApplyPanelSmem = typename std::conditional<sizeof(S) == 4, ApplyPanelSmemU<S>, ApplyPanelSmemS<S>>::type;

// sign(alpha) * ||(alpha, x)|| for a Householder reflector, with an alpha at rounding-noise
// level (|alpha| <= band * ||x||) treated as non-negative: its sign carries no information and
// would otherwise make the reflector flip with the summation order.
template <typename S>
__device__ __forceinline__ S householder_norm_signed(S alpha, S ss) {
    const S nrm = sqrt(alpha * alpha + ss);
    const S band = sizeof(S) == 8 ? S(1e-10) : S(1e-5);
    return (fabs(alpha) <= band * sqrt(ss)) ? nrm : copysign(nrm, alpha);
}


// tau, scal = 1 / (alpha - beta) and beta of the reflector of one panel step, from alpha and
// ss = ||x||^2 > 0.  This scalar chain sits on the critical path of every column step, so FP32
// uses the hardware square root and reciprocal (each within 1 ulp) in place of the correctly
// rounded sequences; FP64 keeps the exact operations.  Same sign rule as householder_norm_signed.
__device__ __forceinline__ void householder_fast(float alpha, float ss, float& tau, float& scal, float& beta) {
    const float nrm = __builtin_amdgcn_sqrtf(alpha * alpha + ss);
    const bool noise = alpha * alpha <= 1e-10f * ss;  // |alpha| <= 1e-5 ||x||
    beta = noise ? -nrm : -copysignf(nrm, alpha);
    tau = (beta - alpha) * __builtin_amdgcn_rcpf(beta);
    scal = __builtin_amdgcn_rcpf(alpha - beta);
}
__device__ __forceinline__ void householder_fast(double alpha, double ss, double& tau, double& scal, double& beta) {
    beta = -householder_norm_signed(alpha, ss);
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
}

// ---------------------------------------------------------------------------------------
// Panel QR.  PENT = false: the panel is A (m x nbp, ld ldr == ldb, Rt == B): reflector j
// is [1; A(j+1:m, j)] acting on rows j..m-1.  PENT = true: the panel is [R (nbp x nbp
// upper, ldr); B (m x nbp, ldb)], reflector j is [e_j; B(:, j)] (dtpqrt2 with L = 0).
// The nbp x nbp T of the panel is written to Tout (ld ldt).
template <typename S, bool PENT>
__device__ __forceinline__ void wg_panel_qr_mem(int m, int nbp, S* Rt, int ldr, S* B, int ldb, S* Tout, int ldt,
                                                MemPanelSmem<S>& sm) {
    const int tid = threadIdx.x;
    const int l = tid >> 3, g = tid & 7;
    for (int e = tid; e < 32 * 33; e += WG) sm.Ts[e] = S(0);
    __syncthreads();
    for (int j = 0; j < nbp; ++j) {
        const int r0 = PENT ? 0 : j + 1;
        S d = S(0);
        if (l < nbp) {
            const S* bj = B + (size_t)j * ldb;
            const S* bl = B + (size_t)l * ldb;
#pragma unroll 8
            for (int r = r0 + g; r < m; r += 8) d += bj[r] * bl[r];
        }
        d = oct_sum(d);
        if (g == 0) sm.d[l] = d;
        __syncthreads();
        if (tid == 0) {
            S alpha = Rt[j + (size_t)j * ldr];
            S ss = sm.d[j];
            S tau, scal, beta;
            if (ss == S(0)) {
                tau = S(0); scal = S(1); beta = alpha;
            } else {
                householder_fast(alpha, ss, tau, scal, beta);
            }
            Rt[j + (size_t)j * ldr] = beta;
            sm.tau = tau;
            sm.scal = scal;
        }
        __syncthreads();
        const S tau = sm.tau, scal = sm.scal;
        if (l > j && l < nbp) {
            S wl = (Rt[j + (size_t)l * ldr] + scal * sm.d[l]) * tau;
            const S ws = wl * scal;
            const S* bj = B + (size_t)j * ldb;
            S* bl = B + (size_t)l * ldb;
#pragma unroll 8
            for (int r = r0 + g; r < m; r += 8) bl[r] -= bj[r] * ws;
            if (g == 0) Rt[j + (size_t)l * ldr] -= wl;
        }
        if (tid < j) {
            S tv = scal * sm.d[tid];
            if (!PENT) tv += Rt[j + (size_t)tid * ldr];
            sm.t[tid] = tv;
        }
        __syncthreads();
        if (l == j) {
            S* bj = B + (size_t)j * ldb;
#pragma unroll 8
            for (int r = r0 + g; r < m; r += 8) bj[r] *= scal;
        }
        if (tid < j) {
            S s = S(0);
            for (int q = tid; q < j; ++q) s += sm.Ts[q * 33 + tid] * sm.t[q];
            sm.Ts[j * 33 + tid] = -tau * s;
        } else if (tid == j) {
            sm.Ts[j * 33 + j] = tau;
        }
        __syncthreads();
    }
    for (int e = tid; e < nbp * nbp; e += WG) {
        int i = e % nbp, jj = e / nbp;
        Tout[i + (size_t)jj * ldt] = (i <= jj) ? sm.Ts[jj * 33 + i] : S(0);
    }
    __syncthreads();
}

// Register-resident panel QR: thread t holds the whole rows t, t + 256, ... (NR of them) of
// the m x nbp panel for all nbp steps, so a step is one 32-column dot-product reduction
// (transposing butterfly + LDS, one barrier) and register updates; nothing goes back to
// memory until the end.  Same reflectors, R and T as wg_panel_qr_mem.
template <typename S, bool PENT, int NR>
__device__ __forceinline__ void wg_panel_qr_reg(int m, int nbp, S* Rt, int ldr, S* B, int ldb, S* Tout, int ldt,
                                                MemPanelSmem<S>& sm) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    S a[NR][32];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int R = tid + i * WG, Rc = min(R, m - 1);
#pragma unroll
        for (int l = 0; l < 32; ++l) a[i][l] = B[Rc + (size_t)min(l, nbp - 1) * ldb];
#pragma unroll
        for (int l = 0; l < 32; ++l) a[i][l] = (R < m && l < nbp) ? a[i][l] : S(0);
    }
    for (int e = tid; e < 32 * 33; e += WG) sm.Ts[e] = S(0);
    __syncthreads();
    for (int j = 0; j < nbp; ++j) {
        const int par = j & 1;
        // PENT: row j of the separate triangle, read before the barrier (written after it)
        S rj[32];
        if (PENT) {
#pragma unroll
            for (int l = 0; l < 32; ++l) rj[l] = Rt[j + (size_t)min(l, nbp - 1) * ldr];
        }
        S vj[NR];
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            S v = S(0);
#pragma unroll
            for (int l = 0; l < 32; ++l) v = (l == j) ? a[i][l] : v;
            vj[i] = v;
        }
        S d[32];
#pragma unroll
        for (int l = 0; l < 32; ++l) d[l] = S(0);
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int R = tid + i * WG;
            const S x = (R < m && (PENT || R > j)) ? vj[i] : S(0);
#pragma unroll
            for (int l = 0; l < 32; ++l) d[l] += x * a[i][l];
        }
        if (!PENT && tid == j) {
#pragma unroll
            for (int l = 0; l < 32; ++l) sm.rowj[par][l] = a[0][l];
        }
        const S tot = wave_sum32<S>(d);
        if (vec_owner(lane)) sm.red[par][wave][vec_col(lane)] = tot;
        __syncthreads();
        S dl[32];
#pragma unroll
        for (int l = 0; l < 32; ++l)
            dl[l] = (sm.red[par][0][l] + sm.red[par][1][l]) + (sm.red[par][2][l] + sm.red[par][3][l]);
        if (!PENT) {
#pragma unroll
            for (int l = 0; l < 32; ++l) rj[l] = sm.rowj[par][l];
        }
        S alpha = S(0), ss = S(0);
#pragma unroll
        for (int l = 0; l < 32; ++l) {
            alpha = (l == j) ? rj[l] : alpha;
            ss = (l == j) ? dl[l] : ss;
        }
        S tau, scal, beta;
        if (ss == S(0)) {
            tau = S(0); scal = S(1); beta = alpha;
        } else {
            householder_fast(alpha, ss, tau, scal, beta);
        }
        S wl[32];
#pragma unroll
        for (int l = 0; l < 32; ++l) wl[l] = (l > j && l < nbp) ? (rj[l] + scal * dl[l]) * tau : S(0);
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int R = tid + i * WG;
            if (R < m && (PENT || R > j)) {
                const S vs = vj[i] * scal;
#pragma unroll
                for (int l = 0; l < 32; ++l) a[i][l] = (l == j) ? vs : a[i][l] - vs * wl[l];
            } else if (!PENT && R == j) {
#pragma unroll
                for (int l = 0; l < 32; ++l) a[i][l] = (l == j) ? beta : a[i][l] - wl[l];
            }
        }
        if (PENT) {
            S mine = S(0);
#pragma unroll
            for (int l = 0; l < 32; ++l) mine = (l == tid) ? rj[l] - wl[l] : mine;
            if (tid == j) Rt[j + (size_t)j * ldr] = beta;
            else if (tid > j && tid < nbp) Rt[j + (size_t)tid * ldr] = mine;
        }
        // column j of T: T(q, j) = -tau sum_{p=q}^{j-1} T(q, p) t_p, t_p = v_p^T v_j
        if (tid < j) {
            S s = S(0);
#pragma unroll
            for (int q = 0; q < 32; ++q) {
                if (q >= tid && q < j) {
                    const S tq = scal * dl[q] + (PENT ? S(0) : rj[q]);
                    s += sm.Ts[q * 33 + tid] * tq;
                }
            }
            sm.Ts[j * 33 + tid] = -tau * s;
        } else if (tid == j) {
            sm.Ts[j * 33 + j] = tau;
        }
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int R = tid + i * WG;
        if (R < m) {
#pragma unroll
            for (int l = 0; l < 32; ++l)
                if (l < nbp) B[R + (size_t)l * ldb] = a[i][l];
        }
    }
    __syncthreads();
    for (int e = tid; e < nbp * nbp; e += WG) {
        int i = e % nbp, jj = e / nbp;
        Tout[i + (size_t)jj * ldt] = (i <= jj) ? sm.Ts[jj * 33 + i] : S(0);
    }
    __syncthreads();
}

// f(i, xs[i]) for i = 0 .. SR-1 in 16-row chunks, the next chunk's LDS loads issued before
// the current chunk's work: one chunk of lookahead.  Without the scheduling fences the
// compiler issues all SR loads first, and long segments then spill.
template <typename S, int SR, typename F>
__device__ __forceinline__ void col_chunks(const S* xs, F&& f) {
    constexpr int CH = SR < 16 ? SR : 16, NCH = SR / CH;
    S x[2][CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) x[0][u] = xs[u];
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
        if (k + 1 < NCH) {
#pragma unroll
            for (int u = 0; u < CH; ++u) x[(k + 1) & 1][u] = xs[(k + 1) * CH + u];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < CH; ++u) f(k * CH + u, x[k & 1][u]);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Register-resident panel QR with the columns across the lanes.  Thread (g, c) -- column
// c = lane % 32, segment g = 2 * wave + lane / 32 -- holds SR consecutive rows of column c,
// rows base + g SR ... (base = nbp for the square panel, whose top nbp x nbp block sits in
// LDS; base = 0 for the pentagonal one, whose top block is the separate triangle R, also
// staged in LDS).  Top row r is handled by segment r % 8.  Step j:
//   - column j's lanes publish their register rows to their segment (LDS, same wave);
//   - every thread forms its column's partial dot product with column j (fixed register
//     indices, plus its four top rows), one barrier, eight partials summed per column;
//   - each thread derives the reflector (all threads alike) and updates its rows with one FMA.
// Reflector columns stay unscaled until the panel's end (scl[j] keeps 1 / (alpha - beta)),
// so column j needs no case of its own.  The same dot products give the Gram column
// G(p, j) = v_p^T v_j of the finished columns, from which wave 0 builds T at the end.
// Same reflectors (sign rule included), R, V and T as wg_panel_qr_mem.
template <typename S, bool PENT, int SR, bool VEC = false>
__device__ __forceinline__ void wg_panel_qr_col(int m, int nbp, S* Rt, int ldr, S* B, int ldb, S* Tout, int ldt,
                                                ColPanelSmem<S>& sm) {
    static_assert(SR % 4 == 0 && SR <= ColPanelSmem<S>::kSegMax, "segment length");
    // short segments keep column j's rows in registers between the dot product and the update
    constexpr bool kKeep = SR <= 32;
    const int tid = threadIdx.x, lane = tid & 63, c = lane & 31;
    const int g = 2 * (tid >> 6) + (lane >> 5);
    int r0 = (PENT ? 0 : nbp) + g * SR;  // this thread's first register row
    // opaque per call: otherwise the row indices r0 + i of every segment length are hoisted out
    // of the callers' link and panel loops and stay live (spilled) for the whole kernel
    asm volatile("" : "+v"(r0));
    const int cc = min(c, nbp - 1);
    __syncthreads();  // the LDS may be shared with the phase before
    for (int e = tid; e < 32 * 32; e += WG) {
        const int r = e & 31, q = e >> 5;
        const S v = Rt[min(r, nbp - 1) + (size_t)min(q, nbp - 1) * ldr];
        sm.top[r * 33 + q] = (r < nbp && q < nbp && (!PENT || r <= q)) ? v : S(0);
        sm.G[e] = S(0);
    }
    // A thread's SR rows of column c are consecutive in memory.  VEC: a segment that lies inside
    // the panel is read four rows per 16-byte load (dword-aligned vector loads) and the one that
    // ends past row m row by row, clamped; the second path costs registers, so only kernels with
    // room for it (the look-ahead panel step) take it.
    S a[SR];
    const S* colc = B + (size_t)cc * ldb;
    if (VEC && r0 + SR <= m) {
#pragma unroll
        for (int i = 0; i < SR; i += 4) __builtin_memcpy(&a[i], colc + r0 + i, 4 * sizeof(S));
    } else {
#pragma unroll
        for (int i = 0; i < SR; ++i) a[i] = colc[min(r0 + i, m - 1)];
    }
#pragma unroll
    for (int i = 0; i < SR; ++i) a[i] = (r0 + i < m && c < nbp) ? a[i] : S(0);
    __syncthreads();
    // square panel: row j of the top block as it stands at step j goes to rowj[j & 1], written by
    // the segment that updates that row (at step j - 1; row 0 here) and read after step j's
    // barrier, so that the row's own update at step j never races with its readers
    if (!PENT && g == 0) sm.rowj[0][c] = sm.top[c];
    S* xs = sm.xb[g];
    for (int j = 0; j < nbp; ++j) {
        const int par = j & 1;
        wave_sync();  // this wave's reads of the previous step's column are done
        if (c == j) {
#pragma unroll
            for (int i = 0; i < SR; ++i) xs[i] = a[i];
        }
        wave_sync();
        // pentagonal panel: row j of R changes only at step j (reflector j is [e_j; B(:, j)]), so
        // it is read before the barrier, ahead of its own update
        S alphaP = S(0), rowcP = S(0);
        if (PENT) {
            alphaP = sm.top[j * 33 + j];
            rowcP = sm.top[j * 33 + c];
        }
        S xr[kKeep ? SR : 1];
        S d = S(0);
        if constexpr (kKeep) {
#pragma unroll
            for (int i = 0; i < SR; ++i) xr[i] = xs[i];
            S d4[4] = {S(0), S(0), S(0), S(0)};
#pragma unroll
            for (int i = 0; i < SR; ++i) d4[i & 3] += xr[i] * a[i];
            d = (d4[0] + d4[1]) + (d4[2] + d4[3]);
        } else {
            S d0 = S(0), d1 = S(0);
            col_chunks<S, SR>(xs, [&](int i, S x) {
                if (i & 1) d1 += x * a[i]; else d0 += x * a[i];
            });
            d = d0 + d1;
        }
        // square panel: the segment's four top rows (this segment's own writes, in program order),
        // read once for the dot product and the update
        S tr[4], tj[4];
        if (!PENT) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = g + 8 * q;
                tr[q] = sm.top[r * 33 + c];
                tj[q] = sm.top[r * 33 + j];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) d += (g + 8 * q > j) ? tj[q] * tr[q] : S(0);
        }
        sm.red[par][g][c] = d;
        __syncthreads();
        S ss, dc;
        {
            S p[8], q[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) { p[u] = sm.red[par][u][j]; q[u] = sm.red[par][u][c]; }
            ss = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
            dc = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
        }
        const S alpha = PENT ? alphaP : sm.rowj[par][j], rowc = PENT ? rowcP : sm.rowj[par][c];
        S tau, scal, beta;
        if (ss == S(0)) {
            tau = S(0); scal = S(1); beta = alpha;
        } else {
            householder_fast(alpha, ss, tau, scal, beta);
        }
        // w_c = tau v_j^T a_c for the trailing columns; v_j = [e_j; scal x_j]
        const S vta = (PENT ? S(0) : rowc) + scal * dc;
        const S w = (c > j && c < nbp) ? (rowc + scal * dc) * tau : S(0);
        const S sw = scal * w;
        if (g == 0 && c < j) sm.G[j * 32 + c] = sm.scl[c] * vta;  // column c is still unscaled
        if (tid == 0) { sm.tau[j] = tau; sm.scl[j] = scal; }
        if constexpr (kKeep) {
#pragma unroll
            for (int i = 0; i < SR; ++i) a[i] -= xr[i] * sw;
        } else {
            col_chunks<S, SR>(xs, [&](int i, S x) { a[i] -= x * sw; });
        }
        if (!PENT) {
            // every write after every read of the step; unchanged entries written back as read.
            // Lanes of one segment are one half-wave, so the column-j reads precede lane j's
            // write of its diagonal.
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = g + 8 * q;
                const S upd = (r > j) ? tr[q] - tj[q] * sw : tr[q] - w;
                const S nv = (r < j) ? tr[q] : ((r == j && c == j) ? beta : upd);
                sm.top[r * 33 + c] = nv;
                if (r == j + 1) sm.rowj[par ^ 1][c] = nv;
            }
        } else if ((j & 7) == g) {
            sm.top[j * 33 + c] = (c == j) ? beta : rowcP - w;
        }
    }
    __syncthreads();
    // T, row q per lane of wave 0: T(q, j) = -tau_j sum_{p=q}^{j-1} T(q, p) G(p, j), T(j, j) = tau_j.
    // Each T(q, j) goes to LDS as it is formed (an unconditional store keeps the arithmetic in
    // place; otherwise it is sunk into the final stores and all 496 G loads stay live).
    if (tid < 32) {
        const int q = tid;
        S Tq[32];
#pragma unroll
        for (int jj = 0; jj < 32; ++jj) {
            S s = S(0);
#pragma unroll
            for (int p = 0; p < 32; ++p)  // fixed trip count: both loops unroll fully, Tq stays in registers
                if (p < jj) s += Tq[p] * sm.G[jj * 32 + p];
            const S tj = sm.tau[jj];
            Tq[jj] = (q < jj) ? -tj * s : (q == jj ? tj : S(0));
            sm.T[jj * 32 + q] = Tq[jj];
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    __syncthreads();
    for (int e = tid; e < 32 * 32; e += WG) {
        const int q = e & 31, jj = e >> 5;
        if (q < nbp && jj < nbp) Tout[q + (size_t)jj * ldt] = sm.T[e];
    }
    // V below the diagonal (scaled now), R on and above it
    if (c < nbp) {
        const S sc = sm.scl[c];
        S* colw = B + (size_t)c * ldb;
#pragma unroll
        for (int i = 0; i < SR; ++i) a[i] *= sc;
        if (VEC && r0 + SR <= m) {
#pragma unroll
            for (int i = 0; i < SR; i += 4) __builtin_memcpy(colw + r0 + i, &a[i], 4 * sizeof(S));
        } else {
#pragma unroll
            for (int i = 0; i < SR; ++i)
                if (r0 + i < m) colw[r0 + i] = a[i];
        }
    }
    for (int e = tid; e < 32 * 32; e += WG) {
        const int r = e & 31, q = e >> 5;
        if (r < nbp && q < nbp) {
            const S v = sm.top[r * 33 + q];
            if (r <= q) Rt[r + (size_t)q * ldr] = v;
            else if (!PENT) Rt[r + (size_t)q * ldr] = v * sm.scl[q];
        }
    }
    __syncthreads();
}

// Row-layout panel for panels taller than the register panels hold: the steps of
// wg_panel_qr_reg (thread t owns rows t, t + WG, ...; one 32-column reduction and one barrier per
// step) with the rows read from memory in every step instead of held in registers.  Step j
// reads the thread's rows (RB at a time, every column, so that a wave's load of one column is
// 64 consecutive rows: whole lines) for the dot products with column j, and again to apply the
// reflector, writing back columns j.. only.  Same reflectors (sign rule included), R, V and T as
// wg_panel_qr_mem.
template <typename S, bool PENT>
__device__ __forceinline__ void wg_panel_qr_rows(int m, int nbp, S* Rt, int ldr, S* B, int ldb, S* Tout, int ldt,
                                                 MemPanelSmem<S>& sm) {
    constexpr int RB = sizeof(S) == 4 ? 4 : 2;  // rows per block, all of their columns in flight
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nrow = (m + WG - 1) / WG;  // rows per thread
    const S* Bg = B;  // read through the same (non-restrict) pointer the stores use
    for (int e = tid; e < 32 * 33; e += WG) sm.Ts[e] = S(0);
    __syncthreads();
    for (int j = 0; j < nbp; ++j) {
        const int par = j & 1;
        // PENT: row j of the separate triangle, read before the barrier (written after it)
        S rj[32];
        if (PENT) {
#pragma unroll
            for (int l = 0; l < 32; ++l) rj[l] = Rt[j + (size_t)min(l, nbp - 1) * ldr];
        }
        S d[32];
#pragma unroll
        for (int l = 0; l < 32; ++l) d[l] = S(0);
        for (int i0 = 0; i0 < nrow; i0 += RB) {
            S a[RB][32], x[RB];
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int R = tid + (i0 + u) * WG, Rc = min(R, m - 1);
#pragma unroll
                for (int l = 0; l < 32; ++l) a[u][l] = Bg[Rc + (size_t)min(l, nbp - 1) * ldb];
                x[u] = Bg[Rc + (size_t)j * ldb];
            }
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int R = tid + (i0 + u) * WG;
                const bool in = R < m && (PENT || R > j);
                const S xu = in ? x[u] : S(0);
#pragma unroll
                for (int l = 0; l < 32; ++l) d[l] += xu * ((l < nbp) ? a[u][l] : S(0));
                // square panel: row j is this thread's first row when tid == j
                if (!PENT && R == j) {
#pragma unroll
                    for (int l = 0; l < 32; ++l) sm.rowj[par][l] = (l < nbp) ? a[u][l] : S(0);
                }
            }
        }
        const S tot = wave_sum32<S>(d);
        if (vec_owner(lane)) sm.red[par][wave][vec_col(lane)] = tot;
        __syncthreads();
        S dl[32];
#pragma unroll
        for (int l = 0; l < 32; ++l)
            dl[l] = (sm.red[par][0][l] + sm.red[par][1][l]) + (sm.red[par][2][l] + sm.red[par][3][l]);
        if (!PENT) {
#pragma unroll
            for (int l = 0; l < 32; ++l) rj[l] = sm.rowj[par][l];
        }
        S alpha = S(0), ss = S(0);
#pragma unroll
        for (int l = 0; l < 32; ++l) {
            alpha = (l == j) ? rj[l] : alpha;
            ss = (l == j) ? dl[l] : ss;
        }
        S tau, scal, beta;
        if (ss == S(0)) {
            tau = S(0); scal = S(1); beta = alpha;
        } else {
            householder_fast(alpha, ss, tau, scal, beta);
        }
        S wl[32];
#pragma unroll
        for (int l = 0; l < 32; ++l) wl[l] = (l > j && l < nbp) ? (rj[l] + scal * dl[l]) * tau : S(0);
        // apply: rows below j (all rows of B when PENT) get v = scal * x in column j and
        // a - v w in the trailing columns; row j of the square panel gets beta and a - w
        for (int i0 = 0; i0 < nrow; i0 += RB) {
            S a[RB][32];
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int R = tid + (i0 + u) * WG, Rc = min(R, m - 1);
#pragma unroll
                for (int l = 0; l < 32; ++l) a[u][l] = Bg[Rc + (size_t)min(l, nbp - 1) * ldb];
            }
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int R = tid + (i0 + u) * WG;
                if (R >= m || (!PENT && R < j)) continue;
                S xj = S(0);
#pragma unroll
                for (int l = 0; l < 32; ++l) xj = (l == j) ? a[u][l] : xj;
                const bool vrow = PENT || R > j;
                const S vs = xj * scal;
#pragma unroll
                for (int l = 0; l < 32; ++l) {
                    if (l < j || l >= nbp) continue;
                    const S nv = vrow ? ((l == j) ? vs : a[u][l] - vs * wl[l]) : ((l == j) ? beta : a[u][l] - wl[l]);
                    B[R + (size_t)l * ldb] = nv;
                }
            }
        }
        if (PENT) {
            S mine = S(0);
#pragma unroll
            for (int l = 0; l < 32; ++l) mine = (l == tid) ? rj[l] - wl[l] : mine;
            if (tid == j) Rt[j + (size_t)j * ldr] = beta;
            else if (tid > j && tid < nbp) Rt[j + (size_t)tid * ldr] = mine;
        }
        // column j of T: T(q, j) = -tau sum_{p=q}^{j-1} T(q, p) t_p, t_p = v_p^T v_j
        if (tid < j) {
            S sacc = S(0);
#pragma unroll
            for (int q = 0; q < 32; ++q) {
                if (q >= tid && q < j) {
                    const S tq = scal * dl[q] + (PENT ? S(0) : rj[q]);
                    sacc += sm.Ts[q * 33 + tid] * tq;
                }
            }
            sm.Ts[j * 33 + tid] = -tau * sacc;
        } else if (tid == j) {
            sm.Ts[j * 33 + j] = tau;
        }
    }
    __syncthreads();
    for (int e = tid; e < nbp * nbp; e += WG) {
        int i = e % nbp, jj = e / nbp;
        Tout[i + (size_t)jj * ldt] = (i <= jj) ? sm.Ts[jj * 33 + i] : S(0);
    }
    __syncthreads();
}

template <typename S, bool PENT, bool VEC = false>
__device__ __forceinline__ void wg_panel_qr(int m, int nbp, S* Rt, int ldr, S* B, int ldb, S* Tout, int ldt,
                                            PanelSmem<S>& sm) {
    if constexpr (sizeof(S) == 4) {
        const int rows = PENT ? m : m - nbp;  // rows below the top block
        if (m <= 0) wg_panel_qr_mem<S, PENT>(m, nbp, Rt, ldr, B, ldb, Tout, ldt, sm.mem);  // reads no row of B
        else if (rows <= 64) wg_panel_qr_col<S, PENT, 8, VEC>(m, nbp, Rt, ldr, B, ldb, Tout, ldt, sm.col);
        else if (rows <= 128) wg_panel_qr_col<S, PENT, 16, VEC>(m, nbp, Rt, ldr, B, ldb, Tout, ldt, sm.col);
        else if (rows <= 256) wg_panel_qr_col<S, PENT, 32, VEC>(m, nbp, Rt, ldr, B, ldb, Tout, ldt, sm.col);
        else if (rows <= 512) wg_panel_qr_col<S, PENT, 64, VEC>(m, nbp, Rt, ldr, B, ldb, Tout, ldt, sm.col);
        else if (rows <= 1024) wg_panel_qr_col<S, PENT, 128, VEC>(m, nbp, Rt, ldr, B, ldb, Tout, ldt, sm.col);
        else wg_panel_qr_rows<S, PENT>(m, nbp, Rt, ldr, B, ldb, Tout, ldt, sm.mem);
    } else {
        // FP64 keeps the row-layout register panel (one row per thread, m <= 256); taller panels
        // stream the same row layout from memory
        if (m <= WG) wg_panel_qr_reg<S, PENT, 1>(m, nbp, Rt, ldr, B, ldb, Tout, ldt, sm.mem);
        else wg_panel_qr_rows<S, PENT>(m, nbp, Rt, ldr, B, ldb, Tout, ldt, sm.mem);
    }
}

// ---------------------------------------------------------------------------------------
// [Atop; C] <- H^T [Atop; C] (PENT, V is the m x kb lower block of [I; V]) or
// C <- H^T C (unit lower V, m x kb, implicit ones / zeros above).  kb <= 32.
// H^T [Atop; C] (or H^T C) on MFMA, FP32: for each 128-column chunk of C (wave w owns its
// 32-column slice), W = Atop + V^T C accumulates over 64-row tiles of V and C staged in LDS,
// W <- T^T W, Atop -= W, then C -= V W tile by tile.  Every MFMA is a 32x32x2 step.
// TRC = true applies to the transpose of the array given: X = C^T with X(r, j) at C[j + r ldc]
// (C X-columns are C rows), which turns the right apply C <- C H into X <- H^T X.
template <bool PENT, bool TRC = false>
__device__ __forceinline__ void wg_apply_lt_mfma(int m, int kb, const float* __restrict__ V, int ldv,
                                                 const float* __restrict__ T, int ldt, float* Atop, int lda, float* C,
                                                 int ldc, int ncols, ApplySmem<float>& sm) {
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
    const int li = l & 31, lk = l >> 5;
    float* Ws = sm.Ws + w * (32 * 33);
    __syncthreads();
    for (int e = tid; e < 32 * 32; e += WG) {  // Ts[i * 33 + k] = T(k, i), upper
        const int k = e & 31, i = e >> 5;
        const float tv = T[min(k, kb - 1) + (size_t)min(i, kb - 1) * ldt];
        sm.Ts[i * 33 + k] = (k <= i && i < kb) ? tv : 0.f;
    }
    auto stage = [&](int r0, int rb, int c0, int nc) {
        float vv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * WG, r = e & 63, i = e >> 6;
            vv[q] = V[(r0 + min(r, rb - 1)) + (size_t)min(i, kb - 1) * ldv];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * WG, r = e & 63, i = e >> 6, R = r0 + r;
            float v = 0.f;
            if (r < rb && i < kb) v = PENT ? vv[q] : ((R > i) ? vv[q] : (R == i ? 1.f : 0.f));
            sm.Vs[i * 65 + r] = v;
        }
#pragma unroll 4
        for (int q0 = 0; q0 < 32; q0 += 8) {
            float cv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = tid + (q0 + q) * WG;
                const int r = TRC ? (e >> 7) : (e & 63), j = TRC ? (e & 127) : (e >> 6);  // coalesced either way
                const int rr = r0 + min(r, rb - 1), jj = c0 + min(j, nc - 1);
                cv[q] = TRC ? C[jj + (size_t)rr * ldc] : C[rr + (size_t)jj * ldc];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = tid + (q0 + q) * WG;
                const int r = TRC ? (e >> 7) : (e & 63), j = TRC ? (e & 127) : (e >> 6);
                sm.Cs[j * 65 + r] = (r < rb && j < nc) ? cv[q] : 0.f;
            }
        }
    };
    if (!TRC && ncols >= 1 && ncols <= 32) {
        // At most 32 columns (a panel's own columns): the waves split each tile's rows for
        // W = Atop + V^T C (wave w: rows 16w .. 16w + 15; partial W summed through LDS), every
        // wave forms T^T W, and waves 0 and 1 update the tile's two 32-row halves.
        const int nc = ncols;
        struct NarrowRegs {
            float v[8];
            float c[8];
        };
        auto load_n = [&](int r0, int rb, NarrowRegs& g) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = tid + q * WG, r = e & 63, i = e >> 6;
                g.v[q] = V[(r0 + min(r, rb - 1)) + (size_t)min(i, kb - 1) * ldv];
                g.c[q] = C[(r0 + min(r, rb - 1)) + (size_t)min(i, nc - 1) * ldc];
            }
        };
        auto store_n = [&](int r0, int rb, const NarrowRegs& g) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = tid + q * WG, r = e & 63, i = e >> 6, R = r0 + r;
                float v = 0.f;
                if (r < rb && i < kb) v = PENT ? g.v[q] : ((R > i) ? g.v[q] : (R == i ? 1.f : 0.f));
                sm.Vs[i * 65 + r] = v;
                sm.Cs[i * 65 + r] = (r < rb && i < nc) ? g.c[q] : 0.f;
            }
        };
        floatx16 acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            float a = 0.f;
            if (PENT && w == 0) {
                const int i = mfma_row(l, r);
                const float av = Atop[min(i, kb - 1) + (size_t)min(li, nc - 1) * lda];
                a = (i < kb && li < nc) ? av : 0.f;
            }
            acc[r] = a;
        }
        NarrowRegs g;
        load_n(0, min(64, m), g);
        for (int r0 = 0; r0 < m; r0 += 64) {
            const int rb = min(64, m - r0);
            __syncthreads();
            store_n(r0, rb, g);
            __syncthreads();
            if (r0 + 64 < m) load_n(r0 + 64, min(64, m - r0 - 64), g);
#pragma unroll
            for (int kk = 16 * w; kk < 16 * w + 16; kk += 2) {
                const float a = sm.Vs[li * 65 + kk + lk];
                const float b = sm.Cs[li * 65 + kk + lk];
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
            }
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) Ws[li * 33 + mfma_row(l, r)] = acc[r];
        load_n(0, min(64, m), g);
        __syncthreads();
        floatx16 w2;
#pragma unroll
        for (int r = 0; r < 16; ++r) w2[r] = 0.f;
#pragma unroll
        for (int kk = 0; kk < 32; kk += 2) {
            const float a = sm.Ts[li * 33 + kk + lk];
            const int o = li * 33 + kk + lk;
            const float b = (sm.Ws[o] + sm.Ws[32 * 33 + o]) + (sm.Ws[2 * 32 * 33 + o] + sm.Ws[3 * 32 * 33 + o]);
            w2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, w2, 0, 0, 0);
        }
        __syncthreads();  // every wave has read the partial blocks
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = mfma_row(l, r);
            const float v = (i < kb) ? w2[r] : 0.f;
            Ws[li * 33 + i] = v;
            if (PENT && w == 0 && i < kb && li < nc) Atop[i + (size_t)li * lda] -= v;
        }
        for (int r0 = 0; r0 < m; r0 += 64) {
            const int rb = min(64, m - r0);
            __syncthreads();
            store_n(r0, rb, g);
            __syncthreads();
            if (r0 + 64 < m) load_n(r0 + 64, min(64, m - r0 - 64), g);
            if (w < 2) {
                floatx16 d;
#pragma unroll
                for (int r = 0; r < 16; ++r) d[r] = 0.f;
#pragma unroll
                for (int kk = 0; kk < 32; kk += 2) {
                    const float a = sm.Vs[(kk + lk) * 65 + w * 32 + li];
                    const float b = Ws[li * 33 + kk + lk];
                    d = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, d, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) sm.Cs[li * 65 + w * 32 + mfma_row(l, r)] -= d[r];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = tid + q * WG, r = e & 63, j = e >> 6;
                if (r < rb && j < nc) C[(r0 + r) + (size_t)j * ldc] = sm.Cs[j * 65 + r];
            }
        }
        __syncthreads();
        return;
    }
    for (int c0 = 0; c0 < ncols; c0 += 128) {
        const int nc = min(128, ncols - c0);
        const int jc = 32 * w + li;  // this lane's column inside the chunk
        floatx16 acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            float a = 0.f;
            if (PENT) {
                const int i = mfma_row(l, r);
                const float av = Atop[min(i, kb - 1) + (size_t)(c0 + min(jc, nc - 1)) * lda];
                a = (i < kb && jc < nc) ? av : 0.f;
            }
            acc[r] = a;
        }
        for (int r0 = 0; r0 < m; r0 += 64) {
            const int rb = min(64, m - r0);
            __syncthreads();
            stage(r0, rb, c0, nc);
            __syncthreads();
#pragma unroll 8
            for (int kk = 0; kk < 64; kk += 2) {
                const float a = sm.Vs[li * 65 + kk + lk];
                const float b = sm.Cs[jc * 65 + kk + lk];
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
            }
        }
        // W <- T^T W (W through the wave's LDS block, column-major by j)
#pragma unroll
        for (int r = 0; r < 16; ++r) Ws[li * 33 + mfma_row(l, r)] = acc[r];
        __syncthreads();
        floatx16 w2;
#pragma unroll
        for (int r = 0; r < 16; ++r) w2[r] = 0.f;
#pragma unroll
        for (int kk = 0; kk < 32; kk += 2) {
            const float a = sm.Ts[li * 33 + kk + lk];
            const float b = Ws[li * 33 + kk + lk];
            w2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, w2, 0, 0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = mfma_row(l, r);
            const float v = (i < kb) ? w2[r] : 0.f;
            Ws[li * 33 + i] = v;
            if (PENT && i < kb && jc < nc) Atop[i + (size_t)(c0 + jc) * lda] -= v;
        }
        // C -= V W, tile by tile
        for (int r0 = 0; r0 < m; r0 += 64) {
            const int rb = min(64, m - r0);
            __syncthreads();
            stage(r0, rb, c0, nc);
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                floatx16 d;
#pragma unroll
                for (int r = 0; r < 16; ++r) d[r] = 0.f;
#pragma unroll
                for (int kk = 0; kk < 32; kk += 2) {
                    const float a = sm.Vs[(kk + lk) * 65 + h * 32 + li];
                    const float b = Ws[li * 33 + kk + lk];
                    d = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, d, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) sm.Cs[jc * 65 + h * 32 + mfma_row(l, r)] -= d[r];
            }
            __syncthreads();
#pragma unroll 4
            for (int q = 0; q < 32; ++q) {
                const int e = tid + q * WG;
                const int r = TRC ? (e >> 7) : (e & 63), j = TRC ? (e & 127) : (e >> 6);
                if (r < rb && j < nc) {
                    if (TRC) C[(c0 + j) + (size_t)(r0 + r) * ldc] = sm.Cs[j * 65 + r];
                    else C[(r0 + r) + (size_t)(c0 + j) * ldc] = sm.Cs[j * 65 + r];
                }
            }
        }
    }
    __syncthreads();
}

template <typename S, bool PENT>
__device__ __forceinline__ void wg_apply_lt(int m, int kb, const S* __restrict__ V, int ldv, const S* __restrict__ T, int ldt,
                            S* Atop, int lda, S* C, int ldc, int ncols, ApplySmem<S>& sm) {
    // Loads are unconditional (indices clamped into the operands, values masked after), so a
    // thread's loads issue back to back instead of one memory latency each.
    const int tid = threadIdx.x;
    if (m <= 0 || kb <= 0) return;
    if constexpr (sizeof(S) == 4) {
        wg_apply_lt_mfma<PENT>(m, kb, V, ldv, T, ldt, Atop, lda, C, ldc, ncols, sm);
        return;
    }
    __syncthreads();
    for (int e = tid; e < 32 * 32; e += WG) {
        int i = e & 31, j = e >> 5;
        sm.Ts[j * 33 + i] = (i <= j && j < kb) ? T[i + (size_t)min(j, kb - 1) * ldt] : S(0);
    }
    const int wi = tid & 31, wq = tid >> 5;
    const int rr = tid & 63, cg = tid >> 6;
    for (int c0 = 0; c0 < ncols; c0 += 32) {
        const int nc = min(32, ncols - c0);
        S acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int j = wq * 4 + q;
            acc[q] = S(0);
            if (PENT) {
                const S a = Atop[min(wi, kb - 1) + (size_t)(c0 + min(j, nc - 1)) * lda];
                acc[q] = (wi < kb && j < nc) ? a : S(0);
            }
        }
        for (int r0 = 0; r0 < m; r0 += 64) {
            const int rb = min(64, m - r0);
            __syncthreads();
            S vv[8], cv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = tid + q * WG, r = e & 63, i = e >> 6;
                const int R = r0 + min(r, rb - 1);
                vv[q] = V[R + (size_t)min(i, kb - 1) * ldv];
                cv[q] = C[R + (size_t)(c0 + min(i, nc - 1)) * ldc];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = tid + q * WG, r = e & 63, i = e >> 6;
                const int R = r0 + r;
                S v = S(0);
                if (r < rb && i < kb) v = PENT ? vv[q] : ((R > i) ? vv[q] : (R == i ? S(1) : S(0)));
                sm.Vs[i * 65 + r] = v;
                sm.Cs[i * 65 + r] = (r < rb && i < nc) ? cv[q] : S(0);
            }
            __syncthreads();
            for (int r = 0; r < rb; ++r) {
                S v = sm.Vs[wi * 65 + r];
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] += v * sm.Cs[(wq * 4 + q) * 65 + r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) sm.W[(wq * 4 + q) * 33 + wi] = acc[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            S s = S(0);
            for (int l2 = 0; l2 <= wi; ++l2) s += sm.Ts[wi * 33 + l2] * sm.W[(wq * 4 + q) * 33 + l2];
            acc[q] = s;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) sm.W[(wq * 4 + q) * 33 + wi] = (wi < kb) ? acc[q] : S(0);
        __syncthreads();
        if (PENT) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int j = wq * 4 + q;
                if (wi < kb && j < nc) Atop[wi + (size_t)(c0 + j) * lda] -= sm.W[j * 33 + wi];
            }
        }
        for (int r0 = 0; r0 < m; r0 += 64) {
            const int R = r0 + rr;
            if (R < m) {
                S vrow[32];
#pragma unroll
                for (int i = 0; i < 32; ++i) vrow[i] = V[R + (size_t)min(i, kb - 1) * ldv];
#pragma unroll
                for (int i = 0; i < 32; ++i)
                    vrow[i] = (i < kb) ? (PENT ? vrow[i] : ((R > i) ? vrow[i] : (R == i ? S(1) : S(0)))) : S(0);
                S cold[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) cold[q] = C[R + (size_t)(c0 + min(cg * 8 + q, nc - 1)) * ldc];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    int j = cg * 8 + q;
                    S s = S(0);
#pragma unroll
                    for (int i = 0; i < 32; ++i) s += vrow[i] * sm.W[j * 33 + i];
                    if (j < nc) C[R + (size_t)(c0 + j) * ldc] = cold[q] - s;
                }
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------------
// C (m x n) <- C H, H = I - V T V^T, V (n x kb unit lower, ldv), kb <= 32.
template <typename S>
__device__ __forceinline__ void wg_apply_rn(int m, int n, int kb, const S* __restrict__ V, int ldv, const S* __restrict__ T, int ldt,
                            S* C, int ldc, ApplySmem<S>& sm) {
    // Two sweeps over 64-column chunks of V staged in LDS (unit lower-trapezoidal, n x kb):
    // W = C V for this 64-row tile, W <- W T, then C -= W V^T.  All global loads are
    // unconditional (clamped, masked), issued in groups of 16 per thread.
    const int tid = threadIdx.x;
    if (m <= 0 || n <= 0 || kb <= 0) return;
    if constexpr (sizeof(S) == 4) {  // FP32: C H = (H^T C^T)^T on MFMA
        wg_apply_lt_mfma<false, true>(n, kb, V, ldv, T, ldt, nullptr, 0, C, ldc, m, sm);
        return;
    }
    __syncthreads();
    for (int e = tid; e < 32 * 32; e += WG) {
        int i = e & 31, j = e >> 5;
        sm.Ts[j * 33 + i] = (i <= j && j < kb) ? T[i + (size_t)min(j, kb - 1) * ldt] : S(0);
    }
    const int rr = tid & 63, cg = tid >> 6;  // row in a 64-row tile, group of 8 reflectors
    auto stage_v = [&](int l0, int lb) {
        S vv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * WG, l = e & 63, i = e >> 6;
            vv[q] = V[(l0 + min(l, lb - 1)) + (size_t)min(i, kb - 1) * ldv];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * WG, l = e & 63, i = e >> 6, L = l0 + l;
            S v = S(0);
            if (l < lb && i < kb) v = (L > i) ? vv[q] : (L == i ? S(1) : S(0));
            sm.Vs[i * 65 + l] = v;
        }
    };
    for (int r0 = 0; r0 < m; r0 += 64) {
        const int R = r0 + rr, Rc = min(R, m - 1);
        S acc[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = S(0);
        for (int l0 = 0; l0 < n; l0 += 64) {
            const int lb = min(64, n - l0);
            __syncthreads();
            stage_v(l0, lb);
            __syncthreads();
            for (int lq = 0; lq < lb; lq += 8) {
                S c[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) c[u] = C[Rc + (size_t)(l0 + min(lq + u, lb - 1)) * ldc];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const S cu = (lq + u < lb) ? c[u] : S(0);
#pragma unroll
                    for (int q = 0; q < 8; ++q) acc[q] += cu * sm.Vs[(cg * 8 + q) * 65 + min(lq + u, 63)];
                }
            }
        }
        // W <- W T : stage W rows in Cs (Cs[i * 65 + rr])
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; ++q) sm.Cs[(cg * 8 + q) * 65 + rr] = acc[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int i = cg * 8 + q;
            S s = S(0);
            for (int l = 0; l <= i; ++l) s += sm.Cs[l * 65 + rr] * sm.Ts[i * 33 + l];
            acc[q] = (i < kb) ? s : S(0);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; ++q) sm.Cs[(cg * 8 + q) * 65 + rr] = acc[q];
        __syncthreads();
        // C -= W V^T over the same chunks: thread (rr, cg) covers chunk columns cg, cg + 4, ...
        S w[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) w[i] = sm.Cs[i * 65 + rr];
        for (int l0 = 0; l0 < n; l0 += 64) {
            const int lb = min(64, n - l0);
            __syncthreads();
            stage_v(l0, lb);
            __syncthreads();
            for (int uh = 0; uh < 16; uh += 8) {
                S cold[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) cold[u] = C[Rc + (size_t)(l0 + min(cg + 4 * (uh + u), lb - 1)) * ldc];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int l = cg + 4 * (uh + u);
                    S sacc = S(0);
#pragma unroll
                    for (int i = 0; i < 32; ++i) sacc += w[i] * sm.Vs[i * 65 + l];
                    if (R < m && l < lb) C[R + (size_t)(l0 + l) * ldc] = cold[u] - sacc;
                }
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------------
// B (m x n) <- B R^{-1}, R upper (ldr).  32-column blocks: B_j -= B_{<j} R_{<j,j} then a
// per-row triangular solve with the 32 x 32 diagonal block in LDS.
template <typename S>
__device__ __forceinline__ void wg_trsm_ru(int m, int n, const S* __restrict__ R, int ldr, S* B, int ldb, ApplySmem<S>& sm) {
    const int tid = threadIdx.x;
    const int rr = tid & 63, cg = tid >> 6;
    for (int j0 = 0; j0 < n; j0 += 32) {
        const int jb = min(32, n - j0);
        if (sizeof(S) == 4 && j0 > 0) {
            // B_j -= B_<j R_<j,j on MFMA: wave w owns rows r0 + 32 w .. of a 128-row tile; the
            // 32 x 32 block of R is staged in LDS per 32-row step of the reduction
            const int l = tid & 63, li = l & 31, lk = l >> 5, w = tid >> 6;
            for (int r0 = 0; r0 < m; r0 += 128) {
                const int row = r0 + 32 * w + li, rowc = min(row, m - 1);
                floatx16 acc;
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[r] = 0.f;
                for (int l0 = 0; l0 < j0; l0 += 32) {
                    __syncthreads();
                    for (int e = tid; e < 32 * 32; e += WG) {
                        const int ll = e & 31, q = e >> 5;
                        const S rv = R[min(l0 + ll, j0 - 1) + (size_t)(j0 + min(q, jb - 1)) * ldr];
                        sm.W[q * 33 + ll] = (l0 + ll < j0 && q < jb) ? rv : S(0);
                    }
                    __syncthreads();
                    float av[16];
#pragma unroll
                    for (int kk = 0; kk < 32; kk += 2) {
                        const int c = l0 + kk + lk;
                        const float v = (float)B[rowc + (size_t)min(c, j0 - 1) * ldb];
                        av[kk / 2] = (c < j0 && row < m) ? v : 0.f;
                    }
#pragma unroll
                    for (int kk = 0; kk < 32; kk += 2)
                        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[kk / 2], (float)sm.W[li * 33 + kk + lk], acc, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int i = r0 + 32 * w + mfma_row(l, r);
                    if (i < m && li < jb) B[i + (size_t)(j0 + li) * ldb] -= (S)acc[r];
                }
            }
        } else if (j0 > 0) {
            for (int r0 = 0; r0 < m; r0 += 64) {
                const int Rw = r0 + rr;
                S acc[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) acc[q] = S(0);
                for (int l0 = 0; l0 < j0; l0 += 32) {
                    __syncthreads();
                    for (int e = tid; e < 32 * 32; e += WG) {
                        int l = e & 31, q = e >> 5;
                        const S rv = R[min(l0 + l, j0 - 1) + (size_t)(j0 + min(q, jb - 1)) * ldr];
                        sm.W[q * 33 + l] = (l0 + l < j0 && q < jb) ? rv : S(0);
                    }
                    __syncthreads();
                    if (Rw < m) {
                        const int lb = min(32, j0 - l0);
                        S bv[32];
#pragma unroll
                        for (int l = 0; l < 32; ++l) bv[l] = B[Rw + (size_t)(l0 + min(l, lb - 1)) * ldb];
#pragma unroll
                        for (int l = 0; l < 32; ++l) {
                            const S b = (l < lb) ? bv[l] : S(0);
#pragma unroll
                            for (int q = 0; q < 8; ++q) acc[q] += b * sm.W[(cg * 8 + q) * 33 + l];
                        }
                    }
                }
                if (Rw < m) {
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (cg * 8 + q < jb) B[Rw + (size_t)(j0 + cg * 8 + q) * ldb] -= acc[q];
                }
            }
        }
        __syncthreads();
        for (int e = tid; e < 32 * 32; e += WG) {
            int i = e & 31, q = e >> 5;
            const S rv = R[(j0 + min(i, jb - 1)) + (size_t)(j0 + min(q, jb - 1)) * ldr];
            sm.Ts[q * 33 + i] = (i < jb && q < jb) ? rv : S(0);
        }
        __syncthreads();
        for (int r = tid; r < m; r += WG) {
            S x[32];
#pragma unroll
            for (int q = 0; q < 32; ++q) x[q] = B[r + (size_t)(j0 + min(q, jb - 1)) * ldb];
#pragma unroll
            for (int q = 0; q < 32; ++q) x[q] = (q < jb) ? x[q] : S(0);
#pragma unroll
            for (int q = 0; q < 32; ++q) {
                if (q < jb) {
                    S s = x[q];
#pragma unroll
                    for (int l = 0; l < q; ++l) s -= x[l] * sm.Ts[q * 33 + l];
                    x[q] = s / sm.Ts[q * 33 + q];
                }
            }
#pragma unroll
            for (int q = 0; q < 32; ++q)
                if (q < jb) B[r + (size_t)(j0 + q) * ldb] = x[q];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Blocked T (NB x k) of the unit-lower reflectors V (m x k, ldv) with scalars tau.
template <typename S>
__device__ __forceinline__ void wg_larft_blocked(int m, int k, const S* V, int ldv, const S* tau, S* Tout, MemPanelSmem<S>& sm) {
    const int tid = threadIdx.x;
    const int l = tid >> 3, g = tid & 7;
    for (int p0 = 0; p0 < k; p0 += NB) {
        const int nbp = min(NB, k - p0);
        __syncthreads();
        for (int e = tid; e < 32 * 33; e += WG) sm.Ts[e] = S(0);
        __syncthreads();
        for (int q = 0; q < nbp; ++q) {
            // G(i, q) for i < q: V(p0+q, p0+i) + sum_{r > p0+q} V(r, p0+i) V(r, p0+q); thread column l = i
            S d = S(0);
            if (l < q) {
                const S* vi = V + (size_t)(p0 + l) * ldv;
                const S* vq = V + (size_t)(p0 + q) * ldv;
                for (int r = p0 + q + 1 + g; r < m; r += 8) d += vi[r] * vq[r];
            }
            d = oct_sum(d);
            if (g == 0 && l < q) sm.t[l] = d + V[(p0 + q) + (size_t)(p0 + l) * ldv];
            __syncthreads();
            const S tq = tau[p0 + q];
            if (tid < q) {
                S s = S(0);
                for (int c = tid; c < q; ++c) s += sm.Ts[c * 33 + tid] * sm.t[c];
                sm.Ts[q * 33 + tid] = (tq == S(0)) ? S(0) : -tq * s;
            } else if (tid == q) {
                sm.Ts[q * 33 + q] = tq;
            }
            __syncthreads();
        }
        for (int e = tid; e < nbp * nbp; e += WG) {
            int i = e % nbp, jj = e / nbp;
            Tout[i + (size_t)(p0 + jj) * NB] = (i <= jj) ? sm.Ts[jj * 33 + i] : S(0);
        }
    }
    __syncthreads();
}

// =======================================================================================
// Batched kernels
// =======================================================================================

template <typename S>
__global__ __launch_bounds__(WG) void k_geqrt(const GeqrtTask<S>* __restrict__ tasks) {
    __shared__ ApplyPanelSmem<S> sm_;
    ApplySmem<S>& asm_ = sm_.ap;
    PanelSmem<S>& psm = sm_.pn;
    const GeqrtTask<S> t = tasks[blockIdx.x];
    const int kmin = min(t.m, t.n);
    for (int p0 = 0; p0 < kmin; p0 += NB) {
        const int nbp = min(NB, kmin - p0);
        S* Ap = t.A + p0 + (size_t)p0 * t.lda;
        wg_panel_qr<S, false>(t.m - p0, nbp, Ap, t.lda, Ap, t.lda, t.T + (size_t)p0 * NB, NB, psm);
        if (p0 + nbp < t.n)
            wg_apply_lt<S, false>(t.m - p0, nbp, Ap, t.lda, t.T + (size_t)p0 * NB, NB, nullptr, 0,
                                  Ap + (size_t)nbp * t.lda, t.lda, t.n - p0 - nbp, asm_);
    }
}

template <typename S, bool PENT>
__global__ __launch_bounds__(WG) void k_panel_qr(const PanelTask<S>* __restrict__ tasks) {
    __shared__ PanelSmem<S> psm;
    const PanelTask<S> t = tasks[blockIdx.x];
    if (PENT) wg_panel_qr<S, true>(t.m, t.nbp, t.A, t.lda, t.B, t.ldb, t.T, NB, psm);
    else wg_panel_qr<S, false>(t.m, t.nbp, t.A, t.lda, t.A, t.lda, t.T, NB, psm);
}

template <typename S, bool PENT>
__global__ __launch_bounds__(WG) void k_panel_apply(const PanelApplyTask<S>* __restrict__ tasks) {
    __shared__ ApplySmem<S> asm_;
    const PanelApplyTask<S> t = tasks[blockIdx.x];
    wg_apply_lt<S, PENT>(t.m, t.kb, t.V, t.ldv, t.T, NB, t.Atop, t.lda, t.C, t.ldc, t.ncols, asm_);
}

// Look-ahead step of the host-driven blocked QR: a workgroup applies the previous panel to
// its columns and, when they hold the next panel, factors that panel at once; the other
// workgroups of the launch update the remaining trailing columns meanwhile.
template <typename S>
__global__ __launch_bounds__(WG) void k_panel_step(const PanelStepTask<S>* __restrict__ tasks) {
    __shared__ ApplyPanelSmem<S> sm_;
    const PanelStepTask<S> t = tasks[blockIdx.x];
    if (t.apply)
        wg_apply_lt<S, false>(t.a.m, t.a.kb, t.a.V, t.a.ldv, t.a.T, NB, nullptr, 0, t.a.C, t.a.ldc, t.a.ncols, sm_.ap);
    if (t.qr) {
        __syncthreads();
        wg_panel_qr<S, false, true>(t.q.m, t.q.nbp, t.q.A, t.q.lda, t.q.A, t.q.lda, t.q.T, NB, sm_.pn);
    }
}

template <typename S>
__global__ __launch_bounds__(WG) void k_larfb(const LarfbTask<S>* __restrict__ tasks) {
    __shared__ ApplySmem<S> asm_;
    const LarfbTask<S> t = tasks[blockIdx.x];
    if (t.V) {
        for (int p0 = 0; p0 < t.k; p0 += NB) {
            const int nbp = min(NB, t.k - p0);
            wg_apply_lt<S, false>(t.m - p0, nbp, t.V + p0 + (size_t)p0 * t.ldv, t.ldv, t.T + (size_t)p0 * NB, NB,
                                  nullptr, 0, t.C + p0, t.ldc, t.n, asm_);
        }
    }
    if (t.R) wg_trsm_ru<S>(t.m, t.k2, t.R, t.ldr, t.C, t.ldc, asm_);
}

template <typename S>
__global__ __launch_bounds__(WG) void k_tpqrt(const TpqrtTask<S>* __restrict__ tasks,
                                              const TpqrtLink<S>* __restrict__ links) {
    __shared__ ApplyPanelSmem<S> sm_;
    ApplySmem<S>& asm_ = sm_.ap;
    PanelSmem<S>& psm = sm_.pn;
    const TpqrtTask<S> t = tasks[blockIdx.x];
    for (int li = 0; li < t.nlinks; ++li) {
        const TpqrtLink<S> L = links[t.link0 + li];
        for (int p0 = 0; p0 < t.k; p0 += NB) {
            const int nbp = min(NB, t.k - p0);
            S* Rp = t.R + p0 + (size_t)p0 * t.ldr;
            S* Bp = L.B + (size_t)p0 * L.ldb;
            wg_panel_qr<S, true>(L.rows, nbp, Rp, t.ldr, Bp, L.ldb, L.T + (size_t)p0 * NB, NB, psm);
            if (p0 + nbp < t.k)
                wg_apply_lt<S, true>(L.rows, nbp, Bp, L.ldb, L.T + (size_t)p0 * NB, NB, Rp + (size_t)nbp * t.ldr,
                                     t.ldr, Bp + (size_t)nbp * L.ldb, L.ldb, t.k - p0 - nbp, asm_);
        }
    }
}

// tpqrt chain of a small interior (k <= 64, every link 0..64 rows) on one 64-thread
// workgroup, column-major in the lanes: lane c holds column c of the link's B in registers,
// R sits in LDS (step j reads and writes row j only).  A step needs no cross-lane reduction:
// lane j forms the reflector from its own column, broadcasts it through LDS, and one
// lane-local dot product per lane gives both the trailing update (columns > j) and the
// panel's Gram column v_p^T v_j (columns < j) from which T is built at the panel's end.
// Same reflectors (sign rule included), R, V and blocked T as the workgroup kernel.
__device__ __forceinline__ int rpos(int j) { return j * 64 - (j * (j - 1)) / 2; }

template <typename S>
struct TpqrtWaveSmem {
    S R[64 * 65 / 2];  // upper triangle of R, row by row: R(j, c) at rpos(j) + c - j
    S v[64];       // reflector of the current step
    S G[32 * 33];  // G(p, j) = v_p^T v_j of the current panel at j * 33 + p
    S Ts[32 * 33]; // T(q, j) of the current panel at j * 33 + q
    S tau[32];
};

// geqrt of a small block (m, n <= 64) on one 64-thread workgroup, column-major in the lanes:
// lane c holds column c (MR >= m rows) in registers.  Step j: lane j forms the reflector
// from its own column (alpha = row j, the rows below it), and writes it to LDS as a full
// MR-vector with zeros above row j and 1 on it, so that every lane's dot product and update
// run over fixed register indices; the same dot product gives the panel's Gram column
// v_p^T v_j for the columns already reflected, from which T is built at each panel's end.
// Same reflectors (sign rule included), R, V and blocked T as the workgroup kernel.
template <typename S>
struct GeqrtWaveSmem {
    S v[64];       // reflector of the current step (zero above its row, one on it)
    S G[32 * 33];  // G(p, j) = v_p^T v_j of the current panel at j * 33 + p
    S Ts[32 * 33]; // T(q, j) of the current panel at j * 33 + q
    S tau[32];
};

template <typename S, int MR>
__global__ __launch_bounds__(64) void k_geqrt_wave(const GeqrtTask<S>* __restrict__ tasks) {
    __shared__ GeqrtWaveSmem<S> sm;
    const GeqrtTask<S> t = tasks[blockIdx.x];
    const int c = threadIdx.x, m = t.m, n = t.n, kmin = min(m, n), cc = min(c, n - 1);
    S* __restrict__ Ac = t.A + (size_t)cc * t.lda;
    S a[MR];
#pragma unroll
    for (int r = 0; r < MR; ++r) a[r] = Ac[min(r, m - 1)];
#pragma unroll
    for (int r = 0; r < MR; ++r) a[r] = (r < m && c < n) ? a[r] : S(0);
    for (int p0 = 0; p0 < kmin; p0 += NB) {
        const int kb = min(NB, kmin - p0);
        for (int jj = 0; jj < kb; ++jj) {
            const int j = p0 + jj;
            wave_sync();  // the previous step's readers of sm.v are done
            if (c == j) {
                S alpha = S(0), s0 = S(0), s1 = S(0);
#pragma unroll
                for (int r = 0; r < MR; ++r) {
                    alpha = (r == j) ? a[r] : alpha;
                    const S x = (r > j) ? a[r] : S(0);
                    if (r & 1) s1 += x * x; else s0 += x * x;
                }
                const S ss = s0 + s1;
                S tau, scal, beta;
                if (ss == S(0)) {
                    tau = S(0); scal = S(1); beta = alpha;
                } else {
                    householder_fast(alpha, ss, tau, scal, beta);
                }
#pragma unroll
                for (int r = 0; r < MR; ++r) {
                    const S v = (r > j) ? a[r] * scal : (r == j ? S(1) : S(0));
                    sm.v[r] = v;
                    a[r] = (r > j) ? v : (r == j ? beta : a[r]);
                }
                sm.tau[jj] = tau;
            }
            wave_sync();
            const S tau = sm.tau[jj];
            S d0 = S(0), d1 = S(0);
#pragma unroll
            for (int r = 0; r < MR; r += 2) {
                d0 += a[r] * sm.v[r];
                d1 += a[r + 1] * sm.v[r + 1];
            }
            const S d = d0 + d1;
            if (c > j && c < n) {
                const S w = tau * d;
#pragma unroll
                for (int r = 0; r < MR; ++r) a[r] -= w * sm.v[r];
            } else if (c >= p0 && c < j) {
                sm.G[jj * 33 + (c - p0)] = d;
            }
        }
        wave_sync();
        // the panel's T: lane q builds row q, T(q, j) = -tau_j sum_{p=q}^{j-1} T(q, p) G(p, j)
        const int q = c;
        if (q < kb) {
            S* Tq = t.T + (size_t)p0 * NB + q;
            for (int jj = 0; jj < kb; ++jj) {
                S v = S(0);
                if (q == jj) {
                    v = sm.tau[jj];
                } else if (q < jj) {
                    S s = S(0);
                    for (int p = q; p < jj; ++p) s += sm.Ts[p * 33 + q] * sm.G[jj * 33 + p];
                    v = -sm.tau[jj] * s;
                }
                sm.Ts[jj * 33 + q] = v;
                Tq[(size_t)jj * NB] = v;
            }
        }
    }
    if (c < n) {
#pragma unroll
        for (int r = 0; r < MR; ++r)
            if (r < m) Ac[r] = a[r];
    }
}

template <typename S>
__global__ __launch_bounds__(64) void k_tpqrt_wave(const TpqrtTask<S>* __restrict__ tasks,
                                                   const TpqrtLink<S>* __restrict__ links) {
    __shared__ TpqrtWaveSmem<S> sm;
    const TpqrtTask<S> t = tasks[blockIdx.x];
    const int c = threadIdx.x, k = t.k, cc = min(c, k - 1);
    for (int j = 0; j <= c && j < k; ++j) sm.R[rpos(j) + c - j] = (c < k) ? t.R[j + (size_t)cc * t.ldr] : S(0);
    for (int li = 0; li < t.nlinks; ++li) {
        const TpqrtLink<S> L = links[t.link0 + li];
        const int rows = L.rows;
        S b[64];
        if (rows > 0) {
            const S* __restrict__ Bc = L.B + (size_t)cc * L.ldb;
#pragma unroll
            for (int i = 0; i < 64; ++i) b[i] = Bc[min(i, rows - 1)];
#pragma unroll
            for (int i = 0; i < 64; ++i) b[i] = (i < rows && c < k) ? b[i] : S(0);
        } else {
#pragma unroll
            for (int i = 0; i < 64; ++i) b[i] = S(0);
        }
        for (int p0 = 0; p0 < k; p0 += NB) {
            const int kb = min(NB, k - p0);
            for (int jj = 0; jj < kb; ++jj) {
                const int j = p0 + jj;
                wave_sync();  // the previous step's readers of sm.v are done
                if (c == j) {
                    S s0 = S(0), s1 = S(0), s2 = S(0), s3 = S(0);
#pragma unroll
                    for (int i = 0; i < 64; i += 4) {
                        s0 += b[i] * b[i];
                        s1 += b[i + 1] * b[i + 1];
                        s2 += b[i + 2] * b[i + 2];
                        s3 += b[i + 3] * b[i + 3];
                    }
                    const S ss = (s0 + s1) + (s2 + s3);
                    const S alpha = sm.R[rpos(j)];
                    S tau, scal, beta;
                    if (ss == S(0)) {
                        tau = S(0); scal = S(1); beta = alpha;
                    } else {
                        householder_fast(alpha, ss, tau, scal, beta);
                    }
#pragma unroll
                    for (int i = 0; i < 64; ++i) {
                        b[i] *= scal;
                        sm.v[i] = b[i];
                    }
                    sm.R[rpos(j)] = beta;
                    sm.tau[jj] = tau;
                }
                wave_sync();
                const S tau = sm.tau[jj];
                S d0 = S(0), d1 = S(0), d2 = S(0), d3 = S(0);
#pragma unroll
                for (int i = 0; i < 64; i += 4) {
                    d0 += b[i] * sm.v[i];
                    d1 += b[i + 1] * sm.v[i + 1];
                    d2 += b[i + 2] * sm.v[i + 2];
                    d3 += b[i + 3] * sm.v[i + 3];
                }
                const S d = (d0 + d1) + (d2 + d3);
                if (c > j && c < k) {
                    const S w = tau * (sm.R[rpos(j) + c - j] + d);
                    sm.R[rpos(j) + c - j] -= w;
#pragma unroll
                    for (int i = 0; i < 64; ++i) b[i] -= w * sm.v[i];
                } else if (c >= p0 && c < j) {
                    sm.G[jj * 33 + (c - p0)] = d;
                }
            }
            wave_sync();
            // the panel's T: lane q builds row q, T(q, j) = -tau_j sum_{p=q}^{j-1} T(q, p) G(p, j);
            // a row depends on nothing but itself, so the lanes run independently
            const int q = c;
            if (q < kb) {
                S* Tq = L.T + (size_t)p0 * NB + q;
                for (int jj = 0; jj < kb; ++jj) {
                    S v = S(0);
                    if (q == jj) {
                        v = sm.tau[jj];
                    } else if (q < jj) {
                        S s = S(0);
                        for (int p = q; p < jj; ++p) s += sm.Ts[p * 33 + q] * sm.G[jj * 33 + p];
                        v = -sm.tau[jj] * s;
                    }
                    sm.Ts[jj * 33 + q] = v;
                    Tq[(size_t)jj * NB] = v;
                }
            }
        }
        if (rows > 0 && c < k) {
            S* Bc = L.B + (size_t)c * L.ldb;
#pragma unroll
            for (int i = 0; i < 64; ++i)
                if (i < rows) Bc[i] = b[i];
        }
    }
    wave_sync();
    for (int j = 0; j <= min(c, k - 1); ++j)
        if (c < k) t.R[j + (size_t)c * t.ldr] = sm.R[rpos(j) + c - j];
}

// Progress counters between the workgroups of one cooperative chain.  The producer makes its
// stores visible (agent-scope release by every thread, then the counter store); a consumer
// spins on the counter (bounded: past the limit it raises *err and goes on, so the kernel
// always ends) and every thread then acquires before reading the producer's data.
constexpr long kChainSpinLimit = 1L << 24;

__device__ __forceinline__ void chain_wait(const int* c, int target, int* err) {
    if (threadIdx.x == 0) {
        long spins = 0;
        while (__hip_atomic_load(c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > kChainSpinLimit) { atomicOr(err, 1); break; }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

__device__ __forceinline__ void chain_signal(int* c, int v) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(c, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// tpqrt chain with one workgroup per 32-column panel.  For each link, workgroup p applies the
// block reflectors of panels 0 .. p-1 of that link to its own columns (each after the panel's
// owner has published it), factors its own panel [R_pp; B(:, panel p)] and publishes it.
// Links pipeline across the panels: workgroup 0 starts link l + 1 while later panels still
// work on link l.  Same reflectors, R and T as the one-workgroup kernel.
template <typename S>
__global__ __launch_bounds__(WG) void k_tpqrt_coop(const TpqrtCoopTask<S>* __restrict__ tasks,
                                                   const int* __restrict__ blk2task,
                                                   const TpqrtLink<S>* __restrict__ links, int* err) {
    __shared__ ApplyPanelSmem<S> sm_;
    ApplySmem<S>& asm_ = sm_.ap;
    PanelSmem<S>& psm = sm_.pn;
    const TpqrtCoopTask<S> t = tasks[blk2task[blockIdx.x]];
    const int p = blockIdx.x - t.b0, p0 = p * NB, kb = min(NB, t.k - p0);
    for (int li = 0; li < t.nlinks; ++li) {
        const TpqrtLink<S> L = links[t.link0 + li];
        S* Bp = L.B + (size_t)p0 * L.ldb;
        for (int q = 0; q < p; ++q) {
            chain_wait(t.cnt + q, li + 1, err);
            const int q0 = q * NB;
            wg_apply_lt<S, true>(L.rows, NB, L.B + (size_t)q0 * L.ldb, L.ldb, L.T + (size_t)q0 * NB, NB,
                                 t.R + q0 + (size_t)p0 * t.ldr, t.ldr, Bp, L.ldb, kb, asm_);
        }
        wg_panel_qr<S, true>(L.rows, kb, t.R + p0 + (size_t)p0 * t.ldr, t.ldr, Bp, L.ldb, L.T + (size_t)p0 * NB, NB, psm);
        chain_signal(t.cnt + p, li + 1);
    }
}

// tpmqrt chain of a small task (k <= 64, links of at most MR <= 64 rows, tile of <= 64
// columns) on one 64-thread workgroup: lane c owns column c of the tile -- its column of
// B_n in registers, its column of Atop in LDS -- and the k reflectors of each link are
// applied one at a time (tau_j = T(j mod 32, j)), the reflector broadcast through LDS with
// the next one's load already in flight.  Every step is lane-local; the same transformation
// as the blocked workgroup kernel.
template <typename S, int MR>
__global__ __launch_bounds__(64) void k_tpmqrt_wave(const TpmqrtTask<S>* __restrict__ tasks,
                                                    const TpmqrtLink<S>* __restrict__ links) {
    __shared__ S At[64 * 65];  // Atop(j, c) at j * 65 + c
    __shared__ S vb[MR];       // reflector j (rows of B_n)
    const TpmqrtTask<S> t = tasks[blockIdx.x];
    const int c = threadIdx.x, k = t.k, nc = t.ncols, cc = min(c, nc - 1);
    {
        const S* __restrict__ Ac = t.Atop + (size_t)cc * t.lda;
        for (int j0 = 0; j0 < k; j0 += 8) {
            S x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = Ac[min(j0 + u, k - 1)];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j0 + u < k) At[(j0 + u) * 65 + c] = x[u];
        }
    }
    for (int li = 0; li < t.nlinks; ++li) {
        const TpmqrtLink<S> L = links[t.link0 + li];
        const int rows = L.rows, rl = min(c, rows - 1);
        S* __restrict__ Bc = L.B + (size_t)(t.col0 + cc) * L.ldb;
        S b[MR];
#pragma unroll
        for (int r = 0; r < MR; ++r) b[r] = Bc[min(r, rows - 1)];
#pragma unroll
        for (int r = 0; r < MR; ++r) b[r] = (r < rows) ? b[r] : S(0);
        const S* __restrict__ V = L.V;
        S vn = V[rl];
        for (int j = 0; j < k; ++j) {
            const S vc = vn;
            if (j + 1 < k) vn = V[rl + (size_t)(j + 1) * L.ldv];
            wave_sync();  // the previous step's readers of vb are done
            if (c < MR) vb[c] = (c < rows) ? vc : S(0);
            wave_sync();
            const S tau = L.T[(j & (NB - 1)) + (size_t)j * NB];
            S d0 = At[j * 65 + c], d1 = S(0);
#pragma unroll
            for (int r = 0; r < MR; r += 2) {
                d0 += vb[r] * b[r];
                d1 += vb[r + 1] * b[r + 1];
            }
            const S w = tau * (d0 + d1);
            At[j * 65 + c] -= w;
#pragma unroll
            for (int r = 0; r < MR; ++r) b[r] -= w * vb[r];
        }
        if (c < nc) {
#pragma unroll
            for (int r = 0; r < MR; ++r)
                if (r < rows) Bc[r] = b[r];
        }
    }
    if (c < nc) {
        S* __restrict__ Ac = t.Atop + (size_t)c * t.lda;
        for (int j = 0; j < k; ++j) Ac[j] = At[j * 65 + c];
    }
}

template <typename S>
__global__ __launch_bounds__(WG) void k_tpmqrt(const TpmqrtTask<S>* __restrict__ tasks,
                                               const TpmqrtLink<S>* __restrict__ links) {
    __shared__ ApplySmem<S> asm_;
    const TpmqrtTask<S> t = tasks[blockIdx.x];
    for (int li = 0; li < t.nlinks; ++li) {
        const TpmqrtLink<S> L = links[t.link0 + li];
        for (int p0 = 0; p0 < t.k; p0 += NB) {
            const int nbp = min(NB, t.k - p0);
            wg_apply_lt<S, true>(L.rows, nbp, L.V + (size_t)p0 * L.ldv, L.ldv, L.T + (size_t)p0 * NB, NB,
                                 t.Atop + p0, t.lda, L.B + (size_t)t.col0 * L.ldb, L.ldb, t.ncols, asm_);
        }
    }
}

// dst <- src (trans: dst <- src^T) by 32 x 32 tiles, coalesced on both sides (a transposed
// tile goes through LDS).  Workgroup (task, part) takes the task's tiles q = part mod parts,
// so large blocks spread over several workgroups.
template <typename S>
__global__ __launch_bounds__(WG) void k_copy(const CopyTask<S>* __restrict__ tasks) {
    __shared__ S tile[32][33];
    const CopyTask<S> t = tasks[blockIdx.x];
    const int tr = (t.rows + 31) / 32, tc = (t.cols + 31) / 32, nt = tr * tc;
    const int tid = threadIdx.x, lx = tid & 31, ly = tid >> 5;  // 32 x (WG / 32)
    constexpr int LY = WG / 32;
    for (int q = blockIdx.y; q < nt; q += gridDim.y) {
        const int i0 = (q % tr) * 32, j0 = (q / tr) * 32;
        if (!t.trans) {
            const int i = i0 + lx;
            for (int jj = ly; jj < 32; jj += LY) {
                const int j = j0 + jj;
                if (i < t.rows && j < t.cols) t.dst[i + (size_t)j * t.ldd] = t.src[i + (size_t)j * t.lds];
            }
        } else {
            // dst(i, j) = src(j, i): src is read along j, dst written along i
            for (int ii = ly; ii < 32; ii += LY) {
                const int i = i0 + ii, j = j0 + lx;
                if (i < t.rows && j < t.cols) tile[ii][lx] = t.src[j + (size_t)i * t.lds];
            }
            __syncthreads();
            for (int jj = ly; jj < 32; jj += LY) {
                const int i = i0 + lx, j = j0 + jj;
                if (i < t.rows && j < t.cols) t.dst[i + (size_t)j * t.ldd] = tile[lx][jj];
            }
            __syncthreads();
        }
    }
}

// Column-pivoted QR (dlaqp2 steps) of C (k x w, ld k), stopped at the first |R_jj|/|R_00| <
// tol (choose_rank, src/util.cpp:830-851, counts exactly the leading columns that pass).
template <typename S>
__global__ __launch_bounds__(WG) void k_geqp3(const Geqp3Task<S>* __restrict__ tasks, double tol, double tie_rel) {
    __shared__ PanelSmem<S> psm;
    __shared__ S red_v[4];
    __shared__ int red_i[4];
    __shared__ int sh_pvt, sh_stop;
    __shared__ S sh_tau, sh_scal, sh_r00, sh_best;
    const Geqp3Task<S> t = tasks[blockIdx.x];
    const int k = t.k, w = t.w, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    S* C = t.C;
    const int kmax = min(k, w);
    int rank;
    if (tol >= 1.0 || kmax == 0) {
        rank = 0;
    } else {
        for (int l = wv; l < w; l += 4) {
            S s = S(0);
            for (int r = lane; r < k; r += 64) { S v = C[r + (size_t)l * k]; s += v * v; }
            s = wave_sum(s);
            if (lane == 0) { S nv = sqrt(s); t.vn1[l] = nv; t.vn2[l] = nv; t.jpvt[l] = l; }
        }
        __syncthreads();
        const S tol3z = sqrt(S(sizeof(S) == 8 ? 1.1102230246251565e-16 : 5.9604644775390625e-08));
        rank = kmax;
        for (int j = 0; j < kmax; ++j) {
            S best = S(-1);
            int bi = w;
            for (int l = j + tid; l < w; l += WG) {
                S v = t.vn1[l];
                if (v > best) { best = v; bi = l; }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                S ov = __shfl_xor(best, o, 64);
                int oi = __shfl_xor(bi, o, 64);
                if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
            }
            if (lane == 0) { red_v[wv] = best; red_i[wv] = bi; }
            __syncthreads();
            if (tid == 0) {
                S b = red_v[0];
                int i0 = red_i[0];
                for (int q = 1; q < 4; ++q)
                    if (red_v[q] > b || (red_v[q] == b && red_i[q] < i0)) { b = red_v[q]; i0 = red_i[q]; }
                sh_pvt = i0;
                sh_best = b;
            }
            __syncthreads();
            if (tie_rel > 0.0) {
                // lowest index inside the tie band below the maximum
                const S thr = sh_best * (S)(1.0 - tie_rel);
                int li = w;
                for (int l = j + tid; l < w; l += WG)
                    if (t.vn1[l] >= thr) { li = l; break; }
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) li = min(li, __shfl_xor(li, o, 64));
                if (lane == 0) red_i[wv] = li;
                __syncthreads();
                if (tid == 0) sh_pvt = min(min(red_i[0], red_i[1]), min(red_i[2], red_i[3]));
                __syncthreads();
            }
            const int pvt = sh_pvt;
            if (pvt != j) {
                for (int r = tid; r < k; r += WG) {
                    S a = C[r + (size_t)pvt * k];
                    C[r + (size_t)pvt * k] = C[r + (size_t)j * k];
                    C[r + (size_t)j * k] = a;
                }
                if (tid == 0) {
                    int ip = t.jpvt[pvt]; t.jpvt[pvt] = t.jpvt[j]; t.jpvt[j] = ip;
                    t.vn1[pvt] = t.vn1[j];
                    t.vn2[pvt] = t.vn2[j];
                }
            }
            __syncthreads();
            // Householder of C(j:k, j)
            S ss = S(0);
            for (int r = j + 1 + tid; r < k; r += WG) { S v = C[r + (size_t)j * k]; ss += v * v; }
            ss = wave_sum(ss);
            if (lane == 0) red_v[wv] = ss;
            __syncthreads();
            if (tid == 0) {
                S s4 = red_v[0] + red_v[1] + red_v[2] + red_v[3];
                S alpha = C[j + (size_t)j * k];
                S tau, scal, beta;
                if (s4 == S(0)) { tau = S(0); scal = S(1); beta = alpha; }
                else {
                    beta = -householder_norm_signed(alpha, s4);
                    tau = (beta - alpha) / beta;
                    scal = S(1) / (alpha - beta);
                }
                int stop = 0;
                if (j == 0) sh_r00 = fabs(beta);
                else if (tol > 0.0 && !((double)(fabs(beta) / sh_r00) >= tol)) stop = 1;
                if (!stop) { C[j + (size_t)j * k] = beta; t.tau[j] = tau; }
                sh_stop = stop;
                sh_tau = tau;
                sh_scal = scal;
            }
            __syncthreads();
            if (sh_stop) { rank = j; break; }
            const S tau = sh_tau, scal = sh_scal;
            for (int r = j + 1 + tid; r < k; r += WG) C[r + (size_t)j * k] *= scal;
            __syncthreads();
            // apply H_j to the trailing columns and downdate their norms (one wave per column)
            const S* vj = C + (size_t)j * k;
            for (int l = j + 1 + wv; l < w; l += 4) {
                S* cl = C + (size_t)l * k;
                S dot = S(0);
                for (int r = j + 1 + lane; r < k; r += 64) dot += vj[r] * cl[r];
                dot = wave_sum(dot);
                const S cjl = cl[j];
                const S wl = (cjl + dot) * tau;
                const S rjl = cjl - wl;
                S nrm = S(0);
                for (int r = j + 1 + lane; r < k; r += 64) {
                    S v = cl[r] - vj[r] * wl;
                    cl[r] = v;
                    nrm += v * v;
                }
                nrm = wave_sum(nrm);
                if (lane == 0) {
                    cl[j] = rjl;
                    S v1 = t.vn1[l];
                    if (v1 != S(0)) {
                        S temp = fabs(rjl) / v1;
                        temp = S(1) - temp * temp;
                        temp = temp > S(0) ? temp : S(0);
                        S q = v1 / t.vn2[l];
                        S temp2 = temp * q * q;
                        if (temp2 <= tol3z) {
                            S nv = (j + 1 < k) ? sqrt(nrm) : S(0);
                            t.vn1[l] = nv;
                            t.vn2[l] = nv;
                        } else {
                            t.vn1[l] = v1 * sqrt(temp);
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    if (tid == 0) *t.rank_out = rank;
    // V = C(:, 0:rank) (unit lower used), T blocked
    for (int e = tid; e < k * rank; e += WG) t.Vout[e] = C[e];
    __syncthreads();
    if (rank > 0) wg_larft_blocked<S>(k, rank, t.Vout, k, t.tau, t.T, psm.mem);
}

// Sparsify apply + scatter into the parent block (src/ptree.cpp:1867-1897, 289-340).
template <typename S>
__global__ __launch_bounds__(WG) void k_spars_apply(const SparsApplyTask<S>* __restrict__ tasks) {
    __shared__ ApplySmem<S> asm_;
    const SparsApplyTask<S> t = tasks[blockIdx.x];
    if (t.ident) {
        for (int e = threadIdx.x; e < t.rc * t.rc; e += WG) {
            int i = e % t.rc, j = e / t.rc;
            t.dst[i + (size_t)j * t.ldd] = (i == j) ? S(1) : S(0);
        }
        return;
    }
    const bool right = t.pad == 0 || t.pad == 1, left = t.pad == 0 || t.pad == 2;
    if (t.Vn && right) {
        for (int p0 = 0; p0 < t.rn; p0 += NB) {
            const int nbp = min(NB, t.rn - p0);
            wg_apply_rn<S>(t.mc, t.kn - p0, nbp, t.Vn + p0 + (size_t)p0 * t.ldvn, t.ldvn, t.Tn + (size_t)p0 * NB, NB,
                           t.A + (size_t)p0 * t.lda, t.lda, asm_);
        }
    }
    if (!left) return;
    if (t.Vc) {
        for (int p0 = 0; p0 < t.rc; p0 += NB) {
            const int nbp = min(NB, t.rc - p0);
            wg_apply_lt<S, false>(t.kc - p0, nbp, t.Vc + p0 + (size_t)p0 * t.ldvc, t.ldvc, t.Tc + (size_t)p0 * NB, NB,
                                  nullptr, 0, t.A + p0, t.lda, t.rn, asm_);
        }
    }
    for (int e = threadIdx.x; e < t.rc * t.rn; e += WG) {
        int i = e % t.rc, j = e / t.rc;
        t.dst[i + (size_t)j * t.ldd] = t.A[i + (size_t)j * t.lda];
    }
}

// ---------------------------------------------------------------------------------------
// Small blocks (at most 64 x 64): one wave per task, four tasks per workgroup, no LDS.  The
// block sits in registers -- a lane per column when reflectors act on columns (Q^T C), a lane
// per row when they act on rows (C Q, C R^{-1}) -- so every step is lane-local; the reflector
// entries are the same for all lanes (wave-uniform reads).  Reflectors are applied one at a
// time with tau_l = T(l mod 32, l), the diagonal of the blocked T: the same transformation as
// the blocked kernels.  MR = register length (16, 32 or 64) >= the block's reflector dimension.

// x (MR entries, rows 0..m-1 live) <- H_{k-1} ... H_0 x for the geqrt reflectors V (m x k, ldv).
// Reflector l is loaded once per wave (lane r holds V(r, l), the next column already in
// flight) and its entries reach every lane through readlane.
template <typename S, int MR>
__device__ __forceinline__ void lane_apply_qt(S (&x)[MR], int m, int k, const S* __restrict__ V, int ldv,
                                              const S* __restrict__ T) {
    const int lane = threadIdx.x & 63, rl = min(lane, m - 1);
    S vn = (k > 0) ? V[rl] : S(0);
    for (int l = 0; l < k; ++l) {
        const S vc = vn;
        if (l + 1 < k) vn = V[rl + (size_t)(l + 1) * ldv];
        const S e = (lane == l) ? S(1) : ((lane > l && lane < m) ? vc : S(0));
        S w0 = S(0), w1 = S(0);
#pragma unroll
        for (int r = 0; r < MR; r += 2) {
            w0 += lane_read(e, r) * x[r];
            w1 += lane_read(e, r + 1) * x[r + 1];
        }
        const S w = (w0 + w1) * T[(l & (NB - 1)) + (size_t)l * NB];
#pragma unroll
        for (int r = 0; r < MR; ++r) x[r] -= w * lane_read(e, r);
    }
}

template <typename S, int MR, bool TRSM>
__global__ __launch_bounds__(WG) void k_larfb_small(const LarfbTask<S>* __restrict__ tasks, int n) {
    const int lane = threadIdx.x & 63, ti = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (ti >= n) return;
    const LarfbTask<S> t = tasks[ti];
    if (!TRSM) {
        // C (m x n) <- Q^T C: lane j owns column j
        const int j = lane, jc = min(j, t.n - 1);
        S* __restrict__ Cj = t.C + (size_t)jc * t.ldc;
        S x[MR];
#pragma unroll
        for (int r = 0; r < MR; ++r) x[r] = Cj[min(r, t.m - 1)];
#pragma unroll
        for (int r = 0; r < MR; ++r) x[r] = (r < t.m) ? x[r] : S(0);
        lane_apply_qt<S, MR>(x, t.m, t.k, t.V, t.ldv, t.T);
        if (j < t.n) {
#pragma unroll
            for (int r = 0; r < MR; ++r)
                if (r < t.m) Cj[r] = x[r];
        }
    } else {
        // C (m x k2) <- C R^{-1}: lane i owns row i, forward substitution along the row
        const int i = lane, ic = min(i, t.m - 1);
        S x[MR];
#pragma unroll
        for (int c = 0; c < MR; ++c) x[c] = t.C[ic + (size_t)min(c, t.k2 - 1) * t.ldc];
        // column c of R: lane l holds R(l, c), broadcast by readlane
        const int ll = min(lane, t.k2 - 1);
        S rn = t.R[ll];
#pragma unroll
        for (int c = 0; c < MR; ++c) {
            if (c < t.k2) {
                const S rc = rn;
                if (c + 1 < t.k2) rn = t.R[ll + (size_t)(c + 1) * t.ldr];
                S s = x[c];
#pragma unroll
                for (int l = 0; l < c; ++l) s -= x[l] * lane_read(rc, l);
                x[c] = s / lane_read(rc, c);
            }
        }
        if (i < t.m) {
#pragma unroll
            for (int c = 0; c < MR; ++c)
                if (c < t.k2) t.C[i + (size_t)c * t.ldc] = x[c];
        }
    }
}

template <typename S, int MR>
__global__ __launch_bounds__(WG) void k_spars_small(const SparsApplyTask<S>* __restrict__ tasks, int n) {
    const int lane = threadIdx.x & 63, ti = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (ti >= n) return;
    const SparsApplyTask<S> t = tasks[ti];
    if (t.ident) {
        if (lane < t.rc)
            for (int i = 0; i < t.rc; ++i) t.dst[i + (size_t)lane * t.ldd] = (i == lane) ? S(1) : S(0);
        return;
    }
    if (t.pad == 1) {
        // A (mc x kn) <- A Q_n: lane i owns row i; reflector l acts on columns l..kn-1
        const int i = lane, ic = min(i, t.mc - 1);
        S x[MR];
#pragma unroll
        for (int c = 0; c < MR; ++c) x[c] = t.A[ic + (size_t)min(c, t.kn - 1) * t.lda];
#pragma unroll
        for (int c = 0; c < MR; ++c) x[c] = (c < t.kn) ? x[c] : S(0);
        lane_apply_qt<S, MR>(x, t.kn, t.rn, t.Vn, t.ldvn, t.Tn);
        if (i < t.mc) {
#pragma unroll
            for (int c = 0; c < MR; ++c)
                if (c < t.kn) t.A[i + (size_t)c * t.lda] = x[c];
        }
        return;
    }
    // A[0:kc, j] <- Q_c^T A[0:kc, j] for the rn columns, then dst(0:rc, j) = A(0:rc, j)
    const int j = lane, jc = min(j, t.rn - 1);
    S* __restrict__ Aj = t.A + (size_t)jc * t.lda;
    S x[MR];
#pragma unroll
    for (int r = 0; r < MR; ++r) x[r] = Aj[min(r, t.kc - 1)];
#pragma unroll
    for (int r = 0; r < MR; ++r) x[r] = (r < t.kc) ? x[r] : S(0);
    if (t.Vc) lane_apply_qt<S, MR>(x, t.kc, t.rc, t.Vc, t.ldvc, t.Tc);
    if (j < t.rn) {
        if (t.Vc) {
#pragma unroll
            for (int r = 0; r < MR; ++r)
                if (r < t.kc) Aj[r] = x[r];
        }
#pragma unroll
        for (int r = 0; r < MR; ++r)
            if (r < t.rc) t.dst[r + (size_t)j * t.ldd] = x[r];
    }
}

template <typename S>
__global__ void k_fill(S* p, size_t n, S v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// =======================================================================================
// Preconditioner apply (vectors)
// =======================================================================================

// Every vector kernel below gives each thread a fixed set of absolute rows (row R belongs
// to thread R mod WG) for the whole task, so a thread only ever reads back entries it wrote
// itself; the threads exchange nothing but the 32 reflector coefficients of the current
// panel, through LDS.

// Global-memory view of a pointer read from a task descriptor (lets the compiler emit global
// rather than flat loads).
template <typename S>
__device__ __forceinline__ const __attribute__((address_space(1))) S* gmem(const S* p) {
    return (const __attribute__((address_space(1))) S*)p;
}

template <typename S, int NT = WG>
struct VecSmem {
    S red[NT / 64][32];  // per-wave partial column sums of one panel
    S t[32];        // panel coefficients t = T^T d (or T d)
    S xb[32];       // triangular-solve block
    S T[32 * 33];   // the panel's T block, column l at T[l * 33]
    S xb2[2][32];   // row-per-thread triangular solve: solved blocks, by block parity
    S part[NT];     // row-per-thread products: partial sums of the column phases
};

// One panel (kb <= 32 reflectors) of a block reflector H = I - V T V^T applied to a vector:
// x <- H^T x (trans) or H x.  T is the panel's upper-triangular kb x kb block (ld NB).
//   PENT = false: V is unit lower-trapezoidal (geqrt layout), its column l acting on rows
//     p0 + l .. m-1 of x (absolute rows; rows < p0 are untouched).
//   PENT = true: V = [I; Vb] (tpqrt layout): xt[p0 .. p0+kb) is the top block and Vb
//     (m rows) acts on x[0 .. m); xt entries are owned by lanes of wave 0.
// The kernels are latency-bound, so every load is unconditional (columns past kb are
// clamped to column kb-1 and weighted by zero): the 32 loads of a row issue together and the
// wave waits once.
template <typename S, bool PENT, int NT = WG>
__device__ __forceinline__ void wg_vec_panel(int m, int p0, int kb, const S* __restrict__ V, int ldv, const S* __restrict__ T,
                             S* xt, S* x, bool trans, VecSmem<S, NT>& sm) {
    constexpr int VH = sizeof(S) == 8 ? 16 : 32;  // loads in flight per row (register budget)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // the panel's T block into LDS (read in phase 2; the previous panel finished with it)
    {
        S tv[(NB * NB) / NT];
#pragma unroll
        for (int i = 0; i < (NB * NB) / NT; ++i) {
            const int e = tid + i * NT, q = e & (NB - 1), l = min(e / NB, kb - 1);
            tv[i] = T[q + (size_t)l * NB];
        }
#pragma unroll
        for (int i = 0; i < (NB * NB) / NT; ++i) {
            const int e = tid + i * NT, q = e & (NB - 1), l = e / NB;
            sm.T[l * 33 + q] = (q < kb && l < kb) ? tv[i] : S(0);
        }
    }
    S xtv = S(0);
    if (PENT && wave == 0 && lane < 32) xtv = xt[p0 + min(lane, kb - 1)];
    // phase 1: d_l = sum_R v(R, l) x[R] over this thread's rows.  Non-PENT: the kb rows
    // p0 .. p0+kb-1 hold the unit diagonal and zeros above it; rows from p0+kb on are full.
    S d[32];
#pragma unroll
    for (int l = 0; l < 32; ++l) d[l] = S(0);
    const int r0 = PENT ? 0 : p0;  // first row the panel touches
    const S* __restrict__ Vp = V + (size_t)(PENT ? 0 : p0) * ldv;  // uniform base, 32-bit offsets below
    for (int R = r0 + ((tid - r0 % NT + NT) % NT); R < m; R += NT) {
        const S xr = x[R];
        const int i = R - r0;  // non-PENT: column l of V is zero above row l and one on it
#pragma unroll
        for (int h = 0; h < 32; h += VH) {
            S v[VH];
#pragma unroll
            for (int l = 0; l < VH; ++l) v[l] = Vp[(unsigned)(R + min(h + l, kb - 1) * ldv)];
#pragma unroll
            for (int l = 0; l < VH; ++l)
                d[h + l] += (PENT || h + l < i ? v[l] : (h + l == i ? S(1) : S(0))) * xr;
        }
    }
    const S tot = wave_sum32<S>(d);
    if (vec_owner(lane)) sm.red[wave][vec_col(lane)] = tot;
    __syncthreads();
    // phase 2 (wave 0): t = T^T d or T d; columns >= kb of d are ignored (T is zero there)
    if (wave == 0) {
        S dl = S(0);
        if (lane < 32) {
#pragma unroll
            for (int w = 0; w < NT / 64; ++w) dl += sm.red[w][lane];
            dl += xtv;
        }
        const int lc = min(lane, 31);
        S tl = S(0);
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const S dq = lane_read(dl, q);
            const S tq = trans ? sm.T[lc * 33 + q] : sm.T[q * 33 + lc];
            const bool use = trans ? (q <= lc) : (q >= lc);
            tl += (use ? tq : S(0)) * dq;
        }
        if (lane < 32) sm.t[lane] = (lane < kb) ? tl : S(0);
        if (PENT && lane < kb) xt[p0 + lane] = xtv - tl;
    }
    __syncthreads();
    // phase 3: x[R] -= sum_l v(R, l) t_l  (t_l = 0 for l >= kb)
    S t[32];
#pragma unroll
    for (int l = 0; l < 32; ++l) t[l] = sm.t[l];
    for (int R = r0 + ((tid - r0 % NT + NT) % NT); R < m; R += NT) {
        const int i = R - r0;
        S s = S(0);
#pragma unroll
        for (int h = 0; h < 32; h += VH) {
            S v[VH];
#pragma unroll
            for (int l = 0; l < VH; ++l) v[l] = Vp[(unsigned)(R + min(h + l, kb - 1) * ldv)];
#pragma unroll
            for (int l = 0; l < VH; ++l) s += (PENT || h + l < i ? v[l] : (h + l == i ? S(1) : S(0))) * t[h + l];
        }
        x[R] -= s;
    }
}

// Row-per-thread panel applies (every row count <= NT): thread R holds x[R] in a register
// and its row of the current panel's 32 columns, and loads the next panel's row (and T
// share) while the current panel is applied, so a panel costs two barriers and no exposed
// load latency.
template <typename S, int NT>
struct VecRowPanel {
    S v[32];
    S tv[(NB * NB) / NT];  // this thread's share of the panel's T block
};

// Loads row r's coefficients for one panel (V: the panel's first column; columns past kb
// clamped, T zero there) and this thread's share of the panel's T block (ld NB).  PENT: the
// row as stored; else unit lower trapezoidal, i = r - p0 the row within the panel (v below
// the diagonal, 1 on it, 0 above).  Dead rows get zeros.  Only the rows of the diagonal
// block need masks, so the compares stay in that branch.
template <typename S, bool PENT, int NT>
__device__ __forceinline__ void vec_row_panel_load(const S* V, int ldv, const S* T, int kb, int r, int i, bool live,
                                                   VecRowPanel<S, NT>& g) {
    const int tid = threadIdx.x;
    int ro = r;
    asm volatile("" : "+v"(ro));  // opaque per call: keeps 32 row offsets out of callers' loops
    asm volatile("" : "+v"(i));   // and the 32 diagonal masks
    const auto Vp = gmem(V);
    if (!live || (!PENT && i < 0)) {
#pragma unroll
        for (int l = 0; l < 32; ++l) g.v[l] = S(0);
    } else {
#pragma unroll
        for (int l = 0; l < 32; ++l) g.v[l] = Vp[(unsigned)(ro + min(l, kb - 1) * ldv)];
        if (!PENT && i < 32) {
#pragma unroll
            for (int l = 0; l < 32; ++l) g.v[l] = l < i ? g.v[l] : (l == i ? S(1) : S(0));
        }
    }
    const auto Tp = gmem(T);
#pragma unroll
    for (int e0 = 0; e0 < (NB * NB) / NT; ++e0) {
        const int e = tid + e0 * NT, q = e & (NB - 1), l = min(e / NB, kb - 1);
        g.tv[e0] = Tp[q + l * NB];
    }
}

// One panel of H = I - V T V^T applied as in wg_vec_panel, from registers (g.v: this
// thread's coefficients, zero for rows the panel does not touch).  xt (PENT, LDS) is the
// top block at p0.  prefetch() runs between the reduction and the first barrier.
template <typename S, bool PENT, int NT, class F>
__device__ __forceinline__ void vec_row_panel_apply(const VecRowPanel<S, NT>& g, int kb, S& xr, S* xt, bool trans,
                                                    VecSmem<S, NT>& sm, F&& prefetch) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int e0 = 0; e0 < (NB * NB) / NT; ++e0) {
        const int e = tid + e0 * NT, q = e & (NB - 1), l = e / NB;
        sm.T[l * 33 + q] = (q < kb && l < kb) ? g.tv[e0] : S(0);
    }
    S d[32];
#pragma unroll
    for (int l = 0; l < 32; ++l) d[l] = g.v[l] * xr;
    const S tot = wave_sum32<S>(d);
    if (vec_owner(lane)) sm.red[wave][vec_col(lane)] = tot;
    __builtin_amdgcn_sched_barrier(0);  // the next panel's loads go after the reduction (registers)
    prefetch();
    __syncthreads();
    if (wave == 0) {
        S dl = S(0), xtv = S(0);
        if (lane < 32) {
#pragma unroll
            for (int w = 0; w < NT / 64; ++w) dl += sm.red[w][lane];
            if (PENT) {
                xtv = xt[min(lane, kb - 1)];
                dl += xtv;
            }
        }
        const int lc = min(lane, 31);
        S tl = S(0);
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const S dq = lane_read(dl, q);
            const S tq = trans ? sm.T[lc * 33 + q] : sm.T[q * 33 + lc];
            const bool use = trans ? (q <= lc) : (q >= lc);
            tl += (use ? tq : S(0)) * dq;
        }
        if (lane < 32) sm.t[lane] = (lane < kb) ? tl : S(0);
        if (PENT && lane < kb) xt[lane] = xtv - tl;
    }
    __syncthreads();
    S s = S(0);
#pragma unroll
    for (int l = 0; l < 32; ++l) s += g.v[l] * sm.t[l];
    xr -= s;
}

// x <- Q^T x (trans) or Q x for m <= NT rows (geqrt layout, as wg_vec_q).
template <typename S, int NT>
__device__ __forceinline__ void wg_vec_q_r1(int m, int k, const S* __restrict__ V, int ldv, const S* __restrict__ T, S* x,
                                            bool trans, VecSmem<S, NT>& sm) {
    if (k <= 0) return;
    const int tid = threadIdx.x, r = min(tid, m - 1);
    const bool live = tid < m;
    S xr = live ? x[tid] : S(0);
    const int np = (k + NB - 1) / NB;
    VecRowPanel<S, NT> cur, nxt;
    {
        const int p0 = (trans ? 0 : np - 1) * NB;
        vec_row_panel_load<S, false, NT>(V + (size_t)p0 * ldv, ldv, T + (size_t)p0 * NB, min(NB, k - p0), r, tid - p0, live,
                                         cur);
    }
    for (int j = 0; j < np; ++j) {
        const int kb = min(NB, k - (trans ? j : np - 1 - j) * NB);
        vec_row_panel_apply<S, false, NT>(cur, kb, xr, nullptr, trans, sm, [&] {
            if (j + 1 < np) {
                const int pn = (trans ? j + 1 : np - 2 - j) * NB;
                vec_row_panel_load<S, false, NT>(V + (size_t)pn * ldv, ldv, T + (size_t)pn * NB, min(NB, k - pn), r, tid - pn,
                                                 live, nxt);
            }
        });
        cur = nxt;
    }
    if (live) x[tid] = xr;
}

// x <- Q^T x (trans) or Q x for Q = H_0 H_1 ... (geqrt layout, k reflectors over m rows,
// blocked T with NB-column panels).
template <typename S, int NT = WG>
__device__ __forceinline__ void wg_vec_q(int m, int k, const S* __restrict__ V, int ldv, const S* __restrict__ T, S* x, bool trans,
                         VecSmem<S, NT>& sm) {
    if (k <= 0) return;
    if (m <= NT) {
        wg_vec_q_r1<S, NT>(m, k, V, ldv, T, x, trans, sm);
        return;
    }
    if (trans) {
        for (int p0 = 0; p0 < k; p0 += NB)
            wg_vec_panel<S, false, NT>(m, p0, min(NB, k - p0), V, ldv, T + (size_t)p0 * NB, nullptr, x, true, sm);
    } else {
        for (int p0 = ((k - 1) / NB) * NB; p0 >= 0; p0 -= NB)
            wg_vec_panel<S, false, NT>(m, p0, min(NB, k - p0), V, ldv, T + (size_t)p0 * NB, nullptr, x, false, sm);
    }
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void k_vec_q(const VecQTask<S>* __restrict__ tasks) {
    __shared__ VecSmem<S, NT> sm;
    const VecQTask<S> t = tasks[blockIdx.x];
    wg_vec_q<S, NT>(t.m, t.k, t.V, t.ldv, t.T, t.x, t.trans != 0, sm);
}

constexpr int kChainTopMax = 1024;  // largest chain top block kept in LDS by k_vec_chain

template <typename S, int NT>
__global__ __launch_bounds__(NT) void k_vec_chain(const VecChainTask<S>* __restrict__ tasks,
                                                  const VecChainLink<S>* __restrict__ links) {
    __shared__ VecSmem<S, NT> sm;
    __shared__ S xc[kChainTopMax];
    const VecChainTask<S> t = tasks[blockIdx.x];
    if (t.k <= 0) return;
    if (t.rmax <= NT && t.k <= kChainTopMax) {
        // row-per-thread: the top block in LDS, each link's rows in registers, the next
        // panel (next link's first panel and x rows included) loaded under the current one
        const int tid = threadIdx.x, np = (t.k + NB - 1) / NB;
        for (int i = tid; i < t.k; i += NT) xc[i] = t.xc[i];
        VecChainLink<S> L = links[t.link0];
        bool live = tid < L.rows;
        S xr = live ? L.xn[tid] : S(0), xnext = S(0);
        VecRowPanel<S, NT> cur, nxt;
        vec_row_panel_load<S, true, NT>(L.V, L.ldv, L.T, min(NB, t.k), min(tid, L.rows - 1), 0, live, cur);
        for (int li = 0; li < t.nlinks; ++li) {
            const bool more = li + 1 < t.nlinks;
            const VecChainLink<S> Ln = links[t.link0 + (more ? li + 1 : li)];
            for (int pi = 0; pi < np; ++pi) {
                const int p0 = pi * NB, kb = min(NB, t.k - p0);
                vec_row_panel_apply<S, true, NT>(cur, kb, xr, xc + p0, true, sm, [&] {
                    if (pi + 1 < np) {
                        const int pn = p0 + NB;
                        vec_row_panel_load<S, true, NT>(L.V + (size_t)pn * L.ldv, L.ldv, L.T + (size_t)pn * NB,
                                                        min(NB, t.k - pn), min(tid, L.rows - 1), 0, live, nxt);
                    } else if (more) {
                        const bool ln = tid < Ln.rows;
                        vec_row_panel_load<S, true, NT>(Ln.V, Ln.ldv, Ln.T, min(NB, t.k), min(tid, Ln.rows - 1), 0, ln, nxt);
                        xnext = ln ? Ln.xn[tid] : S(0);
                    }
                });
                cur = nxt;
            }
            if (live) L.xn[tid] = xr;
            if (more) {
                L = Ln;
                live = tid < L.rows;
                xr = xnext;
            }
        }
        __syncthreads();
        for (int i = tid; i < t.k; i += NT) t.xc[i] = xc[i];
        return;
    }
    for (int li = 0; li < t.nlinks; ++li) {
        const VecChainLink<S> L = links[t.link0 + li];
        for (int p0 = 0; p0 < t.k; p0 += NB)
            wg_vec_panel<S, true, NT>(L.rows, p0, min(NB, t.k - p0), L.V + (size_t)p0 * L.ldv, L.ldv,
                                  L.T + (size_t)p0 * NB, t.xc, L.xn, true, sm);
    }
}

// One link of a chain held in one wave's registers: lane R holds row R of Vb (all k <= 64
// columns) and x_n[R]; lane j holds column j % 32 of panel j / 32's T block, masked to its
// upper triangle (zero past k).
template <typename S>
struct ChainLinkRegs {
    S v[64];
    S tc[NB];
    S xr;
};

template <typename S>
__device__ __forceinline__ void chain_link_load(const VecChainLink<S>& L, int k, int lane, ChainLinkRegs<S>& g) {
    const int r = min(lane, L.rows - 1);
    const auto V = gmem(L.V);
#pragma unroll
    for (int l = 0; l < 64; ++l) g.v[l] = V[(unsigned)(r + min(l, k - 1) * L.ldv)];
    // every T load is issued; the mask multiplies (a select would let the compiler branch
    // around each load and wait on it)
    const int j = min(lane, k - 1), jc = lane & (NB - 1);
    const auto Tj = gmem(L.T) + (size_t)j * NB;
    const bool live = lane < k;
#pragma unroll
    for (int q = 0; q < NB; ++q) g.tc[q] = Tj[q] * ((live && q <= jc) ? S(1) : S(0));
    g.xr = gmem(L.xn)[r];
}

// x_c <- H^T x_c over a chain of tpqrt links (the forward solve's elimination of an interior
// into its neighbours), one wave per chain, for chains whose k <= 64 and whose links have
// <= 64 rows.  The wave uses no workgroup barrier: the 32 panel coefficients pass through
// the wave's own LDS rows.  The next link's V, T and x_n are loaded before the current link
// is applied (FP32; FP64 would spill, so there each link loads its own data first), so each
// link costs one load latency at most.
template <typename S>
__global__ __launch_bounds__(WG) void k_vec_chain_wave(const VecChainTask<S>* __restrict__ tasks,
                                                       const VecChainLink<S>* __restrict__ links, int n) {
    __shared__ S sxc[4][64];  // x_c of the wave's chain
    __shared__ S sd[4][NB];   // panel d = x_c(panel) + Vb^T x_n
    __shared__ S st[4][NB];   // panel t = T^T d
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ti = blockIdx.x * 4 + wave;
    if (ti >= n) return;  // whole waves leave; nothing below synchronises across waves
    const VecChainTask<S> t = tasks[ti];
    const int k = t.k;
    sxc[wave][lane] = (lane < k) ? t.xc[min(lane, k - 1)] : S(0);  // zero past k: masked products stay finite
    constexpr bool PF = sizeof(S) == 4;
    ChainLinkRegs<S> cur, nxt;
    if (PF) chain_link_load<S>(links[t.link0], k, lane, cur);
    for (int li = 0; li < t.nlinks; ++li) {
        const VecChainLink<S> L = links[t.link0 + li];
        if (!PF) chain_link_load<S>(L, k, lane, cur);
        if (PF && li + 1 < t.nlinks) chain_link_load<S>(links[t.link0 + li + 1], k, lane, nxt);
        const bool rv = lane < L.rows;
        S xr = rv ? cur.xr : S(0);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            if (p * NB >= k) break;
            S d[NB];
#pragma unroll
            for (int l = 0; l < NB; ++l) d[l] = cur.v[p * NB + l] * xr;
            const S tot = wave_sum32<S>(d);
            if (vec_owner(lane)) sd[wave][vec_col(lane)] = tot + sxc[wave][p * NB + vec_col(lane)];
            wave_sync();
            if ((lane >> 5) == p) {
                S tj = S(0);
#pragma unroll
                for (int q = 0; q < NB; ++q) tj += cur.tc[q] * sd[wave][q];
                st[wave][lane & (NB - 1)] = tj;
                if (lane < k) sxc[wave][lane] -= tj;
            }
            wave_sync();
            S s = S(0);
#pragma unroll
            for (int l = 0; l < NB; ++l) s += cur.v[p * NB + l] * st[wave][l];
            if (rv) xr -= s;  // lanes past the link's rows keep x = 0 for the next panel's sums
            wave_sync();  // sd / st are rewritten by the next panel
        }
        if (rv) L.xn[lane] = xr;
        if (PF && li + 1 < t.nlinks) cur = nxt;
    }
    wave_sync();
    if (lane < k) t.xc[lane] = sxc[wave][lane];
}

// Chains too large for one wave (k <= kChainPipeK, links of up to 256 * NRW rows): one
// workgroup per chain, the chain flattened into a stream of (link, panel) items.
//  - Thread tid owns rows tid + 256 i (i < NRW) of every link, and keeps x_n for them in
//    registers for the whole link.
//  - Item i's V rows and T block are loaded NBUF - 1 items ahead into a ring of register
//    buffers, so HBM latency overlaps the panels in between.
//  - One workgroup barrier per panel: after it, every wave recomputes the 32 coefficients
//    t = T^T d from the reduced partials itself (identical arithmetic in every wave) and keeps
//    its own copy of x_c in LDS, so no second barrier is needed to publish t.  The partial
//    sums and the T block are double-buffered in LDS: a wave can only be one barrier ahead.
constexpr int kChainPipeK = 512;

template <typename S, int NRW>
struct ChainItemRegs {
    S v[NRW][NB];       // V rows of the panel
    S tb[NB * NB / WG];  // this thread's share of the panel's T block
    S x[NRW];           // x_n rows (used at the link's first panel)
};

template <typename S>
struct ChainPipeSmem {
    S red[2][4][NB];       // per-wave partial column sums
    S Ts[2][NB * (NB + 1)];  // the panel's T block, column l at l * 33 (masked to its triangle)
    S dw[4][NB];           // per-wave full d
    S tw[4][NB];           // per-wave t
    S xc[4][kChainPipeK];  // per-wave copy of x_c
};

template <typename S, int NRW>
__device__ __forceinline__ void chain_item_load(const VecChainLink<S>* __restrict__ links, int link0, int np, int k,
                                                int item, ChainItemRegs<S, NRW>& g) {
    const int li = item / np, p0 = (item - li * np) * NB, kb = min(NB, k - p0);
    const VecChainLink<S> L = links[link0 + li];
    const int tid = threadIdx.x;
    const auto V = gmem(L.V) + (size_t)p0 * L.ldv;
#pragma unroll
    for (int i = 0; i < NRW; ++i) {
        const int r = min(tid + i * WG, L.rows - 1);
#pragma unroll
        for (int l = 0; l < NB; ++l) g.v[i][l] = V[(unsigned)(r + min(l, kb - 1) * L.ldv)];
        g.x[i] = gmem(L.xn)[r];
    }
    const auto T = gmem(L.T) + (size_t)p0 * NB;
#pragma unroll
    for (int e = 0; e < NB * NB / WG; ++e) {
        const int idx = tid + e * WG, q = idx & (NB - 1), l = idx / NB;
        g.tb[e] = T[q + min(l, kb - 1) * NB] * ((q <= l && l < kb) ? S(1) : S(0));
    }
}

template <typename S, int NRW>
__device__ __forceinline__ void chain_item_apply(const VecChainLink<S>* __restrict__ links, int link0, int np, int k,
                                                 int item, const ChainItemRegs<S, NRW>& g, S (&x)[NRW],
                                                 ChainPipeSmem<S>& sm) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, slot = item & 1;
    const int li = item / np, pi = item - li * np, p0 = pi * NB;
    const int rows = links[link0 + li].rows;
    if (pi == 0) {
#pragma unroll
        for (int i = 0; i < NRW; ++i) x[i] = (tid + i * WG < rows) ? g.x[i] : S(0);
    }
#pragma unroll
    for (int e = 0; e < NB * NB / WG; ++e) {
        const int idx = tid + e * WG, q = idx & (NB - 1), l = idx / NB;
        sm.Ts[slot][l * (NB + 1) + q] = g.tb[e];
    }
    S d[NB];
#pragma unroll
    for (int l = 0; l < NB; ++l) {
        S a = S(0);
#pragma unroll
        for (int i = 0; i < NRW; ++i) a += g.v[i][l] * x[i];
        d[l] = a;
    }
    const S tot = wave_sum32<S>(d);
    if (vec_owner(lane)) sm.red[slot][wave][vec_col(lane)] = tot;
    __syncthreads();
    const int j = lane & (NB - 1);
    const S dj = ((sm.red[slot][0][j] + sm.red[slot][1][j]) + (sm.red[slot][2][j] + sm.red[slot][3][j])) +
                 sm.xc[wave][p0 + j];
    if (lane < NB) sm.dw[wave][j] = dj;
    wave_sync();
    S tj = S(0);
#pragma unroll
    for (int q = 0; q < NB; ++q) tj += sm.Ts[slot][j * (NB + 1) + q] * sm.dw[wave][q];
    if (lane < NB) {
        sm.tw[wave][j] = tj;
        sm.xc[wave][p0 + j] -= tj;  // t_j = 0 past the panel (T is masked there)
    }
    wave_sync();
    S tl[NB];
#pragma unroll
    for (int l = 0; l < NB; ++l) tl[l] = sm.tw[wave][l];
#pragma unroll
    for (int i = 0; i < NRW; ++i) {
        S s = S(0);
#pragma unroll
        for (int l = 0; l < NB; ++l) s += g.v[i][l] * tl[l];
        if (tid + i * WG < rows) x[i] -= s;
    }
    if (pi == np - 1) {
        S* xn = links[link0 + li].xn;
#pragma unroll
        for (int i = 0; i < NRW; ++i)
            if (tid + i * WG < rows) xn[tid + i * WG] = x[i];
    }
    wave_sync();  // dw / tw are rewritten by the next item
}

template <typename S, int NRW, int NBUF>
__global__ __launch_bounds__(WG) void k_vec_chain_pipe(const VecChainTask<S>* __restrict__ tasks,
                                                       const VecChainLink<S>* __restrict__ links) {
    __shared__ ChainPipeSmem<S> sm;
    const VecChainTask<S> t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int k = t.k, np = (k + NB - 1) / NB, n = t.nlinks * np;
    for (int i = lane; i < np * NB; i += 64) sm.xc[wave][i] = (i < k) ? t.xc[i] : S(0);
    ChainItemRegs<S, NRW> buf[NBUF];
#pragma unroll
    for (int b = 0; b < NBUF; ++b)
        if (b < n) chain_item_load<S, NRW>(links, t.link0, np, k, b, buf[b]);
    S x[NRW];
#pragma unroll
    for (int i = 0; i < NRW; ++i) x[i] = S(0);
    for (int i0 = 0; i0 < n; i0 += NBUF) {
#pragma unroll
        for (int b = 0; b < NBUF; ++b) {
            const int item = i0 + b;
            if (item < n) {
                chain_item_apply<S, NRW>(links, t.link0, np, k, item, buf[b], x, sm);
                if (item + NBUF < n) chain_item_load<S, NRW>(links, t.link0, np, k, item + NBUF, buf[b]);
            }
        }
    }
    if (wave == 0)
        for (int i = lane; i < k; i += 64) t.xc[i] = sm.xc[0][i];
}

// x[0:k] <- R^{-1} x[0:k], R upper triangular (ldr): 32-row blocks from the bottom; wave 0
// solves the diagonal block with shuffles, then every thread updates its own rows above.
template <typename S, int NT = WG>
__device__ __forceinline__ void wg_trsv_upper(int k, const S* __restrict__ R, int ldr, S* x, VecSmem<S, NT>& sm) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int p0 = ((k - 1) / NB) * NB; p0 >= 0 && k > 0; p0 -= NB) {
        const int kb = min(NB, k - p0);
        const S* __restrict__ Rc = R + (size_t)p0 * ldr;  // the block's columns
        // wave 0 loads its rows of the diagonal block while the right-hand side is gathered
        S row[32];
        const int lr = min(lane, kb - 1);
        if (wave == 0) {
#pragma unroll
            for (int j = 0; j < 32; ++j) row[j] = Rc[(unsigned)(p0 + lr + min(j, kb - 1) * ldr)];
        }
        for (int r = p0 + ((tid - p0 % NT + NT) % NT); r < p0 + kb; r += NT) sm.xb[r - p0] = x[r];
        __syncthreads();
        if (wave == 0) {
            S xi = (lane < kb) ? sm.xb[lr] : S(0);
#pragma unroll
            for (int j = 31; j >= 0; --j) {
                if (j < kb) {
                    if (lane == j) xi = xi / row[j];
                    const S xj = __shfl(xi, j, 64);
                    if (lane < j) xi -= row[j] * xj;
                }
            }
            if (lane < 32) sm.xb[lane] = (lane < kb) ? xi : S(0);
        }
        __syncthreads();
        for (int r = tid; r < p0 + kb; r += NT) {
            if (r >= p0) {
                x[r] = sm.xb[r - p0];
            } else {
                S v[32];
#pragma unroll
                for (int j = 0; j < 32; ++j) v[j] = Rc[(unsigned)(r + min(j, kb - 1) * ldr)];
                S s = S(0);
#pragma unroll
                for (int j = 0; j < 32; ++j) s += v[j] * sm.xb[j];
                x[r] -= s;
            }
        }
        __syncthreads();  // sm.xb is rewritten by the next block
    }
}

// Row-per-thread forms for k <= NT: thread r holds x[r] in a register for the whole task.

// Returns sum_j A_j[r, 0:s_j] y_j over a task's links for thread r < k (0 elsewhere).  All
// waves work: wave w takes the 64-row block w % RB (RB = ceil(k / 64)) and every CP-th
// 32-column chunk of each link (CP = NW / RB column phases), so each wave keeps 32 coalesced
// column loads in flight; the CP partial sums of a row meet in LDS in a fixed order.
template <typename S, int NT>
__device__ __forceinline__ S wg_gemv_links_r1(int k, const VecGemvLink<S>* __restrict__ links, int link0, int nlinks,
                                              VecSmem<S, NT>& sm) {
    constexpr int NW = NT / 64, U = 32;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int RB = (k + 63) >> 6, CP = NW / RB;
    const int rb = wave % RB, cp = wave / RB;
    const int row = rb * 64 + lane, rr = min(row, k - 1);
    S acc = S(0);
    if (cp < CP) {
        for (int li = 0; li < nlinks; ++li) {
            const VecGemvLink<S> L = links[link0 + li];
            const auto A = gmem(L.A);
            const auto Y = gmem(L.y);
            for (int c0 = cp * U; c0 < L.s; c0 += CP * U) {
                S a[U], y[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int c = min(c0 + u, L.s - 1);
                    a[u] = A[(unsigned)(rr + c * L.lda)];
                    y[u] = Y[c];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) acc += (c0 + u < L.s ? y[u] : S(0)) * a[u];
            }
        }
        sm.part[cp * (RB * 64) + row] = acc;
    }
    __syncthreads();
    S tot = S(0);
    if (tid < k)
        for (int c = 0; c < CP; ++c) tot += sm.part[c * (RB * 64) + tid];
    return tot;
}

// x <- R^{-1} x for R upper triangular (k x k, ldr, k <= NT), x[r] in thread r's xr.
// 32-row blocks from the bottom.  Each thread holds its row of the current block's 32
// columns and loads the next block's while this one is solved.  The diagonal block's rows
// p0 .. p0+31 lie in one wave, which solves it by lane reads (no LDS, no barrier); the rows
// above subtract its solution, read from LDS (double-buffered by block parity, so one
// barrier per block).
template <typename S, int NT>
__device__ __forceinline__ void wg_trsv_upper_r1(int k, const S* __restrict__ R, int ldr, S& xr, VecSmem<S, NT>& sm) {
    if (k <= 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = min(tid, k - 1);
    int p0 = ((k - 1) / NB) * NB;
    S v[32], vn[32];
    {
        const int kb = k - p0;
        const auto Rc = gmem(R) + (size_t)p0 * ldr;
#pragma unroll
        for (int j = 0; j < 32; ++j) v[j] = Rc[(unsigned)(r + min(j, kb - 1) * ldr)];
    }
    for (int par = 0; p0 >= 0; p0 -= NB, par ^= 1) {
        const int kb = min(NB, k - p0);  // only the first (bottom) block is partial
        if (p0 > 0 && tid < p0) {  // rows from p0 on are solved
            int ro = r;
            asm volatile("" : "+v"(ro));  // opaque per block: keeps 32 row offsets out of the loop
            const auto Rn = gmem(R) + (size_t)(p0 - NB) * ldr;
#pragma unroll
            for (int j = 0; j < 32; ++j) vn[j] = Rn[(unsigned)(ro + j * ldr)];
        }
        if (wave == (p0 >> 6)) {
            const int base = p0 & 63;
            const int jl = lane - base;  // this lane's row within the block
            S xi = xr;
#pragma unroll
            for (int j = 31; j >= 0; --j) {
                if (j < kb) {
                    const S xj = lane_read(xi, base + j) / lane_read(v[j], base + j);
                    xi = (jl == j) ? xj : ((jl >= 0 && jl < j) ? xi - v[j] * xj : xi);
                }
            }
            if (jl >= 0 && jl < 32) sm.xb2[par][jl] = (jl < kb) ? xi : S(0);
            if (jl >= 0 && jl < kb) xr = xi;
        }
        __syncthreads();
        if (tid < p0) {
            S s = S(0);
#pragma unroll
            for (int j = 0; j < 32; ++j) s += v[j] * sm.xb2[par][j];
            xr -= s;
        }
#pragma unroll
        for (int j = 0; j < 32; ++j) v[j] = vn[j];
    }
}

template <typename S, int NT, bool FWD>
__global__ __launch_bounds__(NT) void k_vec_spars(const VecSparsTask<S>* __restrict__ tasks) {
    __shared__ VecSmem<S, NT> sm;
    const VecSparsTask<S> t = tasks[blockIdx.x];
    if constexpr (FWD) {
        wg_vec_q<S, NT>(t.ldq, t.k, t.Qs, t.ldq, t.Ts, t.x, true, sm);
        __syncthreads();
        wg_vec_q<S, NT>(t.k, t.r, t.Qsp, t.k, t.Tsp, t.x, true, sm);
    } else {
        wg_vec_q<S, NT>(t.k, t.r, t.Qsp, t.k, t.Tsp, t.x, false, sm);
        __syncthreads();
        if (t.k <= NT) {
            const int tid = threadIdx.x;
            S xr = tid < t.k ? t.x[tid] : S(0);
            wg_trsv_upper_r1<S, NT>(t.k, t.Qs, t.ldq, xr, sm);
            if (tid < t.k) t.x[tid] = xr;
        } else {
            wg_trsv_upper<S, NT>(t.k, t.Qs, t.ldq, t.x, sm);
        }
    }
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void k_vec_trsv(const VecTrsvTask<S>* __restrict__ tasks,
                                                 const VecGemvLink<S>* __restrict__ links) {
    __shared__ VecSmem<S, NT> sm;
    const VecTrsvTask<S> t = tasks[blockIdx.x];
    if (t.k <= NT) {
        const int tid = threadIdx.x;
        S xr = tid < t.k ? t.x[tid] : S(0);
        if (t.nlinks > 0) xr -= wg_gemv_links_r1<S, NT>(t.k, links, t.link0, t.nlinks, sm);
        wg_trsv_upper_r1<S, NT>(t.k, t.R, t.ldr, xr, sm);
        if (tid < t.k) t.x[tid] = xr;
        return;
    }
    for (int li = 0; li < t.nlinks; ++li) {
        const VecGemvLink<S> L = links[t.link0 + li];
        if (L.s <= 0) continue;
        for (int i = threadIdx.x; i < t.k; i += NT) {
            S s = S(0);
            for (int j0 = 0; j0 < L.s; j0 += 16) {
                S a[16], y[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int jj = min(j0 + j, L.s - 1);
                    a[j] = L.A[(unsigned)(i + jj * L.lda)];
                    y[j] = (j0 + j < L.s) ? L.y[jj] : S(0);
                }
#pragma unroll
                for (int j = 0; j < 16; ++j) s += a[j] * y[j];
            }
            t.x[i] -= s;
        }
    }
    wg_trsv_upper<S, NT>(t.k, t.R, t.ldr, t.x, sm);
}

template <typename S>
__global__ __launch_bounds__(WG) void k_vec_copy(const VecCopyTask<S>* __restrict__ tasks) {
    const VecCopyTask<S> t = tasks[blockIdx.x];
    for (int i = threadIdx.x; i < t.n; i += WG) t.dst[i] = t.src[i];
}

// =======================================================================================
// FP64 Krylov pieces (GMRES stays FP64 in every mode)
// =======================================================================================

__global__ __launch_bounds__(WG) void k_spmv_csr(int n, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                 const double* __restrict__ val, const double* __restrict__ x,
                                                 double* __restrict__ y) {
    int row = blockIdx.x * WG + threadIdx.x;
    if (row >= n) return;
    double s = 0.0;
    for (int k = rowptr[row]; k < rowptr[row + 1]; ++k) s += val[k] * x[col[k]];
    y[row] = s;
}

// partial[b] = sum over this block's slice of a .* b; fixed grid => deterministic
__global__ __launch_bounds__(WG) void k_dot_partial(int n, const double* __restrict__ a, const double* __restrict__ b,
                                                    double* __restrict__ partial) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) s += a[i] * b[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(WG) void k_dot_final(int nb, const double* __restrict__ partial, double* out, int mode) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += WG) s += partial[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red[0] + red[1] + red[2] + red[3];
        *out = mode ? sqrt(v) : v;
    }
}

// y <- y - (*h) x
__global__ __launch_bounds__(WG) void k_axpy_dev(int n, const double* __restrict__ h, const double* __restrict__ x,
                                                 double* __restrict__ y) {
    const double a = *h;
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) y[i] -= a * x[i];
}

// y <- x / (*h)
__global__ __launch_bounds__(WG) void k_scal_dev(int n, const double* __restrict__ h, const double* __restrict__ x,
                                                 double* __restrict__ y) {
    const double a = *h;
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) y[i] = x[i] / a;
}

// y <- a x + b y (host scalars)
__global__ __launch_bounds__(WG) void k_axpby(int n, double a, const double* __restrict__ x, double b,
                                              double* __restrict__ y) {
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) y[i] = a * x[i] + b * y[i];
}

template <typename S>
__global__ __launch_bounds__(WG) void k_convert_d2s(int n, const double* __restrict__ x, S* __restrict__ y) {
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) y[i] = (S)x[i];
}
template <typename S>
__global__ __launch_bounds__(WG) void k_convert_s2d(int n, const S* __restrict__ x, double* __restrict__ y) {
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) y[i] = (double)x[i];
}

}  // namespace spaqr

// =======================================================================================
// Host-side launchers (C++ linkage, used by the planner and the C ABI)
// =======================================================================================
#include "launch.hpp"

namespace spaqr {

template <typename S>
void launch_geqrt(const GeqrtTask<S>* d, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_geqrt<S>, dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_panel_qr(const PanelTask<S>* d, int n, bool pent, hipStream_t s) {
    if (n <= 0) return;
    if (pent) hipLaunchKernelGGL((k_panel_qr<S, true>), dim3(n), dim3(WG), 0, s, d);
    else hipLaunchKernelGGL((k_panel_qr<S, false>), dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_panel_step(const PanelStepTask<S>* d, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_panel_step<S>, dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_panel_apply(const PanelApplyTask<S>* d, int n, bool pent, hipStream_t s) {
    if (n <= 0) return;
    if (pent) hipLaunchKernelGGL((k_panel_apply<S, true>), dim3(n), dim3(WG), 0, s, d);
    else hipLaunchKernelGGL((k_panel_apply<S, false>), dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_larfb(const LarfbTask<S>* d, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_larfb<S>, dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_tpqrt(const TpqrtTask<S>* d, const TpqrtLink<S>* l, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_tpqrt<S>, dim3(n), dim3(WG), 0, s, d, l);
}
template <typename S>
void launch_tpmqrt_wave(const TpmqrtTask<S>* d, const TpmqrtLink<S>* l, int n, int mr, hipStream_t s) {
    if (n <= 0) return;
    if (mr <= 16) hipLaunchKernelGGL((k_tpmqrt_wave<S, 16>), dim3(n), dim3(64), 0, s, d, l);
    else if (mr <= 32) hipLaunchKernelGGL((k_tpmqrt_wave<S, 32>), dim3(n), dim3(64), 0, s, d, l);
    else hipLaunchKernelGGL((k_tpmqrt_wave<S, 64>), dim3(n), dim3(64), 0, s, d, l);
}
template <typename S>
void launch_geqrt_wave(const GeqrtTask<S>* d, int n, int mr, hipStream_t s) {
    if (n <= 0) return;
    if (mr <= 16) hipLaunchKernelGGL((k_geqrt_wave<S, 16>), dim3(n), dim3(64), 0, s, d);
    else if (mr <= 32) hipLaunchKernelGGL((k_geqrt_wave<S, 32>), dim3(n), dim3(64), 0, s, d);
    else hipLaunchKernelGGL((k_geqrt_wave<S, 64>), dim3(n), dim3(64), 0, s, d);
}
template <typename S>
void launch_tpqrt_coop(const TpqrtCoopTask<S>* d, const int* blk2task, int nblocks, const TpqrtLink<S>* l, int* err,
                       hipStream_t s) {
    if (nblocks > 0) hipLaunchKernelGGL(k_tpqrt_coop<S>, dim3(nblocks), dim3(WG), 0, s, d, blk2task, l, err);
}
template <typename S>
int tpqrt_coop_max_blocks() {
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tpqrt_coop<S>, WG, 0) != hipSuccess) return 0;
    return per_cu * cus;
}
template <typename S>
void launch_tpqrt_wave(const TpqrtTask<S>* d, const TpqrtLink<S>* l, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_tpqrt_wave<S>, dim3(n), dim3(64), 0, s, d, l);
}
template <typename S>
void launch_tpmqrt(const TpmqrtTask<S>* d, const TpmqrtLink<S>* l, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_tpmqrt<S>, dim3(n), dim3(WG), 0, s, d, l);
}
template <typename S>
void launch_copy(const CopyTask<S>* d, int n, hipStream_t s, int parts) {
    if (n > 0) hipLaunchKernelGGL(k_copy<S>, dim3(n, std::max(1, std::min(parts, 64))), dim3(WG), 0, s, d);
}
template <typename S>
void launch_geqp3(const Geqp3Task<S>* d, int n, double tol, double tie_rel, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_geqp3<S>, dim3(n), dim3(WG), 0, s, d, tol, tie_rel);
}
template <typename S>
void launch_spars_apply(const SparsApplyTask<S>* d, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_spars_apply<S>, dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_larfb_small(const LarfbTask<S>* d, int n, int mr, bool trsm, hipStream_t s) {
    if (n <= 0) return;
    const dim3 g((n + 3) / 4);
    if (trsm) {
        if (mr <= 16) hipLaunchKernelGGL((k_larfb_small<S, 16, true>), g, dim3(WG), 0, s, d, n);
        else if (mr <= 32) hipLaunchKernelGGL((k_larfb_small<S, 32, true>), g, dim3(WG), 0, s, d, n);
        else hipLaunchKernelGGL((k_larfb_small<S, 64, true>), g, dim3(WG), 0, s, d, n);
    } else {
        if (mr <= 16) hipLaunchKernelGGL((k_larfb_small<S, 16, false>), g, dim3(WG), 0, s, d, n);
        else if (mr <= 32) hipLaunchKernelGGL((k_larfb_small<S, 32, false>), g, dim3(WG), 0, s, d, n);
        else hipLaunchKernelGGL((k_larfb_small<S, 64, false>), g, dim3(WG), 0, s, d, n);
    }
}
template <typename S>
void launch_spars_small(const SparsApplyTask<S>* d, int n, int mr, hipStream_t s) {
    if (n <= 0) return;
    const dim3 g((n + 3) / 4);
    if (mr <= 16) hipLaunchKernelGGL((k_spars_small<S, 16>), g, dim3(WG), 0, s, d, n);
    else if (mr <= 32) hipLaunchKernelGGL((k_spars_small<S, 32>), g, dim3(WG), 0, s, d, n);
    else hipLaunchKernelGGL((k_spars_small<S, 64>), g, dim3(WG), 0, s, d, n);
}
template <typename S>
void launch_fill(S* p, size_t n, S v, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + WG - 1) / WG;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill<S>, dim3((unsigned)blocks), dim3(WG), 0, s, p, n, v);
}
// Launches of at most kVecWideTasks tasks (the top levels' few, large fronts) run FP32 tasks on
// 1024-thread workgroups: a lone task then has 16 waves of loads in flight on its CU.
constexpr int kVecWideTasks = 128;
template <typename S>
void launch_vec_q(const VecQTask<S>* d, int n, hipStream_t s) {
    if (n <= 0) return;
    if constexpr (sizeof(S) == 4) {
        if (n <= kVecWideTasks) {
            hipLaunchKernelGGL((k_vec_q<S, 1024>), dim3(n), dim3(1024), 0, s, d);
            return;
        }
    }
    hipLaunchKernelGGL((k_vec_q<S, WG>), dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_vec_chain(const VecChainTask<S>* d, const VecChainLink<S>* l, int n, hipStream_t s) {
    if (n <= 0) return;
    if constexpr (sizeof(S) == 4) {
        if (n <= kVecWideTasks) {
            hipLaunchKernelGGL((k_vec_chain<S, 1024>), dim3(n), dim3(1024), 0, s, d, l);
            return;
        }
    }
    hipLaunchKernelGGL((k_vec_chain<S, WG>), dim3(n), dim3(WG), 0, s, d, l);
}
template <typename S>
void launch_vec_chain_pipe(const VecChainTask<S>* d, const VecChainLink<S>* l, int n, int nrw, hipStream_t s) {
    if (n <= 0) return;
    constexpr bool F = sizeof(S) == 4;  // FP64 holds fewer items in flight (registers)
    if (nrw <= 1)
        hipLaunchKernelGGL((k_vec_chain_pipe<S, 1, F ? 3 : 2>), dim3(n), dim3(WG), 0, s, d, l);
    else
        hipLaunchKernelGGL((k_vec_chain_pipe<S, 2, F ? 2 : 1>), dim3(n), dim3(WG), 0, s, d, l);
}
template <typename S>
void launch_vec_chain_wave(const VecChainTask<S>* d, const VecChainLink<S>* l, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_vec_chain_wave<S>, dim3((n + 3) / 4), dim3(WG), 0, s, d, l, n);
}
template <typename S>
void launch_vec_spars(const VecSparsTask<S>* d, int n, int fwd, hipStream_t s) {
    if (n <= 0) return;
    if constexpr (sizeof(S) == 4) {
        if (n <= kVecWideTasks) {
            if (fwd) hipLaunchKernelGGL((k_vec_spars<S, 1024, true>), dim3(n), dim3(1024), 0, s, d);
            else hipLaunchKernelGGL((k_vec_spars<S, 1024, false>), dim3(n), dim3(1024), 0, s, d);
            return;
        }
    }
    if (fwd) hipLaunchKernelGGL((k_vec_spars<S, WG, true>), dim3(n), dim3(WG), 0, s, d);
    else hipLaunchKernelGGL((k_vec_spars<S, WG, false>), dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_vec_trsv(const VecTrsvTask<S>* d, const VecGemvLink<S>* l, int n, hipStream_t s) {
    if (n <= 0) return;
    if constexpr (sizeof(S) == 4) {
        if (n <= kVecWideTasks) {
            hipLaunchKernelGGL((k_vec_trsv<S, 1024>), dim3(n), dim3(1024), 0, s, d, l);
            return;
        }
    }
    hipLaunchKernelGGL((k_vec_trsv<S, WG>), dim3(n), dim3(WG), 0, s, d, l);
}
template <typename S>
void launch_vec_copy(const VecCopyTask<S>* d, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_vec_copy<S>, dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_convert_d2s(int n, const double* x, S* y, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_convert_d2s<S>, dim3(std::min((n + WG - 1) / WG, 4096)), dim3(WG), 0, s, n, x, y);
}
template <typename S>
void launch_convert_s2d(int n, const S* x, double* y, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_convert_s2d<S>, dim3(std::min((n + WG - 1) / WG, 4096)), dim3(WG), 0, s, n, x, y);
}

void launch_spmv(int n, const int* rowptr, const int* col, const double* val, const double* x, double* y,
                 hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_spmv_csr, dim3((n + WG - 1) / WG), dim3(WG), 0, s, n, rowptr, col, val, x, y);
}
void launch_dot(int n, const double* a, const double* b, double* partial, int nblocks, double* out, int sqrt_mode,
                hipStream_t s) {
    hipLaunchKernelGGL(k_dot_partial, dim3(nblocks), dim3(WG), 0, s, n, a, b, partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(WG), 0, s, nblocks, partial, out, sqrt_mode);
}
void launch_axpy_dev(int n, const double* h, const double* x, double* y, hipStream_t s) {
    hipLaunchKernelGGL(k_axpy_dev, dim3(std::min((n + WG - 1) / WG, 2048)), dim3(WG), 0, s, n, h, x, y);
}
void launch_scal_dev(int n, const double* h, const double* x, double* y, hipStream_t s) {
    hipLaunchKernelGGL(k_scal_dev, dim3(std::min((n + WG - 1) / WG, 2048)), dim3(WG), 0, s, n, h, x, y);
}
void launch_axpby(int n, double a, const double* x, double b, double* y, hipStream_t s) {
    hipLaunchKernelGGL(k_axpby, dim3(std::min((n + WG - 1) / WG, 2048)), dim3(WG), 0, s, n, a, x, b, y);
}

#define SPAQR_INSTANTIATE(S)                                                                          \
    template void launch_geqrt<S>(const GeqrtTask<S>*, int, hipStream_t);                            \
    template void launch_larfb<S>(const LarfbTask<S>*, int, hipStream_t);                            \
    template void launch_panel_qr<S>(const PanelTask<S>*, int, bool, hipStream_t);                    \
    template void launch_panel_apply<S>(const PanelApplyTask<S>*, int, bool, hipStream_t);           \
    template void launch_panel_step<S>(const PanelStepTask<S>*, int, hipStream_t);                    \
    template void launch_tpqrt<S>(const TpqrtTask<S>*, const TpqrtLink<S>*, int, hipStream_t);       \
    template void launch_tpqrt_wave<S>(const TpqrtTask<S>*, const TpqrtLink<S>*, int, hipStream_t);       \
    template void launch_tpqrt_coop<S>(const TpqrtCoopTask<S>*, const int*, int, const TpqrtLink<S>*, int*, hipStream_t); \
    template void launch_geqrt_wave<S>(const GeqrtTask<S>*, int, int, hipStream_t);                    \
    template void launch_tpmqrt_wave<S>(const TpmqrtTask<S>*, const TpmqrtLink<S>*, int, int, hipStream_t); \
    template int tpqrt_coop_max_blocks<S>();                                                          \
    template void launch_tpmqrt<S>(const TpmqrtTask<S>*, const TpmqrtLink<S>*, int, hipStream_t);    \
    template void launch_copy<S>(const CopyTask<S>*, int, hipStream_t, int);                         \
    template void launch_geqp3<S>(const Geqp3Task<S>*, int, double, double, hipStream_t);                    \
    template void launch_spars_apply<S>(const SparsApplyTask<S>*, int, hipStream_t);                 \
    template void launch_larfb_small<S>(const LarfbTask<S>*, int, int, bool, hipStream_t);            \
    template void launch_spars_small<S>(const SparsApplyTask<S>*, int, int, hipStream_t);             \
    template void launch_fill<S>(S*, size_t, S, hipStream_t);                                        \
    template void launch_vec_q<S>(const VecQTask<S>*, int, hipStream_t);                             \
    template void launch_vec_chain<S>(const VecChainTask<S>*, const VecChainLink<S>*, int, hipStream_t); \
    template void launch_vec_chain_wave<S>(const VecChainTask<S>*, const VecChainLink<S>*, int, hipStream_t); \
    template void launch_vec_chain_pipe<S>(const VecChainTask<S>*, const VecChainLink<S>*, int, int, hipStream_t); \
    template void launch_vec_spars<S>(const VecSparsTask<S>*, int, int, hipStream_t);                \
    template void launch_vec_trsv<S>(const VecTrsvTask<S>*, const VecGemvLink<S>*, int, hipStream_t); \
    template void launch_vec_copy<S>(const VecCopyTask<S>*, int, hipStream_t);                       \
    template void launch_convert_d2s<S>(int, const double*, S*, hipStream_t);                        \
    template void launch_convert_s2d<S>(int, const S*, double*, hipStream_t);

SPAQR_INSTANTIATE(double)
SPAQR_INSTANTIATE(float)
template void launch_fill<int>(int*, size_t, int, hipStream_t);

}  // namespace spaqr

// =======================================================================================
// Cooperative column-pivoted QR.  One interface's C (k x w) is split into column slices held
// in the LDS of P workgroups (co-resident: cooperative launch).  Per pivot step every block
// publishes its best candidate column (local maximum norm, lowest LAPACK position inside
// the tie band) to HBM, the P blocks of the interface meet at one barrier, and each block
// then picks the global pivot and forms the Householder reflector itself from the published
// column (identical arithmetic in every block), so only one barrier per step is needed and
// the trailing update never leaves LDS.  Column swaps are bookkeeping on LAPACK positions.
// Barrier waits are bounded: a wait that exceeds the spin limit raises *err and returns.
// =======================================================================================
namespace spaqr {

constexpr long kBarrierSpinLimit = 1L << 24;

__device__ __forceinline__ void task_barrier(int* bar, int P, int& passed, int* err) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(bar, 1);
        passed++;
        const int target = P * passed;
        long spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > kBarrierSpinLimit) { atomicOr(err, 1); break; }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
    } else {
        passed++;
    }
    __syncthreads();
}

template <typename S>
__device__ __forceinline__ S ld_coherent(const S* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename S>
__device__ __forceinline__ void st_coherent(S* p, S v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// dst[0:n] = src[0:n] (coherent reads), eight loads in flight per thread before any wait: a
// plain loop waits on every load, and a published column of a few thousand rows would cost
// that many serial memory round trips per pivot step.
template <typename S>
__device__ __forceinline__ void coherent_read(const S* src, S* dst, int n) {
    for (int r0 = threadIdx.x; r0 < n; r0 += 8 * WG) {
        S v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = ld_coherent(src + min(r0 + u * WG, n - 1));
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (r0 + u * WG < n) dst[r0 + u * WG] = v[u];
    }
}
// dst[0:n] = src[0:n] (coherent writes), the source read eight rows at a time.
template <typename S>
__device__ __forceinline__ void coherent_write(const S* src, S* dst, int n) {
    for (int r0 = threadIdx.x; r0 < n; r0 += 8 * WG) {
        S v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = src[min(r0 + u * WG, n - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (r0 + u * WG < n) st_coherent(dst + r0 + u * WG, v[u]);
    }
}

// record layout inside a slot (k + 4 values): M_b, v_cand, pos_cand, unused, column[k]
// HBM = false: the workgroup's column slice lives in LDS for the whole factorization.
// HBM = true: for interfaces whose C cannot be spread over the co-resident workgroups' LDS, each
// workgroup updates its columns of the gathered C in place in HBM (only norms, positions and
// the reflector stay in LDS); the row loops are unrolled so that loads overlap.
template <typename S, bool HBM>
__global__ __launch_bounds__(WG) void k_geqp3_coop(const Geqp3CoopTask<S>* __restrict__ tasks,
                                                   const int* __restrict__ blk2task, int* err, double tol,
                                                   double tie_rel) {
    extern __shared__ __align__(16) unsigned char dyn_smem[];
    __shared__ S red_v[4];
    __shared__ int red_i[4];
    __shared__ S sh_s[4];
    __shared__ int sh_i[4];
    __shared__ PanelSmem<S> psm;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const Geqp3CoopTask<S> t = tasks[blk2task[blockIdx.x]];
    const int k = t.k, w = t.w, P = t.P, cpb = t.cpb;
    const int p = blockIdx.x - t.b0;
    const int q0 = p * cpb;
    const int ncol = max(0, min(w, q0 + cpb) - q0);
    const int rec = k + 4;
    S* slice = HBM ? const_cast<S*>(t.C) + (size_t)q0 * k : reinterpret_cast<S*>(dyn_smem);  // k x cpb, ld k
    S* vn1 = HBM ? reinterpret_cast<S*>(dyn_smem) : slice + (size_t)k * cpb;
    S* vn2 = vn1 + cpb;
    S* vbuf = vn2 + cpb;                          // k: the pivot column, then the reflector
    int* pos = reinterpret_cast<int*>(vbuf + k);  // LAPACK position of each owned column; -1 once pivoted
    const S tol3z = sqrt(S(sizeof(S) == 8 ? 1.1102230246251565e-16 : 5.9604644775390625e-08));
    int passed = 0;

    if (!HBM)
        for (int e = tid; e < k * ncol; e += WG) slice[e] = t.C[(size_t)q0 * k + e];
    __syncthreads();
    // 16 lanes (one DPP row) per column
    const int grp = tid >> 4, gl = tid & 15;
    for (int c0 = 0; c0 < ncol; c0 += WG / 16) {
        const int c = c0 + grp;
        S s = S(0);
        if (c < ncol)
            for (int r = gl; r < k; r += 16) { S v = slice[r + (size_t)c * k]; s += v * v; }
        s = row16_sum(s);
        if (c < ncol && gl == 0) { vn1[c] = sqrt(s); vn2[c] = vn1[c]; pos[c] = q0 + c; }
    }
    __syncthreads();

    int rank = t.kmax;
    S r00 = S(0);
    for (int j = 0; j < t.kmax; ++j) {
        // ---- local maximum and the lowest-position candidate inside the local band
        S mb = S(-1);
        for (int c = tid; c < ncol; c += WG)
            if (pos[c] >= 0 && vn1[c] > mb) mb = vn1[c];
        mb = wave_max(mb);
        if (lane == 0) red_v[wv] = mb;
        __syncthreads();
        mb = fmax(fmax(red_v[0], red_v[1]), fmax(red_v[2], red_v[3]));
        __syncthreads();
        S thr = (tie_rel > 0.0) ? mb * (S)(1.0 - tie_rel) : mb;
        int cand = INT_MAX, cc = -1;
        for (int c = tid; c < ncol; c += WG)
            if (pos[c] >= 0 && vn1[c] >= thr && pos[c] < cand) { cand = pos[c]; cc = c; }
        {
            const int wmin = wave_min(cand);
            const int src = __ffsll((long long)__ballot(cand == wmin)) - 1;
            const int wcc = lane_read(cc, src);
            if (lane == 0) { red_i[wv] = wmin; sh_i[wv] = wcc; }
        }
        __syncthreads();
        if (tid == 0) {
            int b = 0;
            for (int q = 1; q < 4; ++q)
                if (red_i[q] < red_i[b]) b = q;
            sh_i[3] = sh_i[b];  // candidate column (local index) or -1
        }
        __syncthreads();
        int ccol = sh_i[3];
        int pv_col = -1;  // local column of the pivot when this block owns it
        if (P == 1) {
            if (ccol < 0) { rank = j; break; }  // no finite candidate left (NaN input)
            pv_col = ccol;
            for (int r = tid; r < k; r += WG) vbuf[r] = slice[r + (size_t)ccol * k];
            if (tid == 0) sh_i[0] = pos[ccol];
            __syncthreads();
        } else {
            // publish: slot = parity of the barriers passed so far
            S* slot = t.xch + (size_t)(passed & 1) * P * rec;
            S* my = slot + (size_t)p * rec;
            if (tid == 0) {
                st_coherent(my + 0, mb);
                st_coherent(my + 1, ccol >= 0 ? vn1[ccol] : S(-1));
                st_coherent(my + 2, ccol >= 0 ? (S)pos[ccol] : S(-1));
            }
            if (ccol >= 0) coherent_write(slice + (size_t)ccol * k, my + 4, k);
            task_barrier(t.bar, P, passed, err);
            // global maximum, band, ambiguity; the records of q = tid and tid + WG (P <= 2 WG, the
            // co-resident limit) are read once, all three fields in one round trip
            S mrec[2], vrec[2];
            int prec[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const S* rq = slot + (size_t)min(tid + h * WG, P - 1) * rec;
                mrec[h] = ld_coherent(rq);
                vrec[h] = ld_coherent(rq + 1);
                prec[h] = (int)ld_coherent(rq + 2);
            }
            S M = S(-1);
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (tid + h * WG < P) M = fmax(M, mrec[h]);
            for (int q = tid + 2 * WG; q < P; q += WG) M = fmax(M, ld_coherent(slot + (size_t)q * rec));
            M = wave_max(M);
            if (lane == 0) red_v[wv] = M;
            __syncthreads();
            M = fmax(fmax(red_v[0], red_v[1]), fmax(red_v[2], red_v[3]));
            __syncthreads();
            const S gthr = (tie_rel > 0.0) ? M * (S)(1.0 - tie_rel) : M;
            int best = INT_MAX, bq = -1, amb = 0;
            for (int q = tid; q < P; q += WG) {
                S mq, vq;
                int pq;
                if (q < tid + 2 * WG) {
                    const int h = (q - tid) / WG;
                    mq = h ? mrec[1] : mrec[0];
                    vq = h ? vrec[1] : vrec[0];
                    pq = h ? prec[1] : prec[0];
                } else {
                    const S* rq = slot + (size_t)q * rec;
                    mq = ld_coherent(rq);
                    vq = ld_coherent(rq + 1);
                    pq = (int)ld_coherent(rq + 2);
                }
                if (vq >= gthr && pq >= 0 && pq < best) { best = pq; bq = q; }
                if (mq >= gthr && !(vq >= gthr)) amb = 1;
            }
            {
                const int wb = wave_min(best);
                const int src = __ffsll((long long)__ballot(best == wb)) - 1;
                const int wq = lane_read(bq, src);
                const int wamb = __any(amb) ? 1 : 0;
                if (lane == 0) { red_i[wv] = wb; sh_i[wv] = wq; red_v[wv] = (S)wamb; }
            }
            __syncthreads();
            if (tid == 0) {
                int b = 0;
                for (int q = 1; q < 4; ++q)
                    if (red_i[q] < red_i[b]) b = q;
                sh_i[2] = sh_i[b];
                sh_i[3] = (red_v[0] + red_v[1] + red_v[2] + red_v[3]) > S(0) ? 1 : 0;
            }
            __syncthreads();
            int winq = sh_i[2];
            const S* wslot = slot;
            if (sh_i[3]) {
                // a block's band candidate fell outside the global band: second round with gthr
                __syncthreads();
                int c2 = INT_MAX, cc2 = -1;
                for (int c = tid; c < ncol; c += WG)
                    if (pos[c] >= 0 && vn1[c] >= gthr && pos[c] < c2) { c2 = pos[c]; cc2 = c; }
                {
                    const int wmin = wave_min(c2);
                    const int src = __ffsll((long long)__ballot(c2 == wmin)) - 1;
                    const int wcc = lane_read(cc2, src);
                    if (lane == 0) { red_i[wv] = wmin; sh_i[wv] = wcc; }
                }
                __syncthreads();
                if (tid == 0) {
                    int b = 0;
                    for (int q = 1; q < 4; ++q)
                        if (red_i[q] < red_i[b]) b = q;
                    sh_i[3] = sh_i[b];
                }
                __syncthreads();
                ccol = sh_i[3];
                S* slot2 = t.xch + (size_t)(passed & 1) * P * rec;
                S* my2 = slot2 + (size_t)p * rec;
                if (tid == 0) {
                    st_coherent(my2 + 1, ccol >= 0 ? vn1[ccol] : S(-1));
                    st_coherent(my2 + 2, ccol >= 0 ? (S)pos[ccol] : S(-1));
                }
                if (ccol >= 0) coherent_write(slice + (size_t)ccol * k, my2 + 4, k);
                task_barrier(t.bar, P, passed, err);
                int b2 = INT_MAX, q2 = -1;
                for (int q = tid; q < P; q += WG) {
                    int pq = (int)ld_coherent(slot2 + (size_t)q * rec + 2);
                    if (pq >= 0 && pq < b2) { b2 = pq; q2 = q; }
                }
                {
                    const int wb = wave_min(b2);
                    const int src = __ffsll((long long)__ballot(b2 == wb)) - 1;
                    const int wq = lane_read(q2, src);
                    if (lane == 0) { red_i[wv] = wb; sh_i[wv] = wq; }
                }
                __syncthreads();
                if (tid == 0) {
                    int b = 0;
                    for (int q = 1; q < 4; ++q)
                        if (red_i[q] < red_i[b]) b = q;
                    sh_i[2] = sh_i[b];
                }
                __syncthreads();
                winq = sh_i[2];
                wslot = slot2;
            }
            if (winq < 0) { rank = j; break; }  // no finite candidate in any block (NaN input)
            const S* wr = wslot + (size_t)winq * rec;
            coherent_read(wr + 4, vbuf, k);
            if (tid == 0) sh_i[0] = (int)ld_coherent(wr + 2);
            if (winq == p) pv_col = ccol;
            __syncthreads();
        }
        const int pvpos = sh_i[0];
        // ---- reflector from the pivot column (every block, identical arithmetic)
        S ss = S(0);
        for (int r = j + 1 + tid; r < k; r += WG) ss += vbuf[r] * vbuf[r];
        ss = wave_sum(ss);
        if (lane == 0) red_v[wv] = ss;
        __syncthreads();
        if (tid == 0) {
            const S s4 = red_v[0] + red_v[1] + red_v[2] + red_v[3];
            const S alpha = vbuf[j];
            S tau, scal, beta;
            if (s4 == S(0)) { tau = S(0); scal = S(1); beta = alpha; }
            else {
                beta = -householder_norm_signed(alpha, s4);
                tau = (beta - alpha) / beta;
                scal = S(1) / (alpha - beta);
            }
            sh_s[0] = tau;
            sh_s[1] = scal;
            sh_s[2] = beta;
        }
        __syncthreads();
        const S tau = sh_s[0], scal = sh_s[1], beta = sh_s[2];
        if (j == 0) r00 = fabs(beta);
        else if (tol > 0.0 && !((double)(fabs(beta) / r00) >= tol)) { rank = j; break; }
        for (int r = j + 1 + tid; r < k; r += WG) vbuf[r] *= scal;
        __syncthreads();
        // ---- the owner stores the finished column; positions are swapped by bookkeeping
        if (pv_col >= 0) {
            for (int r = tid; r < k; r += WG) t.Vout[r + (size_t)j * k] = (r == j) ? beta : vbuf[r];
            if (tid == 0) {
                t.tau[j] = tau;
                if (t.jpvt_out) t.jpvt_out[j] = q0 + pv_col;
                pos[pv_col] = -1;
            }
        }
        for (int c = tid; c < ncol; c += WG)
            if (pos[c] == j && j != pvpos) pos[c] = pvpos;
        __syncthreads();
        // ---- apply H_j to the owned unpivoted columns, downdate their norms
        for (int c0 = 0; c0 < ncol; c0 += WG / 16) {
            const int c = c0 + grp;
            const bool act = (c < ncol) && pos[c] >= 0;  // uniform over the 16-lane row
            S* cl = slice + (size_t)min(c, ncol - 1) * k;
            S dot = S(0);
            if (act) {
#pragma unroll 8
                for (int r = j + 1 + gl; r < k; r += 16) dot += vbuf[r] * cl[r];
            }
            dot = row16_sum(dot);
            const S cjl = act ? cl[j] : S(0);
            const S wl = (cjl + dot) * tau;
            const S rjl = cjl - wl;
            S nrm = S(0);
            if (act) {
#pragma unroll 8
                for (int r = j + 1 + gl; r < k; r += 16) {
                    S v = cl[r] - vbuf[r] * wl;
                    cl[r] = v;
                    nrm += v * v;
                }
            }
            nrm = row16_sum(nrm);
            if (act && gl == 0) {
                cl[j] = rjl;
                S v1 = vn1[c];
                if (v1 != S(0)) {
                    S temp = fabs(rjl) / v1;
                    temp = S(1) - temp * temp;
                    temp = temp > S(0) ? temp : S(0);
                    S q = v1 / vn2[c];
                    S temp2 = temp * q * q;
                    if (temp2 <= tol3z) {
                        S nv = (j + 1 < k) ? sqrt(nrm) : S(0);
                        vn1[c] = nv;
                        vn2[c] = nv;
                    } else {
                        vn1[c] = v1 * sqrt(temp);
                    }
                }
            }
        }
        __syncthreads();
    }
    // ---- T of the kept reflectors, by the first block once every owner has stored its columns
    if (P > 1) task_barrier(t.bar, P, passed, err);
    if (p == 0) {
        if (tid == 0) *t.rank_out = rank;
        if (rank > 0) wg_larft_blocked<S>(k, rank, t.Vout, k, t.tau, t.T, psm.mem);
    }
}

template <typename S>
int geqp3_coop_max_blocks(size_t lds_bytes, bool hbm) {
    int per_cu = 0, dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    hipError_t e = hbm ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_geqp3_coop<S, true>, WG, lds_bytes)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_geqp3_coop<S, false>, WG, lds_bytes);
    if (e != hipSuccess) return 0;
    return per_cu * ncu;
}

template <typename S>
hipError_t launch_geqp3_coop(const Geqp3CoopTask<S>* d_tasks, const int* d_blk2task, int* d_err, int nblocks,
                             double tol, double tie_rel, size_t lds_bytes, bool hbm, hipStream_t s, bool cooperative) {
    if (nblocks <= 0) return hipSuccess;
    void* args[] = {(void*)&d_tasks, (void*)&d_blk2task, (void*)&d_err, (void*)&tol, (void*)&tie_rel};
    const void* fn = hbm ? (const void*)k_geqp3_coop<S, true> : (const void*)k_geqp3_coop<S, false>;
    if (!cooperative) return hipLaunchKernel(fn, dim3(nblocks), dim3(WG), args, (size_t)lds_bytes, s);
    hipError_t e = hipLaunchCooperativeKernel(fn, dim3(nblocks), dim3(WG), args, (unsigned)lds_bytes, s);
    if (e == hipErrorIllegalState) {
        // The device's cooperative queue belongs to one process at a time: while another
        // process holds it, every cooperative launch here fails with this code.  The grid is
        // within the co-resident limit and the blocks meet at their own barrier, whose wait has
        // a spin limit, so an ordinary launch computes the same result (and a block that is
        // never scheduled ends in an error, not a hang).
        (void)hipGetLastError();
        e = hipLaunchKernel(fn, dim3(nblocks), dim3(WG), args, (size_t)lds_bytes, s);
    }
    return e;
}

template <typename S>
size_t geqp3_coop_lds_budget(int blocks_per_cu) {
    int dev = 0, lds_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess)
        return 0;
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, (const void*)k_geqp3_coop<S, false>) != hipSuccess) return 0;
    const size_t per = (size_t)lds_cu / (size_t)std::max(blocks_per_cu, 1);
    if (per <= fa.sharedSizeBytes + 1024) return 0;
    return ((per - fa.sharedSizeBytes) / 1024) * 1024;  // dynamic bytes, whole KiB
}

template size_t geqp3_coop_lds_budget<double>(int);
template size_t geqp3_coop_lds_budget<float>(int);
template int geqp3_coop_max_blocks<double>(size_t, bool);
template int geqp3_coop_max_blocks<float>(size_t, bool);
template hipError_t launch_geqp3_coop<double>(const Geqp3CoopTask<double>*, const int*, int*, int, double, double,
                                              size_t, bool, hipStream_t, bool);
template hipError_t launch_geqp3_coop<float>(const Geqp3CoopTask<float>*, const int*, int*, int, double, double,
                                             size_t, bool, hipStream_t, bool);

}  // namespace spaqr

// =======================================================================================
// Least squares (rectangular systems): the row bookkeeping of the row reassignment and the
// transposed backward pass of the preconditioner (solve_nrml).  None of these is on the
// square headline's path; each is one workgroup (or one thread) per item, bound by latency.
// =======================================================================================
namespace spaqr {

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// out[i] = || A(i, 0:cols) || in FP64: threads over rows (consecutive rows are consecutive in
// a column, so each column's loads coalesce).
template <typename S>
__global__ __launch_bounds__(WG) void k_row_norms(const RowNormTask<S>* __restrict__ tasks) {
    const RowNormTask<S> t = tasks[blockIdx.x];
    for (int i = threadIdx.x; i < t.rows; i += WG) {
        double s = 0.0;
        for (int j = 0; j < t.cols; ++j) {
            const double a = (double)t.A[i + (size_t)j * t.lda];
            s += a * a;
        }
        t.out[i] = sqrt(s);
    }
}

// dst = [old; the gathered rows]: the grown block of an edge after the row reassignment.
template <typename S>
__global__ __launch_bounds__(WG) void k_row_gather(const RowGatherTask<S>* __restrict__ tasks,
                                                   const RowSrc<S>* __restrict__ src) {
    const RowGatherTask<S> t = tasks[blockIdx.x];
    const int rn = t.rows_old + t.nextra;
    const long total = (long)rn * t.cols;
    for (long e = threadIdx.x; e < total; e += WG) {
        const int i = (int)(e % rn), j = (int)(e / rn);
        S v;
        if (i < t.rows_old) {
            v = t.old[i + (size_t)j * t.lds];
        } else {
            const RowSrc<S> r = src[t.src0 + i - t.rows_old];
            v = r.p ? r.p[(size_t)j * r.ld] : S(0);
        }
        t.dst[i + (size_t)j * t.ldd] = v;
    }
}

// dst[di[q]] = si[q] < 0 ? 0 : src[si[q]]
template <typename S>
__global__ __launch_bounds__(WG) void k_vec_gather(const S* __restrict__ src, S* __restrict__ dst,
                                                   const long long* __restrict__ si, const long long* __restrict__ di,
                                                   int n) {
    const int q = blockIdx.x * WG + threadIdx.x;
    if (q >= n) return;
    const long long s = si[q];
    dst[di[q]] = s < 0 ? S(0) : src[s];
}

// x <- R^{-T} x for R upper triangular (k x k, ldr), x in global memory, 32-row blocks from the
// top: the waves first subtract the solved entries' contributions from the block (one wave per
// column of R, lanes along the column), then wave 0 solves the block's lower triangle R^T by
// lane reads.
template <typename S, int NT>
__device__ void wg_trsv_upper_t(int k, const S* __restrict__ R, int ldr, S* x) {
    constexpr int NW = NT / 64;
    __shared__ double part[32];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int p0 = 0; p0 < k; p0 += NB) {
        const int kb = min(NB, k - p0);
        for (int jj = wave; jj < kb; jj += NW) {
            const S* __restrict__ Rc = R + (size_t)(p0 + jj) * ldr;
            double s = 0.0;
            for (int i = lane; i < p0; i += 64) s += (double)Rc[i] * (double)x[i];
            s = wave_sum(s);
            if (lane == 0) part[jj] = s;
        }
        __syncthreads();
        if (wave == 0) {
            const int l = min(lane, kb - 1);
            S b = (lane < kb) ? x[p0 + l] - (S)part[l] : S(0);
            for (int j = 0; j < kb; ++j) {
                const S xj = lane_read(b, j) / R[(size_t)(p0 + j) * (ldr + 1)];
                if (lane == j) b = xj;
                else if (lane > j && lane < kb) b -= R[(p0 + j) + (size_t)(p0 + lane) * ldr] * xj;
            }
            if (lane < kb) x[p0 + lane] = b;
        }
        __syncthreads();
    }
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void k_vec_trsv_t(const VecTrsvTask<S>* __restrict__ tasks) {
    const VecTrsvTask<S> t = tasks[blockIdx.x];
    wg_trsv_upper_t<S, NT>(t.k, t.R, t.ldr, t.x);
}

// x[0:cols] -= sum over links of A(0:rows, 0:cols)^T y: one wave per column of x.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void k_vec_gemvt(const VecGemvtTask<S>* __restrict__ tasks,
                                                  const VecGemvtLink<S>* __restrict__ links) {
    constexpr int NW = NT / 64;
    const VecGemvtTask<S> t = tasks[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = wave; j < t.cols; j += NW) {
        double s = 0.0;
        for (int li = 0; li < t.nlinks; ++li) {
            const VecGemvtLink<S> L = links[t.link0 + li];
            const S* __restrict__ Aj = L.A + (size_t)j * L.lda;
            for (int i = lane; i < L.rows; i += 64) s += (double)Aj[i] * (double)L.y[i];
        }
        s = wave_sum(s);
        if (lane == 0) t.x[j] -= (S)s;
    }
}

// The transpose of the sparsified interface's backward step: x[0:k] <- R_s^{-T} x[0:k] (R_s the
// upper k x k part of Qs, ld ldq), then x[0:k] <- Qsp^T x[0:k].
template <typename S, int NT>
__global__ __launch_bounds__(NT) void k_vec_spars_t(const VecSparsTask<S>* __restrict__ tasks) {
    __shared__ VecSmem<S, NT> sm;
    const VecSparsTask<S> t = tasks[blockIdx.x];
    wg_trsv_upper_t<S, NT>(t.k, t.Qs, t.ldq, t.x);
    __syncthreads();
    if (t.r > 0) wg_vec_q<S, NT>(t.k, t.r, t.Qsp, t.k, t.Tsp, t.x, true, sm);
}

template <typename S>
void launch_row_norms(const RowNormTask<S>* d, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_row_norms<S>, dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_row_gather(const RowGatherTask<S>* d, const RowSrc<S>* src, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_row_gather<S>, dim3(n), dim3(WG), 0, s, d, src);
}
template <typename S>
void launch_vec_gather(const S* src, S* dst, const long long* si, const long long* di, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_vec_gather<S>, dim3((n + WG - 1) / WG), dim3(WG), 0, s, src, dst, si, di, n);
}
template <typename S>
void launch_vec_trsv_t(const VecTrsvTask<S>* d, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL((k_vec_trsv_t<S, WG>), dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_vec_gemvt(const VecGemvtTask<S>* d, const VecGemvtLink<S>* l, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL((k_vec_gemvt<S, WG>), dim3(n), dim3(WG), 0, s, d, l);
}
template <typename S>
void launch_vec_spars_t(const VecSparsTask<S>* d, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL((k_vec_spars_t<S, WG>), dim3(n), dim3(WG), 0, s, d);
}

#define SPAQR_INSTANTIATE_LSQ(S)                                                                     \
    template void launch_row_norms<S>(const RowNormTask<S>*, int, hipStream_t);                      \
    template void launch_row_gather<S>(const RowGatherTask<S>*, const RowSrc<S>*, int, hipStream_t); \
    template void launch_vec_gather<S>(const S*, S*, const long long*, const long long*, int, hipStream_t); \
    template void launch_vec_trsv_t<S>(const VecTrsvTask<S>*, int, hipStream_t);                    \
    template void launch_vec_gemvt<S>(const VecGemvtTask<S>*, const VecGemvtLink<S>*, int, hipStream_t); \
    template void launch_vec_spars_t<S>(const VecSparsTask<S>*, int, hipStream_t);
SPAQR_INSTANTIATE_LSQ(double)
SPAQR_INSTANTIATE_LSQ(float)

}  // namespace spaqr
