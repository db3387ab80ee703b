// gfx950 block-reflector applies with the updated block resident in MFMA accumulators
// (FP32).  Built with the MFMA accumulators in VGPRs (-amdgpu-mfma-vgpr-form): a resident
// tile is both an accumulator (C -= V W) and a B operand (W = V^T C), and with AGPR
// accumulators the compiler keeps a second, VGPR copy of every tile.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "launch.hpp"
#include "mfma.hpp"

namespace spaqr {

// ---------------------------------------------------------------------------------------
// FP32 block-reflector applies with C resident in MFMA accumulators.  Wave w owns the
// 32-column slab 32 w .. 32 w + 31 of C as MT 32 x 32 accumulator tiles (entry r of lane l
// in tile t is C(32 t + mfma_row(l, r), 32 w + l % 32)) for all panels of the reflector, so
// per panel only V_p (staged once in LDS, shared by the waves) and T_p move, and C is read
// and written once per task instead of twice per panel.  Each product's contraction index is
// permuted to the accumulator layout, so the three products chain register to register: in
// step r the lanes of half h = l / 32 contribute index mfma_row(32 h, r), and the A operand
// is read from LDS at that index:
//   W  = V_p^T C  : A(i, .) = V_p(row 32 t + mfma_row(l, r), i),  B = C tile entry r
//   W2 = T_p^T W  : A(i, .) = T_p(mfma_row(l, r), i),             B = W entry r
//   C -= V_p W2   : A(i, .) = V_p(row 32 t + i, mfma_row(l, r)),  B = W2 entry r
template <int MT>
struct ResSmem {
    float Vs[32 * MT * 33];  // V_p rows p0 .. p0 + 32 MT - 1, row-major with 33-float rows
    float Ts[32 * 33];       // T_p(q, i) at q * 33 + i
};

// Stage one panel of V -- Vp points at its first row and column, `rows` rows, kb <= 32
// columns -- into sm.Vs (UNIT: unit lower, ones on the diagonal and zeros above; zeros past
// the rows and columns) and its T block (T, ld NB) into sm.Ts (zeros outside the upper kb x kb
// triangle).  NT threads.
template <int MT, bool UNIT, int NT>
__device__ __forceinline__ void res_stage(const float* __restrict__ Vp, int ldv, const float* __restrict__ T, int rows,
                                          int kb, ResSmem<MT>& sm) {
    // RS row slots (consecutive threads, consecutive rows: coalesced), CG column groups, RPT
    // rows per thread; columns go 8 at a time so few loads are live next to the accumulators
    constexpr int RS = 32 * MT < NT ? 32 * MT : NT, CG = NT / RS, RPT = 32 * MT / RS, CPG = 32 / CG;
    constexpr int JG = RPT >= 8 ? 1 : 8 / RPT;  // columns per group: 8 loads in flight per thread
    const int tid = threadIdx.x, lr0 = tid % RS, cg = tid / RS;
#pragma unroll
    for (int j0 = 0; j0 < CPG; j0 += JG) {
        float v[JG][RPT];
#pragma unroll
        for (int jj = 0; jj < JG; ++jj) {
            const int j = cg * CPG + j0 + jj;
            const float* __restrict__ col = Vp + (size_t)min(j, kb - 1) * ldv;
#pragma unroll
            for (int q = 0; q < RPT; ++q) v[jj][q] = col[min(lr0 + q * RS, rows - 1)];
        }
#pragma unroll
        for (int jj = 0; jj < JG; ++jj) {
            const int j = cg * CPG + j0 + jj;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int lr = lr0 + q * RS;
                float x = (lr < rows && j < kb) ? v[jj][q] : 0.f;
                if (UNIT) x = (lr > j) ? x : ((lr == j && j < kb) ? 1.f : 0.f);
                sm.Vs[lr * 33 + j] = x;
            }
        }
    }
    for (int e = tid; e < 32 * 32; e += NT) {
        const int q = e & 31, i = e >> 5;
        const float tv = T[min(q, kb - 1) + (size_t)min(i, kb - 1) * NB];
        sm.Ts[q * 33 + i] = (q <= i && i < kb) ? tv : 0.f;
    }
}

// Resident tiles from / to one column of C, a tile at a time: entry `row` of this lane's
// column is base[row] (TRC: base[row * ld], C stored transposed).  Full tiles use fixed
// offsets; the partial one is clamped (loads) or masked (stores) at m rows.
template <int MT, bool TRC>
__device__ __forceinline__ void res_load(floatx16 (&acc)[MT], const float* __restrict__ base, int ld, int m, bool ok) {
    int l = threadIdx.x & 63;
    asm volatile("" : "+v"(l));  // opaque per call: keeps the per-lane row offsets out of callers' loops
    const size_t rs = TRC ? (size_t)ld : 1;
    const float* __restrict__ bh = base + (size_t)(4 * (l >> 5)) * rs;
#pragma unroll
    for (int tt = 0; tt < MT; ++tt) {
        if (32 * tt + 32 <= m) {
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[tt][r] = bh[(size_t)(32 * tt + (r >> 2) * 8 + (r & 3)) * rs];
        } else {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 32 * tt + mfma_row(l, r);
                const float v = base[(size_t)min(row, m - 1) * rs];
                acc[tt][r] = (row < m) ? v : 0.f;
            }
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[tt][r] = ok ? acc[tt][r] : 0.f;
        __builtin_amdgcn_sched_barrier(0);
    }
}
template <int MT, bool TRC>
__device__ __forceinline__ void res_store(const floatx16 (&acc)[MT], float* __restrict__ base, int ld, int m) {
    int l = threadIdx.x & 63;
    asm volatile("" : "+v"(l));
    const size_t rs = TRC ? (size_t)ld : 1;
    float* __restrict__ bh = base + (size_t)(4 * (l >> 5)) * rs;
#pragma unroll
    for (int tt = 0; tt < MT; ++tt) {
        if (32 * tt + 32 <= m) {
#pragma unroll
            for (int r = 0; r < 16; ++r) bh[(size_t)(32 * tt + (r >> 2) * 8 + (r & 3)) * rs] = acc[tt][r];
        } else if (32 * tt < m) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 32 * tt + mfma_row(l, r);
                if (row < m) base[(size_t)row * rs] = acc[tt][r];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// One panel on this wave's resident tiles: acc[tt] holds tile tb + tt of the block (rows
// 32 (tb + tt) ..); tiles t0 .. ntiles-1 take part, and Vs row 0 is tile t0's row 0.  Wtop, if
// given, holds the pentagonal top rows A (added to W = A + V_p^T C first) and receives A - W2.
// XCH: two waves hold the same columns in two row ranges; each adds the other's partial W
// through LDS (wx_mine, wx_other: 16 x 64 floats each), so both hold the same W -- every wave
// of the workgroup must make the call (it has a barrier); `active` waves do the arithmetic.
template <int MT, bool XCH>
__device__ __forceinline__ void res_panel(floatx16 (&acc)[MT], int tb, int t0, int ntiles, const float* Vs,
                                          const float* Ts, floatx16* Wtop, bool active, float* wx_mine,
                                          const float* wx_other) {
    const int l = threadIdx.x & 63, li = l & 31;
    // one accumulation chain: a 32x32x2 FP32 MFMA holds the SIMD for as long as its latency
    floatx16 W;
#pragma unroll
    for (int r = 0; r < 16; ++r) W[r] = Wtop ? (*Wtop)[r] : 0.f;
    if (active) {
#pragma unroll
        for (int tt = 0; tt < MT; ++tt) {
            const int g = tb + tt;
            if (g >= t0 && g < ntiles) {
                const float* vt = Vs + (size_t)(32 * (g - t0)) * 33;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const float a = vt[mfma_row(l, r) * 33 + li];
                    W = __builtin_amdgcn_mfma_f32_32x32x2f32(a, acc[tt][r], W, 0, 0, 0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);  // one tile's operands at a time (else all are hoisted)
        }
    }
    if constexpr (XCH) {
#pragma unroll
        for (int r = 0; r < 16; ++r) wx_mine[r * 64 + l] = W[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; ++r) W[r] += wx_other[r * 64 + l];  // a + b == b + a: both waves agree
    }
    if (!active) return;
    floatx16 W2;
#pragma unroll
    for (int r = 0; r < 16; ++r) W2[r] = 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const float a = Ts[mfma_row(l, r) * 33 + li];
        W2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, W[r], W2, 0, 0, 0);
    }
    if (Wtop) *Wtop -= W2;
#pragma unroll
    for (int tt = 0; tt < MT; ++tt) {
        const int g = tb + tt;
        if (g >= t0 && g < ntiles) {
            const float* vt = Vs + (size_t)(32 * (g - t0)) * 33;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float a = -vt[li * 33 + mfma_row(l, r)];
                acc[tt] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, W2[r], acc[tt], 0, 0, 0);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// C <- Q^T C (larfb 'L','T', Q from V m x k unit lower with blocked T), C m x n with
// m <= 32 MT and n <= 128, FP32.  Same transformation as k_larfb without a trsm.
template <int MT>
__global__ __launch_bounds__(WG) void k_larfb_res(const LarfbTask<float>* __restrict__ tasks) {
    __shared__ ResSmem<MT> sm;
    const LarfbTask<float> t = tasks[blockIdx.x];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, li = l & 31;
    const int m = t.m, n = t.n, k = t.k, ntiles = (m + 31) >> 5;
    const int col = 32 * w + li;
    const bool active = 32 * w < n;  // this wave has columns (uniform)
    float* __restrict__ Cc = t.C + (size_t)min(col, n - 1) * t.ldc;
    floatx16 acc[MT];
    res_load<MT, false>(acc, Cc, 0, m, col < n);
    for (int p0 = 0; p0 < k; p0 += 32) {
        const int kb = min(32, k - p0);
        __syncthreads();  // the previous panel's readers of sm are done
        res_stage<MT, true, WG>(t.V + p0 + (size_t)p0 * t.ldv, t.ldv, t.T + (size_t)p0 * NB, m - p0, kb, sm);
        __syncthreads();
        res_panel<MT, false>(acc, 0, p0 >> 5, ntiles, sm.Vs, sm.Ts, nullptr, active, nullptr, nullptr);
    }
    if (col < n) res_store<MT, false>(acc, Cc, 0, m);
}

// tpmqrt chain (PENT) of one column tile (ncols <= 64) on two waves: for each link j,
// [Atop; B_j] <- Q_j^T [Atop; B_j] with B_j's tile (rows_j <= 32 MT) resident in the
// accumulators for all of the link's panels; Atop's 32 rows of a panel are read into the W
// accumulator and written back as Atop - W2.  Same transformation as k_tpmqrt.
constexpr int kResTpmThreads = 128;
template <int MT>
__global__ __launch_bounds__(kResTpmThreads) void k_tpmqrt_res(const TpmqrtTask<float>* __restrict__ tasks,
                                                               const TpmqrtLink<float>* __restrict__ links) {
    __shared__ ResSmem<MT> sm;
    const TpmqrtTask<float> t = tasks[blockIdx.x];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, li = l & 31;
    const int k = t.k, nc = t.ncols, col = 32 * w + li, cc = min(col, nc - 1);
    const bool active = 32 * w < nc;  // this wave has columns (uniform)
    float* __restrict__ Acol = t.Atop + (size_t)cc * t.lda;
    floatx16 acc[MT];
    for (int j = 0; j < t.nlinks; ++j) {
        const TpmqrtLink<float> L = links[t.link0 + j];
        const int rows = L.rows, ntiles = (rows + 31) >> 5;
        if (rows <= 0) continue;  // an empty link's reflectors are the identity
        float* __restrict__ Cc = L.B + (size_t)(t.col0 + cc) * L.ldb;
        if (active) res_load<MT, false>(acc, Cc, 0, rows, col < nc);
        for (int p0 = 0; p0 < k; p0 += 32) {
            const int kb = min(32, k - p0);
            __syncthreads();  // the previous panel's readers of sm are done
            res_stage<MT, false, kResTpmThreads>(L.V + (size_t)p0 * L.ldv, L.ldv, L.T + (size_t)p0 * NB, rows, kb, sm);
            __syncthreads();
            if (active) {
                floatx16 top;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int q = mfma_row(l, r);
                    const float v = Acol[p0 + min(q, kb - 1)];
                    top[r] = (q < kb && col < nc) ? v : 0.f;
                }
                res_panel<MT, false>(acc, 0, 0, ntiles, sm.Vs, sm.Ts, &top, true, nullptr, nullptr);
                if (col < nc) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int q = mfma_row(l, r);
                        if (q < kb) Acol[p0 + q] = top[r];
                    }
                }
            }
        }
        if (active && col < nc) res_store<MT, false>(acc, Cc, 0, rows);
    }
}

// Sparsify apply of one interface block chunk (FP32), C resident in the accumulators:
//   RIGHT (pad 1): A <- A Q_n on a chunk of mc <= 128 rows, as Q_n^T A^T (C = A^T, kn <= 32 MT
//                  rows, read and written transposed: coalesced over A's rows);
//   left  (pad 2): A(:, chunk) <- Q_c^T A(:, chunk) on rn <= 128 columns, kc <= 32 MT rows, and
//                  rows 0 .. rc-1 go straight to their place in the parent block (dst).
// Same transformation as k_spars_apply for those phases.
template <int MT, bool RIGHT>
__global__ __launch_bounds__(WG) void k_spars_res(const SparsApplyTask<float>* __restrict__ tasks) {
    __shared__ ResSmem<MT> sm;
    const SparsApplyTask<float> t = tasks[blockIdx.x];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, li = l & 31;
    const float* __restrict__ V = RIGHT ? t.Vn : t.Vc;
    const float* __restrict__ T = RIGHT ? t.Tn : t.Tc;
    const int ldv = RIGHT ? t.ldvn : t.ldvc;
    const int m = RIGHT ? t.kn : t.kc, k = RIGHT ? t.rn : t.rc, n = RIGHT ? t.mc : t.rn;
    const int ntiles = (m + 31) >> 5, col = 32 * w + li, cc = min(col, n - 1);
    const bool active = 32 * w < n;
    float* __restrict__ base = RIGHT ? t.A + cc : t.A + (size_t)cc * t.lda;
    floatx16 acc[MT];
    if (active) res_load<MT, RIGHT>(acc, base, t.lda, m, col < n);
    for (int p0 = 0; p0 < k; p0 += 32) {
        const int kb = min(32, k - p0);
        __syncthreads();
        res_stage<MT, true, WG>(V + p0 + (size_t)p0 * ldv, ldv, T + (size_t)p0 * NB, m - p0, kb, sm);
        __syncthreads();
        res_panel<MT, false>(acc, 0, p0 >> 5, ntiles, sm.Vs, sm.Ts, nullptr, active, nullptr, nullptr);
    }
    if (active && col < n) {
        if (RIGHT) res_store<MT, true>(acc, base, t.lda, m);
        else res_store<MT, false>(acc, t.dst + (size_t)col * t.ldd, 0, t.rc);
    }
}

// C <- C R^{-1} (trsm 'R','U','N','N'; R k2 x k2 upper, ldr) for a chunk of m <= 128 rows of C
// (FP32), as the lower solve R^T X^T = C^T with C^T resident: wave w holds C rows 32 w .. as
// accumulator columns (read and written transposed, coalesced over C's rows).  Per 32-row
// block J: wave 0 inverts R_JJ (a column of the inverse per lane, in LDS), X_J = R_JJ^{-T} C_J
// and C_K -= R_JK^T X_J for the later blocks, all on MFMA with B from the resident tiles.
template <int MT>
struct TrsmSmem {
    static constexpr int kStride = 32 * MT + 4;  // strip row stride (16-byte rows)
    float Rs[32 * kStride];                      // R(J0 + q, J0 + c) at q * kStride + c, zero below the diagonal
    float Li[32 * 33];                           // R_JJ^{-1}(q, i) at q * 33 + i
};

template <int MT>
__global__ __launch_bounds__(WG) void k_trsm_res(const LarfbTask<float>* __restrict__ tasks) {
    __shared__ TrsmSmem<MT> sm;
    constexpr int ST = TrsmSmem<MT>::kStride;
    const LarfbTask<float> t = tasks[blockIdx.x];
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6, li = l & 31;
    const int m = t.m, k2 = t.k2, nJ = (k2 + 31) >> 5;
    const int row = 32 * w + li;
    const bool active = 32 * w < m;
    float* __restrict__ base = t.C + min(row, m - 1);
    floatx16 acc[MT];
    if (active) res_load<MT, true>(acc, base, t.ldc, k2, row < m);
    for (int J = 0; J < nJ; ++J) {
        const int J0 = 32 * J, wid = k2 - J0, kb = min(32, wid);
        __syncthreads();  // the previous block's readers of sm are done
        // strip R(J0 .. J0 + 31, J0 .. k2 - 1): q = tid % 32, columns tid / 32 + 8 i
        {
            const int q = tid & 31, c0 = tid >> 5;
            const float* __restrict__ Rq = t.R + J0 + min(q, kb - 1) + (size_t)J0 * t.ldr;
#pragma unroll 4
            for (int c = c0; c < 32 * MT; c += 8) {
                const float v = Rq[(size_t)min(c, wid - 1) * t.ldr];
                sm.Rs[q * ST + c] = (q < kb && c < wid && c >= q) ? v : 0.f;
            }
        }
        __syncthreads();
        if (tid < 32) {  // column c of R_JJ^{-1} by back substitution (identity past kb)
            const int c = tid;
            float s[32];
#pragma unroll
            for (int j = 0; j < 32; ++j) s[j] = 0.f;
#pragma unroll
            for (int i = 31; i >= 0; --i) {
                const float d = (i < kb) ? sm.Rs[i * ST + i] : 1.f;
                const float x = ((i == c ? 1.f : 0.f) - s[i]) / d;
#pragma unroll
                for (int j = 0; j < 32; ++j)
                    if (j < i) s[j] += sm.Rs[j * ST + i] * x;
                sm.Li[i * 33 + c] = x;
                __builtin_amdgcn_sched_barrier(0);  // this step's R loads only (else all 496 are hoisted)
            }
        }
        __syncthreads();
        if (active) {
            floatx16 X;
#pragma unroll
            for (int r = 0; r < 16; ++r) X[r] = 0.f;
#pragma unroll
            for (int JJ = 0; JJ < MT; ++JJ) {
                if (JJ == J) {  // tile J by a static index: X(i, .) = sum_q R_JJ^{-1}(q, i) C_J(q, .)
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        X = __builtin_amdgcn_mfma_f32_32x32x2f32(sm.Li[mfma_row(l, r) * 33 + li], acc[JJ][r], X, 0, 0, 0);
                    acc[JJ] = X;
                }
            }
#pragma unroll
            for (int K = 0; K < MT; ++K) {
                if (K > J && K < nJ) {  // C_K(i, .) -= sum_q R(J0 + q, K0 + i) X(q, .)
                    const float* rk = sm.Rs + 32 * (K - J);
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        acc[K] = __builtin_amdgcn_mfma_f32_32x32x2f32(-rk[mfma_row(l, r) * ST + li], X[r], acc[K], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    if (active && row < m) res_store<MT, true>(acc, base, t.ldc, k2);
}

// Blocks of 513 .. 1024 rows: four waves, two 32-column slabs (w % 2) times two 512-row halves
// (w / 2); the halves of a slab sum their partial W through LDS once per panel.
struct Res2Smem {
    ResSmem<32> v;        // V_p rows for the whole block (1024 rows)
    float Wx[4 * 1024];   // per-wave partial W
};

// larfb (Q^T C) with 512 < m <= 1024 and n <= 64.
__global__ __launch_bounds__(WG) void k_larfb_res2(const LarfbTask<float>* __restrict__ tasks) {
    __shared__ Res2Smem sm;
    const LarfbTask<float> t = tasks[blockIdx.x];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, li = l & 31, sl = w & 1, h = w >> 1;
    const int m = t.m, n = t.n, k = t.k, ntiles = (m + 31) >> 5;
    const int col = 32 * sl + li;
    const bool active = 32 * sl < n;
    float* __restrict__ Cc = t.C + (size_t)min(col, n - 1) * t.ldc + 512 * h;
    floatx16 acc[16];
    if (active) res_load<16, false>(acc, Cc, 0, m - 512 * h, col < n);
    for (int p0 = 0; p0 < k; p0 += 32) {
        const int kb = min(32, k - p0);
        __syncthreads();
        res_stage<32, true, WG>(t.V + p0 + (size_t)p0 * t.ldv, t.ldv, t.T + (size_t)p0 * NB, m - p0, kb, sm.v);
        __syncthreads();
        res_panel<16, true>(acc, 16 * h, p0 >> 5, ntiles, sm.v.Vs, sm.v.Ts, nullptr, active, sm.Wx + w * 1024,
                            sm.Wx + (w ^ 2) * 1024);
    }
    if (active && col < n) res_store<16, false>(acc, Cc, 0, m - 512 * h);
}

// tpmqrt chain of a tile (ncols <= 64) whose links have up to 1024 rows; half 0 carries Atop.
__global__ __launch_bounds__(WG) void k_tpmqrt_res2(const TpmqrtTask<float>* __restrict__ tasks,
                                                    const TpmqrtLink<float>* __restrict__ links) {
    __shared__ Res2Smem sm;
    const TpmqrtTask<float> t = tasks[blockIdx.x];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, li = l & 31, sl = w & 1, h = w >> 1;
    const int k = t.k, nc = t.ncols, col = 32 * sl + li, cc = min(col, nc - 1);
    const bool active = 32 * sl < nc;
    float* __restrict__ Acol = t.Atop + (size_t)cc * t.lda;
    floatx16 acc[16];
    for (int j = 0; j < t.nlinks; ++j) {
        const TpmqrtLink<float> L = links[t.link0 + j];
        const int rows = L.rows, ntiles = (rows + 31) >> 5;
        if (rows <= 0) continue;  // an empty link's reflectors are the identity
        const int rh = rows - 512 * h;  // this half's rows (<= 0: none)
        float* __restrict__ Cc = L.B + (size_t)(t.col0 + cc) * L.ldb + 512 * h;
        if (active && rh > 0) res_load<16, false>(acc, Cc, 0, rh, col < nc);
        for (int p0 = 0; p0 < k; p0 += 32) {
            const int kb = min(32, k - p0);
            __syncthreads();
            res_stage<32, false, WG>(L.V + (size_t)p0 * L.ldv, L.ldv, L.T + (size_t)p0 * NB, rows, kb, sm.v);
            __syncthreads();
            floatx16 top;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int q = mfma_row(l, r);
                const float v = (h == 0) ? Acol[p0 + min(q, kb - 1)] : 0.f;
                top[r] = (q < kb && col < nc) ? v : 0.f;
            }
            res_panel<16, true>(acc, 16 * h, 0, ntiles, sm.v.Vs, sm.v.Ts, &top, active, sm.Wx + w * 1024,
                                sm.Wx + (w ^ 2) * 1024);
            if (h == 0 && active && col < nc) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int q = mfma_row(l, r);
                    if (q < kb) Acol[p0 + q] = top[r];
                }
            }
        }
        if (active && col < nc && rh > 0) res_store<16, false>(acc, Cc, 0, rh);
    }
}

void launch_larfb_res(const LarfbTask<float>* d, int n, int mt, hipStream_t s) {
    if (n <= 0) return;
    if (mt == 4) hipLaunchKernelGGL(k_larfb_res<4>, dim3(n), dim3(WG), 0, s, d);
    else if (mt == 8) hipLaunchKernelGGL(k_larfb_res<8>, dim3(n), dim3(WG), 0, s, d);
    else hipLaunchKernelGGL(k_larfb_res<16>, dim3(n), dim3(WG), 0, s, d);
}
void launch_larfb_res2(const LarfbTask<float>* d, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_larfb_res2, dim3(n), dim3(WG), 0, s, d);
}
void launch_tpmqrt_res2(const TpmqrtTask<float>* d, const TpmqrtLink<float>* l, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_tpmqrt_res2, dim3(n), dim3(WG), 0, s, d, l);
}
void launch_trsm_res(const LarfbTask<float>* d, int n, int mt, hipStream_t s) {
    if (n <= 0) return;
    if (mt == 4) hipLaunchKernelGGL(k_trsm_res<4>, dim3(n), dim3(WG), 0, s, d);
    else if (mt == 8) hipLaunchKernelGGL(k_trsm_res<8>, dim3(n), dim3(WG), 0, s, d);
    else hipLaunchKernelGGL(k_trsm_res<16>, dim3(n), dim3(WG), 0, s, d);
}
void launch_spars_res(const SparsApplyTask<float>* d, int n, int mt, bool right, hipStream_t s) {
    if (n <= 0) return;
    if (right) {
        if (mt == 4) hipLaunchKernelGGL((k_spars_res<4, true>), dim3(n), dim3(WG), 0, s, d);
        else if (mt == 8) hipLaunchKernelGGL((k_spars_res<8, true>), dim3(n), dim3(WG), 0, s, d);
        else hipLaunchKernelGGL((k_spars_res<16, true>), dim3(n), dim3(WG), 0, s, d);
    } else {
        if (mt == 4) hipLaunchKernelGGL((k_spars_res<4, false>), dim3(n), dim3(WG), 0, s, d);
        else if (mt == 8) hipLaunchKernelGGL((k_spars_res<8, false>), dim3(n), dim3(WG), 0, s, d);
        else hipLaunchKernelGGL((k_spars_res<16, false>), dim3(n), dim3(WG), 0, s, d);
    }
}
void launch_tpmqrt_res(const TpmqrtTask<float>* d, const TpmqrtLink<float>* l, int n, int mt, hipStream_t s) {
    if (n <= 0) return;
    const dim3 b(kResTpmThreads);
    if (mt == 4) hipLaunchKernelGGL(k_tpmqrt_res<4>, dim3(n), b, 0, s, d, l);
    else if (mt == 8) hipLaunchKernelGGL(k_tpmqrt_res<8>, dim3(n), b, 0, s, d, l);
    else hipLaunchKernelGGL(k_tpmqrt_res<16>, dim3(n), b, 0, s, d, l);
}

}  // namespace spaqr
