// gfx950 kernels of the randomized / learned sparsification (see learned.hpp) and of the rank
// predictor's training (forward, backward, Adam) -- all one-workgroup-per-task batched launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>

#include "dpp.hpp"
#include "learned.hpp"

namespace spaqr {

typedef float floatx16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------------------
// Gaussian test matrix.  Entry (i, j) is a function of (i, j, seed) only -- splitmix64
// (S. Vigna's public-domain mixer) of the packed index gives two uniforms, Box-Muller gives
// the normal -- so a larger matrix regenerated later agrees with a smaller one on the overlap.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double gauss_entry(int i, int j, uint64_t seed) {
    const uint64_t h1 = splitmix64(seed ^ (((uint64_t)j << 32) | (uint32_t)i));
    const uint64_t h2 = splitmix64(h1);
    const double u1 = 1.0 - (double)(h1 >> 11) * 0x1.0p-53;  // (0, 1]
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;        // [0, 1)
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

template <typename S>
__global__ __launch_bounds__(WG) void k_gauss(S* om, int rows, int cols, int ld, uint64_t seed) {
    const size_t total = (size_t)rows * cols;
    for (size_t e = (size_t)blockIdx.x * WG + threadIdx.x; e < total; e += (size_t)gridDim.x * WG) {
        const int i = (int)(e % rows), j = (int)(e / rows);
        om[i + (size_t)j * ld] = (S)gauss_entry(i, j, seed);
    }
}

// The same matrix rounded to bf16 (round to nearest even) from its FP32 value, so that it is
// exactly the FP32 Omega rounded.  The rounding is done on the float's bits: written as
// (__bf16)(float)x the compiler fuses the two conversions into one from double, which rounds
// the float's ties by the discarded double bits instead of to even.
__device__ __forceinline__ uint16_t bf16_bits_rne(float f) {  // finite f
    const uint32_t b = __float_as_uint(f);
    return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

__global__ __launch_bounds__(WG) void k_gauss_bf16(uint16_t* om, int rows, int cols, int ld, uint64_t seed) {
    const size_t total = (size_t)rows * cols;
    for (size_t e = (size_t)blockIdx.x * WG + threadIdx.x; e < total; e += (size_t)gridDim.x * WG) {
        const int i = (int)(e % rows), j = (int)(e / rows);
        om[i + (size_t)j * ld] = bf16_bits_rne((float)gauss_entry(i, j, seed));
    }
}

// ---------------------------------------------------------------------------------------
// Column norms (16 lanes per column).
// ---------------------------------------------------------------------------------------
// Sum of squares of col[gl], col[gl + 16], ... (rows < k) accumulated in A, eight loads in
// flight per lane (a plain loop waits on every load).
template <typename A, typename S>
__device__ __forceinline__ A col_sumsq16(const S* __restrict__ col, int k, int gl) {
    A acc[4] = {A(0), A(0), A(0), A(0)};
    for (int r0 = gl; r0 < k; r0 += 16 * 8) {
        S v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = col[min(r0 + 16 * u, k - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const A x = (r0 + 16 * u < k) ? (A)v[u] : A(0);
            acc[u & 3] += x * x;
        }
    }
    return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

template <typename S>
__global__ __launch_bounds__(WG) void k_colnorm(const ColNormTask<S>* __restrict__ tasks) {
    __shared__ S red[4];
    const ColNormTask<S> t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, grp = tid >> 4, gl = tid & 15;
    S mx = S(0);
    for (int c0 = 0; c0 < t.w; c0 += WG / 16) {
        const int c = c0 + grp;
        S s = col_sumsq16<S>(t.C + (size_t)min(c, t.w - 1) * t.k, t.k, gl);
        s = row16_sum(s);
        if (c < t.w) {
            if (gl == 0) t.cn[c] = s;
            mx = max(mx, s);
        }
    }
    mx = wave_max(mx);
    if (lane == 0) red[wv] = mx;
    __syncthreads();
    if (tid == 0) {
        const S v = max(max(red[0], red[1]), max(red[2], red[3]));
        if constexpr (sizeof(S) == 8)
            atomicMax((unsigned long long*)t.cmax2, (unsigned long long)__double_as_longlong((double)v));
        else
            atomicMax((unsigned int*)t.cmax2, (unsigned int)__float_as_int((float)v));
    }
}

// ---------------------------------------------------------------------------------------
// Sketch Y = C Omega: 64 x 64 output tile per workgroup, w consumed in chunks of 32 through
// LDS; FP32 on MFMA (v_mfma_f32_32x32x2f32, one 32 x 32 block per wave: accumulator entry q of
// lane l is row 8(q/4) + 4(l/32) + q%4, column l%32), FP64 on VALU (lane l of wave q owns row
// r0 + l and columns c0 + 16q ..).  Loads are clamped in range and zero-weighted, so they issue
// unconditionally.
// ---------------------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(WG) void k_sketch(const SketchTask<S>* __restrict__ tasks,
                                               const SketchTile* __restrict__ tiles, const S* __restrict__ om,
                                               int ldo) {
    __shared__ S Cs[32][65];
    __shared__ S Os[64][33];
    const SketchTile tl = tiles[blockIdx.x];
    const SketchTask<S> t = tasks[tl.task];
    const int tid = threadIdx.x, row = tid & 63, wv = tid >> 6;
    S acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = S(0);
    // the next 32-column step of C and Omega is loaded into registers while this one computes
    S cv[8], ov[8];
    auto load_step = [&](int w0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * WG, rr = e & 63, cc = e >> 6;
            const int gr = min(tl.r0 + rr, t.k - 1), gc = min(w0 + cc, t.w - 1);
            cv[q] = t.C[gr + (size_t)gc * t.k];
            const int i = e & 31, j = e >> 5;
            ov[q] = om[min(w0 + i, t.w - 1) + (size_t)min(tl.c0 + j, t.s - 1) * ldo];
        }
    };
    load_step(0);
    for (int w0 = 0; w0 < t.w; w0 += 32) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * WG, rr = e & 63, cc = e >> 6;
            Cs[cc][rr] = (w0 + cc < t.w) ? cv[q] : S(0);
            const int i = e & 31, j = e >> 5;
            Os[j][i] = (w0 + i < t.w) ? ov[q] : S(0);
        }
        __syncthreads();
        if (w0 + 32 < t.w) load_step(w0 + 32);
        if constexpr (sizeof(S) == 4) {
            // MFMA 32x32x2: wave (wm, wn) owns the 32 x 32 block (rows 32 wm.., columns 32 wn..)
            const int li = tid & 31, lk = (tid >> 5) & 1, wm = wv & 1, wn = wv >> 1;
            floatx16& m = *reinterpret_cast<floatx16*>(acc);
#pragma unroll
            for (int kk = 0; kk < 32; kk += 2)
                m = __builtin_amdgcn_mfma_f32_32x32x2f32((float)Cs[kk + lk][32 * wm + li],
                                                         (float)Os[32 * wn + li][kk + lk], m, 0, 0, 0);
        } else {
#pragma unroll 8
            for (int i = 0; i < 32; ++i) {
                const S c = Cs[i][row];
#pragma unroll
                for (int j = 0; j < 16; ++j) acc[j] += c * Os[wv * 16 + j][i];
            }
        }
        __syncthreads();
    }
    if constexpr (sizeof(S) == 4) {
        const int li = tid & 31, wm = wv & 1, wn = wv >> 1, l = tid & 63;
        const int c = tl.c0 + 32 * wn + li;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int r = tl.r0 + 32 * wm + (q >> 2) * 8 + (l >> 5) * 4 + (q & 3);
            if (r < t.k && c < t.s) t.Y[r + (size_t)c * t.k] = acc[q];
        }
    } else {
        const int r = tl.r0 + row;
        if (r < t.k) {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int c = tl.c0 + wv * 16 + j;
                if (c < t.s) t.Y[r + (size_t)c * t.k] = acc[j];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// BF16 sketch Y = C Omega for FP32 C: C and Omega rounded to bf16 (round to nearest even) as
// they are staged in LDS, products accumulated in FP32 on v_mfma_f32_32x32x16_bf16.  One
// workgroup computes a 64 x TN tile (TN = 64 NC): wave (wm, wn) owns rows 32 wm .. and the NC
// 32-column blocks from TN/2 wn.  Operand layout of the instruction: lane l holds row
// (A) / column (B) l % 32 at the 8 consecutive k from 8 (l / 32); accumulator entry q of lane l
// is row 8 (q / 4) + 4 (l / 32) + q % 4, column l % 32.  Both LDS tiles keep k contiguous
// (rows of 40 bf16 = 80 bytes: 16-byte aligned, no bank conflicts for 16-byte accesses), so
// each operand is one 16-byte LDS read.  Omega is bf16 in HBM, column-major with ldo a
// multiple of 8 and at least w rounded up to 32 (one 16-byte load per thread and group; rows
// from w on are read and discarded, so their values do not matter); C is read once per TN
// columns of Y.  Tiles that share rows of C are consecutive in the tile list and run on one
// XCD (workgroups go to the 8 XCDs round-robin in blockIdx order), where the rows stay in L2.
// ---------------------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int full = n & ~7;
    return b >= full ? b : (b & 7) * (full >> 3) + (b >> 3);
}

template <int NC>
__global__ __launch_bounds__(WG) void k_sketch_bf16(const SketchTask<float>* __restrict__ tasks,
                                                    const SketchTile* __restrict__ tiles,
                                                    const __bf16* __restrict__ om, int ldo) {
    constexpr int TN = 64 * NC, KP = 40;
    __shared__ __attribute__((aligned(16))) __bf16 Cs[64 * KP];  // Cs[row][kk]
    __shared__ __attribute__((aligned(16))) __bf16 Os[TN * KP];  // Os[column][kk]
    const SketchTile tl = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const SketchTask<float> t = tasks[tl.task];
    const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6, wm = wv & 1, wn = wv >> 1;
    // C staging: thread owns row cr of the tile and the 8 columns 8 cg .. of each 32-column step
    const int cr = tid & 63, cg = tid >> 6;
    const float* Crow = t.C + min(tl.r0 + cr, t.k - 1);
    float cv[8];
    bf16x8 ov[NC];
    auto load_step = [&](int w0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) cv[q] = Crow[(size_t)min(w0 + 8 * cg + q, t.w - 1) * t.k];
#pragma unroll
        for (int u = 0; u < NC; ++u) {  // Omega group (column oc, rows 8 og ..) of the step
            const int p = tid + WG * u, og = p & 3, oc = p >> 2;
            ov[u] = *reinterpret_cast<const bf16x8*>(om + (size_t)min(tl.c0 + oc, t.s - 1) * ldo + w0 + 8 * og);
        }
    };
    floatx16 acc[NC];
#pragma unroll
    for (int nb = 0; nb < NC; ++nb)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[nb][q] = 0.f;
    load_step(0);
    for (int w0 = 0; w0 < t.w; w0 += 32) {
        bf16x8 cb;
#pragma unroll
        for (int q = 0; q < 8; ++q) cb[q] = (w0 + 8 * cg + q < t.w) ? (__bf16)cv[q] : (__bf16)0.f;
        *reinterpret_cast<bf16x8*>(&Cs[cr * KP + 8 * cg]) = cb;
#pragma unroll
        for (int u = 0; u < NC; ++u) {
            const int p = tid + WG * u, og = p & 3, oc = p >> 2;
            bf16x8 v = ov[u];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (w0 + 8 * og + q >= t.w) v[q] = (__bf16)0.f;
            *reinterpret_cast<bf16x8*>(&Os[oc * KP + 8 * og]) = v;
        }
        __syncthreads();
        if (w0 + 32 < t.w) load_step(w0 + 32);
#pragma unroll
        for (int sub = 0; sub < 2; ++sub) {
            const int kk = 16 * sub + 8 * (l >> 5);
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&Cs[(32 * wm + (l & 31)) * KP + kk]);
#pragma unroll
            for (int nb = 0; nb < NC; ++nb) {
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Os[(wn * (TN / 2) + 32 * nb + (l & 31)) * KP + kk]);
                acc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[nb], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int nb = 0; nb < NC; ++nb) {
        const int c = tl.c0 + wn * (TN / 2) + 32 * nb + (l & 31);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int r = tl.r0 + 32 * wm + (q >> 2) * 8 + (l >> 5) * 4 + (q & 3);
            if (r < t.k && c < t.s) t.Y[r + (size_t)c * t.k] = acc[nb][q];
        }
    }
}

// ---------------------------------------------------------------------------------------
// Rank features of one interface.
//   x[0..15]  decay of the R diagonal of the QR'd 16-column sketch: 1 + log10(|R_jj|/|R_00|)/8
//   x[16..26] log sizes, aspect, level, tolerance, neighbour counts, share of columns above
//             the tolerance, column-norm concentration and effective-rank ratio
// ---------------------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(WG) void k_rank_features(const FeatTask<S>* __restrict__ tasks, double tol) {
    __shared__ double red_s[4], red_c[4];
    const FeatTask<S> t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double cmax2 = (double)*t.cmax2;
    const double thr2 = tol * tol * cmax2;
    double sum = 0.0, cnt = 0.0;
    for (int j = tid; j < t.w; j += WG) {
        const double c = (double)t.cn[j];
        sum += c;
        cnt += (c >= thr2 && c > 0.0) ? 1.0 : 0.0;
    }
    sum = wave_sum(sum);
    cnt = wave_sum(cnt);
    if (lane == 0) { red_s[wv] = sum; red_c[wv] = cnt; }
    __syncthreads();
    if (tid < kRankFeatures) {
        const double tot = (red_s[0] + red_s[1]) + (red_s[2] + red_s[3]);
        const double count = (red_c[0] + red_c[1]) + (red_c[2] + red_c[3]);
        const int kw = min(t.k, t.w), nd = min(min(kSketchFeatureCols, t.k), t.w);
        const double r0 = nd > 0 ? fabs((double)t.Y16[0]) : 0.0;
        double f = 0.0;
        if (tid < kSketchFeatureCols) {
            if (tid < nd && r0 > 0.0) {
                const double rj = fabs((double)t.Y16[tid + (size_t)tid * t.k]);
                f = fmin(1.0, fmax(0.0, 1.0 + log10(rj / r0 + 1e-30) / 8.0));
            }
        } else {
            switch (tid) {
                case 16: f = log2(t.k + 1.0) / 11.0; break;
                case 17: f = log2(t.w + 1.0) / 14.0; break;
                case 18: f = (double)kw / (double)max(max(t.k, t.w), 1); break;
                case 19: f = t.level / 16.0; break;
                case 20: f = tol > 0.0 ? -log10(tol) / 8.0 : 1.0; break;
                case 21: f = t.nsout / 32.0; break;
                case 22: f = t.nsin / 32.0; break;
                case 23: f = t.w > 0 ? count / t.w : 0.0; break;
                case 24: f = kw > 0 ? fmin(1.0, count / kw) : 0.0; break;
                case 25: f = tot > 0.0 ? cmax2 / tot : 0.0; break;
                case 26: f = (cmax2 > 0.0 && kw > 0) ? fmin(1.0, tot / cmax2 / kw) : 0.0; break;
                default: f = 0.0;
            }
        }
        t.x[tid] = (float)f;
    }
}

// ---------------------------------------------------------------------------------------
// Tail selection on B = Q^T C: per column (16 lanes) the smallest r <= rY whose tail
// ||B(r:k, j)||^2 <= tol^2 max_j ||C(:, j)||^2, by an inclusive DPP scan of the squared
// entries in 16-row chunks; the interface's rank is the largest of these, or -1 when some
// column's tail stays above the threshold at rY.  Sums in FP64.
// ---------------------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(WG) void k_tail_select(const TailTask<S>* __restrict__ tasks, double tol) {
    __shared__ int red[4];
    const TailTask<S> t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, grp = tid >> 4, gl = tid & 15;
    const int rY = *t.ry;
    const double thr2 = tol * tol * (double)*t.cmax2;
    int need = 0;
    for (int c0 = 0; c0 < t.w; c0 += WG / 16) {
        const int c = c0 + grp;
        const S* col = t.B + (size_t)min(c, t.w - 1) * t.k;
        double tot = col_sumsq16<double>(col, t.k, gl);
        tot = row16_sum(tot);
        int rc = 0;
        if (c < t.w && tot > thr2) {
            rc = INT_MAX;
            double run = 0.0;
            S vnext = col[min(gl, t.k - 1)];  // the next chunk's entry is loaded one chunk ahead
            for (int b = 0; b <= rY; b += 16) {
                const int i = b + gl;
                const S vcur = vnext;
                vnext = col[min(i + 16, t.k - 1)];
                const double v = (i < rY && i < t.k) ? (double)vcur : 0.0;
                const double sq = v * v;
                const double incl = row16_scan(sq);
                const double tail = tot - (run + incl - sq);  // rows i .. k-1
                const unsigned long long bal = __ballot(i <= rY && tail <= thr2);
                const unsigned grpbits = (unsigned)((bal >> (lane & 48)) & 0xFFFFull);
                if (grpbits) {
                    rc = b + __ffs(grpbits) - 1;
                    break;
                }
                run += __shfl(incl, (lane & 48) | 15, 64);
            }
        }
        need = max(need, rc);
    }
    const int wn = max(max(lane_read(need, 0), lane_read(need, 16)), max(lane_read(need, 32), lane_read(need, 48)));
    if (lane == 0) red[wv] = wn;
    __syncthreads();
    if (tid == 0) atomicMax(t.rank, max(max(red[0], red[1]), max(red[2], red[3])));
}

// The same selection from the leading rows of B (ld r = rY) and the squared column norms of C.
template <typename S>
__global__ __launch_bounds__(WG) void k_tail_lead(const TailLeadTask<S>* __restrict__ tasks, double tol) {
    __shared__ int red[4];
    const TailLeadTask<S> t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, grp = tid >> 4, gl = tid & 15;
    const int rY = min(*t.ry, t.r);
    const double thr2 = tol * tol * (double)*t.cmax2;
    int need = 0;
    for (int c0 = 0; c0 < t.w; c0 += WG / 16) {
        const int c = c0 + grp;
        const S* col = t.B + (size_t)min(c, t.w - 1) * t.r;
        const double tot = (double)t.cn[min(c, t.w - 1)];
        int rc = 0;
        if (c < t.w && tot > thr2) {
            rc = INT_MAX;
            double run = 0.0;
            S vnext = col[min(gl, t.r - 1)];  // the next chunk's entry is loaded one chunk ahead
            for (int b = 0; b <= rY; b += 16) {
                const int i = b + gl;
                const S vcur = vnext;
                vnext = col[min(i + 16, t.r - 1)];
                const double v = (i < rY) ? (double)vcur : 0.0;
                const double sq = v * v;
                const double incl = row16_scan(sq);
                const double tail = tot - (run + incl - sq);  // rows i .. k-1
                const unsigned long long bal = __ballot(i <= rY && tail <= thr2);
                const unsigned grpbits = (unsigned)((bal >> (lane & 48)) & 0xFFFFull);
                if (grpbits) {
                    rc = b + __ffs(grpbits) - 1;
                    break;
                }
                run += __shfl(incl, (lane & 48) | 15, 64);
            }
        }
        need = max(need, rc);
    }
    const int wn = max(max(lane_read(need, 0), lane_read(need, 16)), max(lane_read(need, 32), lane_read(need, 48)));
    if (lane == 0) red[wv] = wn;
    __syncthreads();
    if (tid == 0) atomicMax(t.rank, max(max(red[0], red[1]), max(red[2], red[3])));
}

// B = M(0:r, 0:k) C: the FP32 sketch kernel's tiling (64 x 64 output tile per workgroup, the
// inner dimension in chunks of 32 through LDS, v_mfma_f32_32x32x2f32, one 32 x 32 block per
// wave) with M in place of C and C in place of Omega.  Loads are clamped in range and
// zero-weighted.
__global__ __launch_bounds__(WG) void k_tail_gemm(const TailGemmTask* __restrict__ tasks,
                                                  const SketchTile* __restrict__ tiles) {
    __shared__ float Ms[32][65];
    __shared__ float Cs[64][33];
    const SketchTile tl = tiles[blockIdx.x];
    const TailGemmTask t = tasks[tl.task];
    const int tid = threadIdx.x, wv = tid >> 6;
    float acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
    float mv[8], cv[8];
    auto load_step = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * WG, rr = e & 63, cc = e >> 6;
            mv[q] = t.M[min(tl.r0 + rr, t.r - 1) + (size_t)min(k0 + cc, t.k - 1) * t.ldm];
            const int i = e & 31, j = e >> 5;
            cv[q] = t.C[min(k0 + i, t.k - 1) + (size_t)min(tl.c0 + j, t.w - 1) * t.k];
        }
    };
    load_step(0);
    for (int k0 = 0; k0 < t.k; k0 += 32) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + q * WG, rr = e & 63, cc = e >> 6;
            Ms[cc][rr] = (k0 + cc < t.k) ? mv[q] : 0.0f;
            const int i = e & 31, j = e >> 5;
            Cs[j][i] = (k0 + i < t.k) ? cv[q] : 0.0f;
        }
        __syncthreads();
        if (k0 + 32 < t.k) load_step(k0 + 32);
        const int li = tid & 31, lk = (tid >> 5) & 1, wm = wv & 1, wn = wv >> 1;
        floatx16& m = *reinterpret_cast<floatx16*>(acc);
#pragma unroll
        for (int kk = 0; kk < 32; kk += 2)
            m = __builtin_amdgcn_mfma_f32_32x32x2f32(Ms[kk + lk][32 * wm + li], Cs[32 * wn + li][kk + lk], m, 0, 0, 0);
        __syncthreads();
    }
    const int li = tid & 31, wm = wv & 1, wn = wv >> 1, l = tid & 63;
    const int c = tl.c0 + 32 * wn + li;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = tl.r0 + 32 * wm + (q >> 2) * 8 + (l >> 5) * 4 + (q & 3);
        if (r < t.r && c < t.w) t.B[r + (size_t)c * t.r] = acc[q];
    }
}

template <typename S>
__global__ __launch_bounds__(WG) void k_eye(const EyeTask<S>* __restrict__ tasks) {
    const EyeTask<S> t = tasks[blockIdx.x];
    for (long e = threadIdx.x; e < (long)t.k * t.k; e += WG) {
        const int i = (int)(e % t.k), j = (int)(e / t.k);
        t.M[i + (size_t)j * t.ld] = (i == j) ? S(1) : S(0);
    }
}

// ---------------------------------------------------------------------------------------
// MLP: prediction (one thread per sample, weights in LDS).
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_mlp_predict(const float* __restrict__ X, int n, const float* __restrict__ W,
                                                    float* __restrict__ out) {
    __shared__ float w[kMlpParams];
    for (int p = threadIdx.x; p < kMlpParams; p += WG) w[p] = W[p];
    __syncthreads();
    const int i = blockIdx.x * WG + threadIdx.x;
    if (i >= n) return;
    float x[kRankFeatures];
#pragma unroll
    for (int f = 0; f < kRankFeatures; ++f) x[f] = X[(size_t)i * kRankFeatures + f];
    float h1[kMlpHidden];
#pragma unroll
    for (int h = 0; h < kMlpHidden; ++h) {
        float a = w[kMlpOffB1 + h];
#pragma unroll
        for (int f = 0; f < kRankFeatures; ++f) a += w[kMlpOffW1 + h * kRankFeatures + f] * x[f];
        h1[h] = tanhf(a);
    }
    float o = w[kMlpOffB3];
    for (int h = 0; h < kMlpHidden; ++h) {
        float a = w[kMlpOffB2 + h];
#pragma unroll
        for (int g = 0; g < kMlpHidden; ++g) a += w[kMlpOffW2 + h * kMlpHidden + g] * h1[g];
        o += w[kMlpOffW3 + h] * tanhf(a);
    }
    out[i] = 1.0f / (1.0f + expf(-o));
}

// ---------------------------------------------------------------------------------------
// MLP: gradient of L = mean_i [softplus(o_i) - y_i o_i] (logistic loss on rank fractions).
// Each workgroup walks its chunk of samples in order; threads 0..63 carry one hidden unit
// through the forward and backward sweep and every thread accumulates a fixed set of
// parameter gradients in registers, written once per workgroup.
// ---------------------------------------------------------------------------------------
constexpr int kParamsPerThread = (kMlpParams + WG - 1) / WG;  // 25

int mlp_grad_blocks(int n) { return std::max(1, std::min(256, (n + 31) / 32)); }

__global__ __launch_bounds__(WG) void k_mlp_grad(const float* __restrict__ X, const float* __restrict__ y, int n,
                                                 const float* __restrict__ W, float* __restrict__ partial) {
    __shared__ float w[kMlpParams];
    __shared__ float xs[kRankFeatures], h1[kMlpHidden], h2[kMlpHidden], da1[kMlpHidden], da2[kMlpHidden];
    __shared__ float dos;
    const int tid = threadIdx.x;
    for (int p = tid; p < kMlpParams; p += WG) w[p] = W[p];
    const int per = (n + gridDim.x - 1) / gridDim.x;
    const int s0 = blockIdx.x * per, s1 = min(n, s0 + per);
    float g[kParamsPerThread];
#pragma unroll
    for (int q = 0; q < kParamsPerThread; ++q) g[q] = 0.f;
    float loss = 0.f;
    const float inv_n = 1.0f / (float)n;
    __syncthreads();
    for (int sm = s0; sm < s1; ++sm) {
        if (tid < kRankFeatures) xs[tid] = X[(size_t)sm * kRankFeatures + tid];
        __syncthreads();
        if (tid < kMlpHidden) {
            float a = w[kMlpOffB1 + tid];
            for (int f = 0; f < kRankFeatures; ++f) a += w[kMlpOffW1 + tid * kRankFeatures + f] * xs[f];
            h1[tid] = tanhf(a);
        }
        __syncthreads();
        if (tid < kMlpHidden) {  // wave 0 exactly
            float a = w[kMlpOffB2 + tid];
            for (int g2 = 0; g2 < kMlpHidden; ++g2) a += w[kMlpOffW2 + tid * kMlpHidden + g2] * h1[g2];
            const float hh = tanhf(a);
            h2[tid] = hh;
            const float o = wave_sum(w[kMlpOffW3 + tid] * hh) + w[kMlpOffB3];
            const float p = 1.0f / (1.0f + expf(-o));
            const float yt = y[sm];
            const float d_o = (p - yt) * inv_n;
            da2[tid] = d_o * w[kMlpOffW3 + tid] * (1.0f - hh * hh);
            if (tid == 0) {
                dos = d_o;
                loss += (fmaxf(o, 0.f) - o * yt + log1pf(expf(-fabsf(o)))) * inv_n;
            }
        }
        __syncthreads();
        if (tid < kMlpHidden) {
            float d = 0.f;
            for (int h = 0; h < kMlpHidden; ++h) d += w[kMlpOffW2 + h * kMlpHidden + tid] * da2[h];
            da1[tid] = d * (1.0f - h1[tid] * h1[tid]);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kParamsPerThread; ++q) {
            const int p = tid + q * WG;
            float c = 0.f;
            if (p < kMlpOffB1) c = da1[p / kRankFeatures] * xs[p % kRankFeatures];
            else if (p < kMlpOffW2) c = da1[p - kMlpOffB1];
            else if (p < kMlpOffB2) { const int e = p - kMlpOffW2; c = da2[e / kMlpHidden] * h1[e % kMlpHidden]; }
            else if (p < kMlpOffW3) c = da2[p - kMlpOffB2];
            else if (p < kMlpOffB3) c = dos * h2[p - kMlpOffW3];
            else if (p == kMlpOffB3) c = dos;
            g[q] += c;
        }
        __syncthreads();
    }
    float* out = partial + (size_t)blockIdx.x * (kMlpParams + 1);
#pragma unroll
    for (int q = 0; q < kParamsPerThread; ++q) {
        const int p = tid + q * WG;
        if (p < kMlpParams) out[p] = g[q];
    }
    if (tid == 0) out[kMlpParams] = loss;
}

// grad[p] = sum_b partial[b][p] in block order; loss likewise; optional Adam step
// (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected).
__global__ __launch_bounds__(WG) void k_mlp_reduce(const float* __restrict__ partial, int nblocks, float* grad,
                                                   float* loss, float* W, float* m, float* v, int step, float lr,
                                                   int adam) {
    const int p = blockIdx.x * WG + threadIdx.x;
    if (p > kMlpParams) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += partial[(size_t)b * (kMlpParams + 1) + p];
    if (p == kMlpParams) {
        if (loss) *loss = s;
        return;
    }
    if (grad) grad[p] = s;
    if (adam) {
        const float b1 = 0.9f, b2 = 0.999f;
        const float mm = b1 * m[p] + (1.f - b1) * s;
        const float vv = b2 * v[p] + (1.f - b2) * s * s;
        m[p] = mm;
        v[p] = vv;
        const float mh = mm / (1.f - powf(b1, (float)step));
        const float vh = vv / (1.f - powf(b2, (float)step));
        W[p] -= lr * mh / (sqrtf(vh) + 1e-8f);
    }
}

// ---------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------
template <typename S>
void launch_gauss(S* om, int rows, int cols, int ld, uint64_t seed, hipStream_t s) {
    const size_t total = (size_t)rows * cols;
    if (total == 0) return;
    const int blocks = (int)std::min<size_t>((total + WG - 1) / WG, 4096);
    hipLaunchKernelGGL(k_gauss<S>, dim3(blocks), dim3(WG), 0, s, om, rows, cols, ld, seed);
}
template <typename S>
void launch_colnorm(const ColNormTask<S>* d, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_colnorm<S>, dim3(n), dim3(WG), 0, s, d);
}
template <typename S>
void launch_sketch(const SketchTask<S>* d, const SketchTile* tiles, int ntiles, const S* om, int ldo, hipStream_t s) {
    if (ntiles > 0) hipLaunchKernelGGL(k_sketch<S>, dim3(ntiles), dim3(WG), 0, s, d, tiles, om, ldo);
}
void launch_gauss_bf16(void* om, int rows, int cols, int ld, uint64_t seed, hipStream_t s) {
    const size_t total = (size_t)rows * cols;
    if (total == 0) return;
    const int blocks = (int)std::min<size_t>((total + WG - 1) / WG, 4096);
    hipLaunchKernelGGL(k_gauss_bf16, dim3(blocks), dim3(WG), 0, s, (uint16_t*)om, rows, cols, ld, seed);
}
void launch_sketch_bf16(const SketchTask<float>* d, const SketchTile* tiles, int ntiles, const void* om, int ldo,
                        int tile_cols, hipStream_t s) {
    if (ntiles <= 0) return;
    const __bf16* o = (const __bf16*)om;
    if (tile_cols == 256) hipLaunchKernelGGL(k_sketch_bf16<4>, dim3(ntiles), dim3(WG), 0, s, d, tiles, o, ldo);
    else if (tile_cols == 128) hipLaunchKernelGGL(k_sketch_bf16<2>, dim3(ntiles), dim3(WG), 0, s, d, tiles, o, ldo);
    else hipLaunchKernelGGL(k_sketch_bf16<1>, dim3(ntiles), dim3(WG), 0, s, d, tiles, o, ldo);
}
template <typename S>
void launch_rank_features(const FeatTask<S>* d, int n, double tol, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_rank_features<S>, dim3(n), dim3(WG), 0, s, d, tol);
}
template <typename S>
void launch_tail_select(const TailTask<S>* d, int n, double tol, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_tail_select<S>, dim3(n), dim3(WG), 0, s, d, tol);
}
template <typename S>
void launch_tail_lead(const TailLeadTask<S>* d, int n, double tol, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_tail_lead<S>, dim3(n), dim3(WG), 0, s, d, tol);
}
void launch_tail_gemm(const TailGemmTask* d, const SketchTile* tiles, int ntiles, hipStream_t s) {
    if (ntiles > 0) hipLaunchKernelGGL(k_tail_gemm, dim3(ntiles), dim3(WG), 0, s, d, tiles);
}
template <typename S>
void launch_eye(const EyeTask<S>* d, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_eye<S>, dim3(n), dim3(WG), 0, s, d);
}
void launch_mlp_predict(const float* X, int n, const float* W, float* out, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_mlp_predict, dim3((n + WG - 1) / WG), dim3(WG), 0, s, X, n, W, out);
}
void launch_mlp_grad(const float* X, const float* y, int n, const float* W, float* partial, int nblocks,
                     hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(WG), 0, s, X, y, n, W, partial);
}
void launch_mlp_reduce(const float* partial, int nblocks, float* grad, float* loss, float* W, float* m, float* v,
                       int step, float lr, int adam, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 1 + WG - 1) / WG), dim3(WG), 0, s, partial, nblocks, grad,
                       loss, W, m, v, step, lr, adam);
}

#define SPAQR_INST(S)                                                                                       \
    template void launch_gauss<S>(S*, int, int, int, uint64_t, hipStream_t);                                 \
    template void launch_colnorm<S>(const ColNormTask<S>*, int, hipStream_t);                                \
    template void launch_sketch<S>(const SketchTask<S>*, const SketchTile*, int, const S*, int, hipStream_t); \
    template void launch_rank_features<S>(const FeatTask<S>*, int, double, hipStream_t);                     \
    template void launch_tail_select<S>(const TailTask<S>*, int, double, hipStream_t);                     \
    template void launch_tail_lead<S>(const TailLeadTask<S>*, int, double, hipStream_t);                   \
    template void launch_eye<S>(const EyeTask<S>*, int, hipStream_t);
SPAQR_INST(float)
SPAQR_INST(double)
#undef SPAQR_INST

}  // namespace spaqr
