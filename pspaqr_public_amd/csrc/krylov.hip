// FP64 Krylov pieces over a batch of systems: the unknowns of system s occupy positions
// [off[s], off[s + 1]) of every vector, and each system has its own scalars.  One launch
// serves all systems (a fixed number of workgroups per system), and every reduction runs in a
// fixed order, so the results do not depend on timing.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dpp.hpp"
#include "kernels.hpp"
#include "launch.hpp"

namespace spaqr {

namespace {
constexpr int kSegBlocks = 32;  // workgroups per system and launch
}

// partial[s * kSegBlocks + c] = sum over chunk c of system s of a .* b
__global__ __launch_bounds__(WG) void k_seg_dot_partial(const int* __restrict__ off, const double* __restrict__ a,
                                                        const double* __restrict__ b, double* __restrict__ partial) {
    __shared__ double red[4];
    const int s = blockIdx.x / kSegBlocks, c = blockIdx.x % kSegBlocks;
    const int lo = off[s], hi = off[s + 1];
    double acc = 0.0;
    for (int i = lo + c * WG + (int)threadIdx.x; i < hi; i += kSegBlocks * WG) acc += a[i] * b[i];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// out[s * ostride] = sum of system s's partials (its square root if sqrt_mode)
__global__ __launch_bounds__(WG) void k_seg_dot_final(int nseg, const double* __restrict__ partial,
                                                      double* __restrict__ out, int ostride, int sqrt_mode) {
    for (int s = blockIdx.x * WG + threadIdx.x; s < nseg; s += gridDim.x * WG) {
        double v = 0.0;
        for (int c = 0; c < kSegBlocks; ++c) v += partial[s * kSegBlocks + c];
        out[(size_t)s * ostride] = sqrt_mode ? sqrt(v) : v;
    }
}

// y <- y - h[s * hstride] x on system s
__global__ __launch_bounds__(WG) void k_seg_axpy(const int* __restrict__ off, const double* __restrict__ h, int hstride,
                                                 const double* __restrict__ x, double* __restrict__ y) {
    const int s = blockIdx.x / kSegBlocks, c = blockIdx.x % kSegBlocks;
    const int lo = off[s], hi = off[s + 1];
    const double a = h[(size_t)s * hstride];
    for (int i = lo + c * WG + (int)threadIdx.x; i < hi; i += kSegBlocks * WG) y[i] -= a * x[i];
}

// y <- x / h[s * hstride] on systems with keep[s] != 0 and h > 0; y <- 0 on the others
__global__ __launch_bounds__(WG) void k_seg_scal(const int* __restrict__ off, const double* __restrict__ h, int hstride,
                                                 const int* __restrict__ keep, const double* __restrict__ x,
                                                 double* __restrict__ y) {
    const int s = blockIdx.x / kSegBlocks, c = blockIdx.x % kSegBlocks;
    const int lo = off[s], hi = off[s + 1];
    const double d = h[(size_t)s * hstride];
    const bool on = keep[s] != 0 && d > 0.0;
    const double inv = on ? 1.0 / d : 0.0;
    for (int i = lo + c * WG + (int)threadIdx.x; i < hi; i += kSegBlocks * WG) y[i] = on ? x[i] * inv : 0.0;
}

// x <- x + coef[s * cstride] v on system s
__global__ __launch_bounds__(WG) void k_seg_axpy_coef(const int* __restrict__ off, const double* __restrict__ coef,
                                                      int cstride, const double* __restrict__ v, double* __restrict__ x) {
    const int s = blockIdx.x / kSegBlocks, c = blockIdx.x % kSegBlocks;
    const int lo = off[s], hi = off[s + 1];
    const double a = coef[(size_t)s * cstride];
    if (a == 0.0) return;
    for (int i = lo + c * WG + (int)threadIdx.x; i < hi; i += kSegBlocks * WG) x[i] += a * v[i];
}

int seg_partials(int nseg) { return nseg * kSegBlocks; }

void launch_seg_dot(int nseg, const int* off, const double* a, const double* b, double* partial, double* out,
                    int ostride, int sqrt_mode, hipStream_t s) {
    if (nseg <= 0) return;
    hipLaunchKernelGGL(k_seg_dot_partial, dim3(nseg * kSegBlocks), dim3(WG), 0, s, off, a, b, partial);
    hipLaunchKernelGGL(k_seg_dot_final, dim3((nseg + WG - 1) / WG), dim3(WG), 0, s, nseg, partial, out, ostride,
                       sqrt_mode);
}
void launch_seg_axpy(int nseg, const int* off, const double* h, int hstride, const double* x, double* y, hipStream_t s) {
    if (nseg > 0) hipLaunchKernelGGL(k_seg_axpy, dim3(nseg * kSegBlocks), dim3(WG), 0, s, off, h, hstride, x, y);
}
void launch_seg_scal(int nseg, const int* off, const double* h, int hstride, const int* keep, const double* x,
                     double* y, hipStream_t s) {
    if (nseg > 0) hipLaunchKernelGGL(k_seg_scal, dim3(nseg * kSegBlocks), dim3(WG), 0, s, off, h, hstride, keep, x, y);
}
void launch_seg_axpy_coef(int nseg, const int* off, const double* coef, int cstride, const double* v, double* x,
                          hipStream_t s) {
    if (nseg > 0)
        hipLaunchKernelGGL(k_seg_axpy_coef, dim3(nseg * kSegBlocks), dim3(WG), 0, s, off, coef, cstride, v, x);
}

// ---------------------------------------------------------------------------------------
// One system: classical Gram-Schmidt applied twice (CGS2) in fused passes, and the GMRES
// Givens / least-squares steps on the device (single-thread kernels: a few hundred flops).
// Row chunks of kRowsPerBlock rows; the partial sums of every chunk are reduced in a fixed
// order by one wave per coefficient, so results do not depend on timing.
// ---------------------------------------------------------------------------------------
namespace {
constexpr int kRowsPerThread = 4, kRowsPerBlock = kRowsPerThread * WG;
}

int mdot_blocks(int n) { return std::max(1, (n + kRowsPerBlock - 1) / kRowsPerBlock); }

// part[i * nb + blk] = sum over chunk blk of Q(:, i) .* v, i < nvec (Q columns of length n)
__global__ __launch_bounds__(WG) void k_mdot_partial(int n, const double* __restrict__ Q, int nvec,
                                                     const double* __restrict__ v, double* __restrict__ part) {
    __shared__ double red[4][64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nb = gridDim.x;
    const int r0 = blockIdx.x * kRowsPerBlock + tid;
    double vr[kRowsPerThread];
#pragma unroll
    for (int u = 0; u < kRowsPerThread; ++u) vr[u] = (r0 + u * WG < n) ? v[r0 + u * WG] : 0.0;
    for (int i0 = 0; i0 < nvec; i0 += 64) {
        const int ni = min(64, nvec - i0);
        for (int ii = 0; ii < ni; ++ii) {
            const double* q = Q + (size_t)(i0 + ii) * n;
            double s = 0.0;
#pragma unroll
            for (int u = 0; u < kRowsPerThread; ++u)
                if (r0 + u * WG < n) s += q[r0 + u * WG] * vr[u];
            s = wave_sum(s);
            if (lane == 0) red[wv][ii] = s;
        }
        __syncthreads();
        if (tid < ni) part[(size_t)(i0 + tid) * nb + blockIdx.x] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
        __syncthreads();
    }
}

// c[i] = sum of part[i * nb ..]; h[i] = c[i] (accumulate == 0) or h[i] + c[i]; one wave per i
__global__ __launch_bounds__(WG) void k_mdot_final(int nvec, int nb, const double* __restrict__ part,
                                                   double* __restrict__ c, double* __restrict__ h, int accumulate) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= nvec) return;  // whole waves leave together
    double s = 0.0;
    for (int b = lane; b < nb; b += 64) s += part[(size_t)i * nb + b];
    s = wave_sum(s);
    if (lane == 0) {
        c[i] = s;
        h[i] = accumulate ? h[i] + s : s;
    }
}

// v <- v - Q(:, 0:nvec) c; with normpart, normpart[blk] = sum over the chunk of v_new^2
__global__ __launch_bounds__(WG) void k_cgs_update(int n, const double* __restrict__ Q, int nvec,
                                                   const double* __restrict__ c, double* __restrict__ v,
                                                   double* __restrict__ normpart) {
    __shared__ double red[4];
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * kRowsPerBlock + tid;
    double acc[kRowsPerThread];
#pragma unroll
    for (int u = 0; u < kRowsPerThread; ++u) acc[u] = (r0 + u * WG < n) ? v[r0 + u * WG] : 0.0;
    for (int i = 0; i < nvec; ++i) {
        const double ci = c[i];
        const double* q = Q + (size_t)i * n;
#pragma unroll
        for (int u = 0; u < kRowsPerThread; ++u)
            if (r0 + u * WG < n) acc[u] -= q[r0 + u * WG] * ci;
    }
    double ss = 0.0;
#pragma unroll
    for (int u = 0; u < kRowsPerThread; ++u)
        if (r0 + u * WG < n) {
            v[r0 + u * WG] = acc[u];
            ss += acc[u] * acc[u];
        }
    if (!normpart) return;
    ss = wave_sum(ss);
    if ((tid & 63) == 0) red[tid >> 6] = ss;
    __syncthreads();
    if (tid == 0) normpart[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Step k of GMRES's Givens QR of the Hessenberg matrix (the rotations of dist_gmres,
// src/is.cpp:489-509): hcol[0..k] is column k from the orthogonalisation, the norm of the new
// vector is the sum of normpart (nb chunks, one wave, fixed order) and goes to hcol[k + 1]; the
// previous rotations and the new one are applied, the rotated column is stored as column k of
// H (ld ldh), beta is updated and est[0] = |beta[k + 1]| / ref_norm.
__global__ __launch_bounds__(64) void k_givens_step(int k, int nb, const double* __restrict__ normpart,
                                                    double* __restrict__ hcol, double* __restrict__ H, int ldh,
                                                    double* __restrict__ cs, double* __restrict__ sn,
                                                    double* __restrict__ beta, double ref_norm,
                                                    double* __restrict__ est) {
    const int lane = threadIdx.x;
    double s = 0.0;
    for (int b = lane; b < nb; b += 64) s += normpart[b];
    s = wave_sum(s);
    if (lane != 0) return;
    const double hk1 = sqrt(s);
    hcol[k + 1] = hk1;
    double* h = H + (size_t)k * ldh;
    for (int i = 0; i <= k; ++i) h[i] = hcol[i];
    h[k + 1] = hk1;
    for (int i = 0; i < k; ++i) {
        const double tmp = cs[i] * h[i] + sn[i] * h[i + 1];
        h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
        h[i] = tmp;
    }
    double c, sv;
    if (h[k] == 0.0) {
        c = 0.0;
        sv = 1.0;
    } else {
        const double tt = sqrt(h[k] * h[k] + h[k + 1] * h[k + 1]);
        c = fabs(h[k]) / tt;
        sv = c * h[k + 1] / h[k];
    }
    cs[k] = c;
    sn[k] = sv;
    h[k] = c * h[k] + sv * h[k + 1];
    h[k + 1] = 0.0;
    beta[k + 1] = -sv * beta[k];
    beta[k] = c * beta[k];
    est[0] = fabs(beta[k + 1]) / ref_norm;
}

// y(0:kk) solves the upper triangular H(0:kk, 0:kk) y = beta(0:kk) (back substitution)
__global__ __launch_bounds__(64) void k_lsq_solve(int kk, const double* __restrict__ H, int ldh,
                                                  const double* __restrict__ beta, double* __restrict__ y) {
    if (threadIdx.x != 0) return;
    for (int i = kk - 1; i >= 0; --i) {
        double t = beta[i];
        for (int j = i + 1; j < kk; ++j) t -= H[i + (size_t)j * ldh] * y[j];
        y[i] = t / H[i + (size_t)i * ldh];
    }
}

// x <- x + V(:, 0:kk) y, the columns added in order
__global__ __launch_bounds__(WG) void k_basis_update(int n, const double* __restrict__ V, int kk,
                                                     const double* __restrict__ y, double* __restrict__ x) {
    for (int r = blockIdx.x * WG + threadIdx.x; r < n; r += gridDim.x * WG) {
        double a = x[r];
        for (int j = 0; j < kk; ++j) a += y[j] * V[(size_t)j * n + r];
        x[r] = a;
    }
}

void launch_cgs2(int n, const double* Q, int nvec, double* v, double* part, double* c, double* hcol, double* normpart,
                 hipStream_t s) {
    if (nvec <= 0) return;
    const int nb = mdot_blocks(n);
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_mdot_partial, dim3(nb), dim3(WG), 0, s, n, Q, nvec, v, part);
        hipLaunchKernelGGL(k_mdot_final, dim3((nvec + 3) / 4), dim3(WG), 0, s, nvec, nb, part, c, hcol, pass);
        hipLaunchKernelGGL(k_cgs_update, dim3(nb), dim3(WG), 0, s, n, Q, nvec, c, v, pass == 1 ? normpart : nullptr);
    }
}
void launch_givens_step(int k, int nb, const double* normpart, double* hcol, double* H, int ldh, double* cs, double* sn,
                        double* beta, double ref_norm, double* est, hipStream_t s) {
    hipLaunchKernelGGL(k_givens_step, dim3(1), dim3(64), 0, s, k, nb, normpart, hcol, H, ldh, cs, sn, beta, ref_norm,
                       est);
}
void launch_lsq_update(int n, int kk, const double* H, int ldh, const double* beta, double* y, const double* V,
                       double* x, hipStream_t s) {
    if (kk <= 0) return;
    hipLaunchKernelGGL(k_lsq_solve, dim3(1), dim3(64), 0, s, kk, H, ldh, beta, y);
    hipLaunchKernelGGL(k_basis_update, dim3(std::min((n + WG - 1) / WG, 2048)), dim3(WG), 0, s, n, V, kk, y, x);
}

}  // namespace spaqr
