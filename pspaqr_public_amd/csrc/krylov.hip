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

}  // namespace spaqr
