// Per-block numerical status of a factorization (the ABI's info[]): one wavefront per interior
// scans the diagonal of its R factor for non-finite entries and for pivots at or below
// rel * max_i |R_ii| (rel = 0: exactly zero pivots only).  Flags go to the cluster's entry and,
// with the first cluster order that raised them, to the level's entry, which the host reads at
// its next synchronisation.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "dpp.hpp"
#include "kernels.hpp"
#include "status.hpp"

namespace spaqr {

template <typename S>
__global__ __launch_bounds__(64) void k_check_pivots(const PivotCheck<S>* __restrict__ tasks, double rel,
                                                     int* __restrict__ cinfo, int* __restrict__ lflags,
                                                     int* __restrict__ lfirst) {
    const PivotCheck<S> t = tasks[blockIdx.x];
    const int lane = threadIdx.x;
    double mx = 0.0;
    bool bad = false;
    for (int i = lane; i < t.k; i += 64) {
        const double d = fabs((double)t.R[i + (size_t)i * t.ld]);
        if (!(d <= DBL_MAX)) bad = true;  // NaN or infinite
        else mx = fmax(mx, d);
    }
    mx = wave_max(mx);
    bool tiny = false;
    const double thr = rel * mx;
    for (int i = lane; i < t.k; i += 64) {
        const double d = fabs((double)t.R[i + (size_t)i * t.ld]);
        if (d <= DBL_MAX && (d == 0.0 || d <= thr)) tiny = true;
    }
    const int f = (__ballot(bad) ? kInfoNonFinite : 0) | (__ballot(tiny) ? kInfoZeroPivot : 0);
    if (lane == 0) {
        cinfo[t.order] = f;
        if (f) {
            atomicOr(lflags + t.level, f);
            atomicMin(lfirst + t.level, t.order);
        }
    }
}

template <typename S>
void launch_check_pivots(const PivotCheck<S>* d, int n, double rel, int* cinfo, int* lflags, int* lfirst, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_check_pivots<S>, dim3(n), dim3(64), 0, s, d, rel, cinfo, lflags, lfirst);
}

template void launch_check_pivots<float>(const PivotCheck<float>*, int, double, int*, int*, int*, hipStream_t);
template void launch_check_pivots<double>(const PivotCheck<double>*, int, double, int*, int*, int*, hipStream_t);

}  // namespace spaqr
