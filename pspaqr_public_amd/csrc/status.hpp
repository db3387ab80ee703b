// Per-cluster status flags of a factorization (spaqr_tree_get_info).
#pragma once
#include <hip/hip_runtime.h>

namespace spaqr {

constexpr int kInfoNonFinite = 1;      // a non-finite entry on the diagonal of the R factor
constexpr int kInfoZeroPivot = 2;      // a zero pivot (or one at or below the set relative threshold)
constexpr int kInfoPredictorMiss = 4;  // the sketch failed the tail test; geqp3 gave the rank

// Diagonal of an interior's R (k x k upper, ld) after its elimination.
template <typename S> struct PivotCheck { const S* R; int k, ld, order, level; };

template <typename S>
void launch_check_pivots(const PivotCheck<S>* d, int n, double rel, int* cinfo, int* lflags, int* lfirst, hipStream_t s);

}  // namespace spaqr
