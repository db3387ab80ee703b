// Grouped geqp3 driver: packs the interfaces of one level into cooperative launches whose
// workgroups hold LDS-resident column slices (k_geqp3_coop); an interface too large for one
// cooperative launch runs on the single-workgroup kernel (k_geqp3) instead.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <vector>

#include "kernels.hpp"

namespace spaqr {

struct RrqrAlloc {
    std::function<void*(size_t)> alloc;                  // device scratch, alive until the stream syncs
    std::function<void*(const void*, size_t)> upload;    // async H2D copy of a host buffer, returns device ptr
    // the launches are being recorded into a hip graph: counters that must start at zero are
    // cleared by a kernel of the graph (an upload is not replayed), and the cooperative geqp3 is
    // an ordinary launch (its grid never exceeds the co-resident limit; a wait that exceeds its
    // spin limit raises the error flag instead of hanging)
    bool graph = false;
};

// Cooperative geqp3 workgroups resident per CU: each gets an equal share of the CU's LDS for
// its column slice, norms and reflector.
constexpr int kGeqp3CoopBlocksPerCU = 2;

// Returns the number of tasks that ran on the single-workgroup fallback.  *d_err (device int,
// zero on entry) becomes non-zero if a cooperative barrier wait hit its spin limit.
// *dispatches (if non-null) receives the number of kernel launches issued.
template <typename S>
int geqp3_grouped(const std::vector<Geqp3Task<S>>& tasks, double tol, double tie_rel, hipStream_t stream,
                  const RrqrAlloc& mem, int* d_err, int* dispatches = nullptr);

}  // namespace spaqr
