// Definitions shared by the translation units of the level planner (tree.cpp, tree_rank.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "tree.hpp"

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t _e = (x);                                                                         \
        if (_e != hipSuccess)                                                                        \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " #x); \
    } while (0)

namespace spaqr {

// geqp3 pivot candidates within this relative band of the largest norm count as tied and
// the lowest index wins (LAPACK's plain argmax lets rounding order decide exact ties).
constexpr double kPivotTieRel = 1e-10;

template <class F>
void Tree::timed(int cat, long launches, long tasks, F&& f) {
    if (!profiling) {
        f();
        return;
    }
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, stream_));
    f();
    HIPCHK(hipEventRecord(b, stream_));
    prof_pending.push_back({cat, {a, b}});
    prof[cat].launches += launches;
    prof[cat].tasks += tasks;
}

}  // namespace spaqr
