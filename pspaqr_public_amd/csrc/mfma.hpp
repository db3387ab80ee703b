// MFMA lane layout shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace spaqr {

typedef float floatx16 __attribute__((ext_vector_type(16)));
// v_mfma_f32_32x32x2f32: lane l supplies A(l % 32, l / 32) and B(l / 32, l % 32); accumulator
// entry r of lane l is D(mfma_row(l, r), l % 32).
__device__ __forceinline__ int mfma_row(int l, int r) { return (r >> 2) * 8 + (l >> 5) * 4 + (r & 3); }

}  // namespace spaqr
