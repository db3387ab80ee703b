// Cross-lane reductions for 64-wide gfx950 wavefronts built on DPP lane moves (ALU-speed)
// and readlane, instead of LDS permutes.
#pragma once
#include <hip/hip_runtime.h>

namespace spaqr {

// ---- cross-lane reductions on DPP (ALU-speed lane moves) instead of LDS permutes.
// dpp_ctrl codes: quad_perm [1,0,3,2] = 0xB1, quad_perm [2,3,0,1] = 0x4E,
// row_half_mirror = 0x141, row_mirror = 0x140.  All lanes of the wave must be active.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ float lane_read(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_read(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ int lane_read(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Sum over each row of 16 lanes; every lane of the row gets the same (bitwise) total.
template <typename S>
__device__ __forceinline__ S row16_sum(S v) {
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    v += dpp_mov<0x141>(v);
    v += dpp_mov<0x140>(v);
    return v;
}
template <typename S>
__device__ __forceinline__ S row16_max(S v) {
    v = max(v, dpp_mov<0xB1>(v));
    v = max(v, dpp_mov<0x4E>(v));
    v = max(v, dpp_mov<0x141>(v));
    v = max(v, dpp_mov<0x140>(v));
    return v;
}
template <typename S>
__device__ __forceinline__ S row16_min(S v) {
    v = min(v, dpp_mov<0xB1>(v));
    v = min(v, dpp_mov<0x4E>(v));
    v = min(v, dpp_mov<0x141>(v));
    v = min(v, dpp_mov<0x140>(v));
    return v;
}
// Sum over each group of 8 lanes (the same total in all 8).
template <typename S>
__device__ __forceinline__ S oct_sum(S v) {
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    v += dpp_mov<0x141>(v);
    return v;
}
// Whole-wave reductions; the result is uniform across the wave.
template <typename S>
__device__ __forceinline__ S wave_sum(S v) {
    v = row16_sum(v);
    return (lane_read(v, 0) + lane_read(v, 16)) + (lane_read(v, 32) + lane_read(v, 48));
}
template <typename S>
__device__ __forceinline__ S wave_max(S v) {
    v = row16_max(v);
    return max(max(lane_read(v, 0), lane_read(v, 16)), max(lane_read(v, 32), lane_read(v, 48)));
}
__device__ __forceinline__ int wave_min(int v) {
    v = row16_min(v);
    return min(min(lane_read(v, 0), lane_read(v, 16)), min(lane_read(v, 32), lane_read(v, 48)));
}

// Row-of-16 shift right by N lanes with zero fill (inclusive-scan building block).
template <int N>
__device__ __forceinline__ float dpp_shr(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x110 + N, 0xF, 0xF, true));
}
template <int N>
__device__ __forceinline__ double dpp_shr(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x110 + N, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x110 + N, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Inclusive prefix sum within each row of 16 lanes.
template <typename S>
__device__ __forceinline__ S row16_scan(S v) {
    v += dpp_shr<1>(v);
    v += dpp_shr<2>(v);
    v += dpp_shr<4>(v);
    v += dpp_shr<8>(v);
    return v;
}

// Orders one wave's LDS writes before its later LDS reads by other lanes, without a
// workgroup barrier: the wavefront-scope fences stop the compiler from moving LDS accesses
// across the point (the hardware executes one wave's LDS instructions in order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum the 32 per-lane values d[0..31] over the wave: a transposing butterfly.  On return
// lane L holds the total of column vec_col(L); lanes L and L ^ 32 hold the same column and
// vec_owner(L) (L < 32) picks one of them.  The four steps inside a 16-lane row pair lanes
// with DPP moves (row_mirror, row_half_mirror, quad_perm), which flip the step's bit plus
// only bits that are still to be reduced, so the partners always hold the same columns; the
// two cross-row steps (16, 32) use shuffles.  30 DPP moves and 2 shuffles in all.
__device__ __forceinline__ int vec_col(int lane) {
    return ((lane >> 3) & 1) * 16 + ((lane >> 2) & 1) * 8 + ((lane >> 1) & 1) * 4 + (lane & 1) * 2 +
           ((lane >> 4) & 1);
}
__device__ __forceinline__ bool vec_owner(int lane) { return lane < 32; }

template <int H, int BIT, int CTRL, typename S>
__device__ __forceinline__ void butterfly_dpp(S (&d)[32], int lane) {
    const bool up = ((lane >> BIT) & 1) != 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const S send = up ? d[i] : d[i + H];
        const S keep = up ? d[i + H] : d[i];
        d[i] = keep + dpp_mov<CTRL>(send);
    }
}

template <typename S>
__device__ __forceinline__ S wave_sum32(S (&d)[32]) {
    const int lane = threadIdx.x & 63;
    butterfly_dpp<16, 3, 0x140>(d, lane);  // row_mirror: flips bits 0-3
    butterfly_dpp<8, 2, 0x141>(d, lane);   // row_half_mirror: flips bits 0-2
    butterfly_dpp<4, 1, 0x4E>(d, lane);    // quad_perm [2,3,0,1]: xor 2
    butterfly_dpp<2, 0, 0xB1>(d, lane);    // quad_perm [1,0,3,2]: xor 1
    {
        const bool up = ((lane >> 4) & 1) != 0;
        const S send = up ? d[0] : d[1];
        const S keep = up ? d[1] : d[0];
        d[0] = keep + __shfl_xor(send, 16, 64);
    }
    return d[0] + __shfl_xor(d[0], 32, 64);
}

}  // namespace spaqr
