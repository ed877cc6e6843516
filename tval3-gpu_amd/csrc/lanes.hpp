// lanes.hpp -- sums across the lanes of a wavefront without LDS traffic.
//
// DPP moves within rows of 16 lanes: swap neighbours (quad_perm 1,0,3,2),
// swap pairs (quad_perm 2,3,0,1), mirror within 8 lanes, mirror within 16.
// Applied in that order to a sum, each step leaves in every lane the sum of
// its aligned group of 2, 4, 8, 16 lanes (a + b and b + a round alike, so
// the lanes of a group agree bit for bit).  Every lane taking part must be
// active.  The orders are fixed, so the sums are deterministic.
#pragma once

#include <hip/hip_runtime.h>

namespace tv3 {

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double lane_f64(double x, int lane) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum of a float over aligned groups of L lanes (L a power of two, 1..64);
// every lane returns its group's sum.  Groups above 16 lanes finish with
// lane exchanges across rows.
template <int L>
__device__ __forceinline__ float group_sum_f32(float v) {
    static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "L: power of two up to 64");
    if constexpr (L >= 2) v += dpp_f32<0xB1>(v);
    if constexpr (L >= 4) v += dpp_f32<0x4E>(v);
    if constexpr (L >= 8) v += dpp_f32<0x141>(v);
    if constexpr (L >= 16) v += dpp_f32<0x140>(v);
    if constexpr (L >= 32) v += __shfl_xor(v, 16, 64);
    if constexpr (L >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

// Sum of a double over the 64 lanes of a wavefront: the four row sums of
// the DPP steps, then (r0 + r1) + (r2 + r3).  Every lane returns the total.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    return (lane_f64(v, 0) + lane_f64(v, 16)) + (lane_f64(v, 32) + lane_f64(v, 48));
}

}  // namespace tv3
