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

__device__ __forceinline__ float lane_xor(float x, int off) { return __shfl_xor(x, off, 64); }
__device__ __forceinline__ double lane_xor(double x, int off) { return __shfl_xor(x, off, 64); }

template <int CTRL>
__device__ __forceinline__ float dpp(float x) { return dpp_f32<CTRL>(x); }
template <int CTRL>
__device__ __forceinline__ double dpp(double x) { return dpp_f64<CTRL>(x); }

// Sum of a float or double over aligned groups of L lanes (L a power of two,
// 1..64, uniform over the wavefront); every lane returns its group's sum.
// Groups above 16 lanes finish with lane exchanges across rows.  A group's
// lanes must all be active (whole groups may be inactive).
template <typename T>
__device__ __forceinline__ T lane_sum(T v, int L) {
    if (L >= 2) v += dpp<0xB1>(v);
    if (L >= 4) v += dpp<0x4E>(v);
    if (L >= 8) v += dpp<0x141>(v);
    if (L >= 16) v += dpp<0x140>(v);
    if (L >= 32) v += lane_xor(v, 16);
    if (L >= 64) v += lane_xor(v, 32);
    return v;
}

template <int L, typename T>
__device__ __forceinline__ T lane_sum(T v) {
    static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "L: power of two up to 64");
    return lane_sum(v, L);
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

// Keeps loads issued ahead of the branch that follows: an empty asm taking
// their values on the path that otherwise does not use them, so the compiler
// cannot sink them past the branch (the other path waits for them only if it
// is taken).
template <typename T>
__device__ __forceinline__ void keep_one(const T &v) {
    asm volatile("" ::"v"(v));
}

template <typename... T>
__device__ __forceinline__ void keep_issued(const T &...v) {
    (keep_one(v), ...);
}

}  // namespace tv3
