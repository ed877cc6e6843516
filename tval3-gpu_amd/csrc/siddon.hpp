// siddon.hpp -- the Siddon ray traversal every Siddon product uses (the
// matrix-free products, the device CSR build and the per-tile products).
//
// Exact intersection lengths of the segment p0 -> p1 with a grid of
// N[0] x N[1] x N[2] voxels (axes x, y, z) of size h from corner org (Siddon,
// Med. Phys. 12(2), 1985), traversed in increasing ray parameter t:
//   plane k (0..N) of axis a is crossed at t(a, k) = (org + k h - p0) / d,
//   evaluated as fma(k, A, B) with A = h * (1 / d), B = (org - p0) * (1 / d)
//   (each rounded once; no other contraction): one fused operation per
//   plane, monotone in k;
//   the ray is clipped to [amin, amax] by planes 0 and N of every axis;
//   the walk keeps, per axis, the next plane with t > t_cur and the voxel
//   index it tracks, and emits each voxel it crosses with length
//   (t_next - t_cur) L, L = |p1 - p0|, when that length is positive.
// Two facts make a walk over part of a ray exact:
//   * the walk's state after any event t is a function of t alone (per axis
//     the first plane beyond t), so a walk started at the entry of a tile
//     (an 8^3 block, whose faces are planes 8j) emits exactly the entries the
//     walk over the whole ray emits inside that tile;
//   * the segment of a voxel is its box intersection: t_cur is the largest
//     entry-plane parameter of the voxel (or amin), t_next the smallest
//     exit-plane parameter (or amax), so siddon_box(ray, voxel) reproduces
//     the walk's length for that voxel bit for bit.
// The oracle restates this traversal in FP32 and FP64 (orc_siddon_walk);
// the textbook merge of sorted plane crossings (orc_siddon_ray) pins it.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace tv3 {

template <typename R>
struct SidRay {
    R A[3], B[3];    // plane k of axis a at t = fma(k, A, B); a parallel axis: A = 0, B = +inf
    R rA[3];         // 1 / A: seeds the plane search only
    R amin, amax, L;
    int N[3];
    int st[3];       // +1 / -1: direction of travel along the axis; 0: parallel
    int cfix[3];     // voxel index of a parallel axis
    bool hit;        // the clipped ray is not empty
};

__host__ __device__ __forceinline__ float sid_fma(float k, float a, float b) { return fmaf(k, a, b); }
__host__ __device__ __forceinline__ double sid_fma(double k, double a, double b) {
    return fma(k, a, b);
}

template <typename R>
__host__ __device__ __forceinline__ R sid_plane(const SidRay<R> &s, int a, int k) {
    return sid_fma((R)k, s.A[a], s.B[a]);
}

template <typename R>
__host__ __device__ __forceinline__ int sid_clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

template <typename R>
__host__ __device__ inline void sid_setup(SidRay<R> &s, const R p0[3], const R p1[3],
                                          const R org[3], const R h[3], const int N[3]) {
#pragma clang fp contract(off)
    s.amin = 0;
    s.amax = 1;
    s.hit = true;
    R d[3];
    for (int a = 0; a < 3; ++a) {
        s.N[a] = N[a];
        d[a] = p1[a] - p0[a];
        s.cfix[a] = 0;
        if (d[a] == 0) {
            s.st[a] = 0;
            s.A[a] = 0;
            s.B[a] = (R)INFINITY;
            s.rA[a] = 0;
            const R lo = org[a], hi = org[a] + (R)N[a] * h[a];
            if (!(p0[a] > lo && p0[a] < hi)) s.hit = false;
            s.cfix[a] = sid_clampi<R>((int)floor((p0[a] - org[a]) / h[a]), 0, N[a] - 1);
        } else {
            s.st[a] = d[a] > 0 ? 1 : -1;
            const R inv = (R)1 / d[a];
            s.A[a] = h[a] * inv;
            s.B[a] = (org[a] - p0[a]) * inv;
            s.rA[a] = (R)1 / s.A[a];
            const R t0 = sid_plane(s, a, 0), t1 = sid_plane(s, a, N[a]);
            const R lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
            s.amin = lo > s.amin ? lo : s.amin;
            s.amax = hi < s.amax ? hi : s.amax;
        }
    }
    if (!(s.amin < s.amax)) s.hit = false;
    s.L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

// First plane along the direction of travel of axis a crossed after t: the
// smallest k with t(a, k) > t (st = +1) or the largest (st = -1).  The
// position estimate only seeds the search; the result is decided by the
// plane parameters themselves.
template <typename R>
__host__ __device__ inline int sid_next_plane(const SidRay<R> &s, int a, R t) {
#pragma clang fp contract(off)
    const int N = s.N[a];
    const R f = (t - s.B[a]) * s.rA[a];
    if (s.st[a] > 0) {
        int k = sid_clampi<R>((int)floor(f) + 1, 1, N);
        while (k > 1 && sid_plane(s, a, k - 1) > t) --k;
        while (k < N && !(sid_plane(s, a, k) > t)) ++k;
        return k;
    }
    int k = sid_clampi<R>((int)ceil(f) - 1, 0, N - 1);
    while (k < N - 1 && sid_plane(s, a, k + 1) > t) ++k;
    while (k > 0 && !(sid_plane(s, a, k) > t)) --k;
    return k;
}

// Walk [ts, te] (ts an event of the traversal: amin or a plane crossing,
// te = amax or a plane crossing); emit(ix, iy, iz, length) per voxel.
template <typename R, typename F>
__host__ __device__ inline void sid_walk(const SidRay<R> &s, R ts, R te, F emit) {
#pragma clang fp contract(off)
    const R INF = (R)1e30;
    int kn[3], c[3];
    R tk[3];
    for (int a = 0; a < 3; ++a) {
        if (s.st[a] == 0) {
            kn[a] = 0;
            c[a] = s.cfix[a];
            tk[a] = INF;
            continue;
        }
        kn[a] = sid_next_plane(s, a, ts);
        c[a] = s.st[a] > 0 ? kn[a] - 1 : kn[a];
        tk[a] = sid_plane(s, a, kn[a]);
    }
    R tcur = ts;
    const int maxsteps = s.N[0] + s.N[1] + s.N[2] + 4;
    for (int step = 0; step < maxsteps; ++step) {
        R tnext = te;
        for (int a = 0; a < 3; ++a) tnext = tk[a] < tnext ? tk[a] : tnext;
        const R seg = (tnext - tcur) * s.L;
        if (seg > 0) emit(c[0], c[1], c[2], seg);
        if (!(tnext < te)) break;
        for (int a = 0; a < 3; ++a) {
            if (tk[a] == tnext) {
                c[a] += s.st[a];
                kn[a] += s.st[a];
                tk[a] = (kn[a] < 0 || kn[a] > s.N[a]) ? INF : sid_plane(s, a, kn[a]);
            }
        }
        tcur = tnext;
    }
}

// Length of the ray inside voxel (cx, cy, cz): the walk's segment of that
// voxel, bit for bit; 0 when the ray misses it.
template <typename R>
__host__ __device__ inline R sid_box(const SidRay<R> &s, const int c[3]) {
#pragma clang fp contract(off)
    R lo = s.amin, hi = s.amax;
    for (int a = 0; a < 3; ++a) {
        if (s.st[a] == 0) {
            if (c[a] != s.cfix[a]) return 0;
            continue;
        }
        const int ke = s.st[a] > 0 ? c[a] : c[a] + 1, kx = s.st[a] > 0 ? c[a] + 1 : c[a];
        const R te = sid_plane(s, a, ke), tx = sid_plane(s, a, kx);
        lo = te > lo ? te : lo;
        hi = tx < hi ? tx : hi;
    }
    const R w = (hi - lo) * s.L;
    return w > 0 ? w : (R)0;
}

// Parameter range of the ray inside the block of voxels [b0, b1) (per axis):
// [ts, te] with ts < te when the ray crosses it (the walk over [ts, te]
// emits exactly the whole-ray entries inside the block).
template <typename R>
__host__ __device__ inline bool sid_block_range(const SidRay<R> &s, const int b0[3],
                                                const int b1[3], R &ts, R &te) {
#pragma clang fp contract(off)
    ts = s.amin;
    te = s.amax;
    for (int a = 0; a < 3; ++a) {
        if (s.st[a] == 0) {
            if (s.cfix[a] < b0[a] || s.cfix[a] >= b1[a]) return false;
            continue;
        }
        const R t0 = sid_plane(s, a, b0[a]), t1 = sid_plane(s, a, b1[a]);
        const R lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
        ts = lo > ts ? lo : ts;
        te = hi < te ? hi : te;
    }
    return ts < te;
}

}  // namespace tv3
