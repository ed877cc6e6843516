// solver_batched.hip -- many TVAL3 problems of uneven size solved together.
//
// Every problem follows the fused solver's algorithm (the reference loop
// 3d/tval3_gpu_3d.cpp:1088-1345 with the quadratic line search and FP64
// device scalars, tval3_step.hpp); what is batched is the work: each kernel
// of an inner iteration covers every problem of the batch, so the number of
// launches per iteration does not grow with the batch, and K iterations are
// captured in one graph.
//
//  * Problems are concatenated, each in the fused solver's numbering (voxels
//    in 8^3 tiles of 4^3 bricks where its dims allow, rays in a locality
//    order); each problem's voxel range starts at a multiple of 512, its ray
//    range at a multiple of 256.
//    The system matrix is the block-diagonal concatenation of the problems'
//    renumbered CSR matrices (global voxel indices), its CSC likewise.
//  * A workgroup works inside one problem (256-element groups never cross a
//    range start); it finds its problem by binary search over the range
//    starts and returns at once when that problem is not active.
//  * Each problem has its control block (FCtl) and its own buffer parity: the
//    accepted U, g and Au alternate between two copies per problem, while the
//    trial vectors -- input and outputs of the products -- are single.
//  * Per-problem reductions: every workgroup stores its partial sums; one
//    workgroup per problem sums its problem's slice in a fixed order and runs
//    the scalar step of that problem.
//  * Products: the fused solver's kernels over the block-diagonal matrix --
//    row groups of L lanes per row, row blocks of NB entries staged in LDS,
//    or the panel forms (A^T per 512-voxel tile with the tile's rays in LDS;
//    A as segment sums per tile, then per ray), chosen by timing at set-up
//    -- whose row groups, blocks and panels skip inactive problems; update_W
//    runs on 8^3 tiles for the B = 8 problems; problems sharing a dense BF16
//    operator run as one matrix-core product with up to 64 right-hand sides.
//
// One inner iteration: descend (trial U_t, V statistics), A U_t, residual
// statistics, A^T A U_t - Atb, line search, apply alpha to Au, update_W
// (accepts U and g at alpha; d, BB sums), end of iteration -- a fixed number
// of launches for any batch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "panels.hpp"
#include "tiles.hpp"
#include "tval3_step.hpp"
#include "vec.hpp"

namespace tv3 {

namespace {

constexpr int GB = 256;   // elements per group = threads per workgroup

struct BDims {
    int32_t p, q, r;
    int32_t B, nby, nbx;  // voxel numbering of the problem (brick_index)
    int64_t v0, n;        // voxel range
    int64_t r0, m;        // ray range
};

// problem holding group g: starts[k] <= g < starts[k + 1] (starts[P] = end)
__device__ __forceinline__ int find_prob(const int32_t *__restrict__ starts, int P, int g) {
    int lo = 0, hi = P;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (starts[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// The groups of this workgroup (grp = blockIdx.x + i gridDim.x < ng) whose
// problem passes want(k), found in one parallel pass over up to 256
// candidates at a time (one thread each) instead of one dependent lookup
// per group; then body(grp, k) for each of them, in whatever order they
// were found (each group's partial sum has a slot of its own).  For
// 256-thread workgroups.
template <typename W, typename Bd>
__device__ __forceinline__ void for_active_groups(int ng, const int32_t *__restrict__ starts, int P,
                                                  W want, Bd body) {
    __shared__ int32_t lg[256], lk[256];
    __shared__ int32_t cnt;
    const int G = (int)gridDim.x;
    for (int base = 0; (int)blockIdx.x + base * G < ng; base += 256) {
        if (threadIdx.x == 0) cnt = 0;
        __syncthreads();
        const int grp = (int)blockIdx.x + (base + (int)threadIdx.x) * G;
        if (grp < ng) {
            const int k = find_prob(starts, P, grp);
            if (want(k)) {
                const int slot = atomicAdd(&cnt, 1);
                lg[slot] = grp;
                lk[slot] = k;
            }
        }
        __syncthreads();
        const int n = cnt;
        for (int i = 0; i < n; ++i) body(lg[i], lk[i]);
        __syncthreads();                 // lg, lk and cnt are rewritten by the next pass
    }
}

struct Vox {
    int y, x, z;
    int p, q, r, B, nby, nbx;
    int64_t v0, pq;
    __device__ __forceinline__ int64_t at(int yy, int xx, int zz) const {
        return v0 + brick_index(B, pq, p, nby, nbx, yy, xx, zz);
    }
};

// voxel (y, x, z) of internal index i of the batch (false: padding); the
// inverse of brick_index for the problem's numbering
__device__ __forceinline__ bool voxel_at(const BDims &dm, int64_t i, Vox &v) {
    const int64_t li = i - dm.v0;
    if (li >= dm.n) return false;
    const uint32_t l = (uint32_t)li;
    v.p = dm.p; v.q = dm.q; v.r = dm.r; v.B = dm.B; v.nby = dm.nby; v.nbx = dm.nbx;
    v.v0 = dm.v0; v.pq = (int64_t)dm.p * dm.q;
    if (dm.B == 8) {
        const uint32_t tile = l >> 9, sub = (l >> 6) & 7, ib = l & 63;
        const uint32_t ty = tile % (uint32_t)dm.nby, rest = tile / (uint32_t)dm.nby;
        const uint32_t tx = rest % (uint32_t)dm.nbx, tz = rest / (uint32_t)dm.nbx;
        v.y = (int)(ty * 8 + ((sub & 1) << 2) + (ib & 3));
        v.x = (int)(tx * 8 + (((sub >> 1) & 1) << 2) + ((ib >> 2) & 3));
        v.z = (int)(tz * 8 + ((sub >> 2) << 2) + (ib >> 4));
    } else if (dm.B == 4) {
        const uint32_t brick = l >> 6, ib = l & 63;
        const uint32_t by = brick % (uint32_t)dm.nby, rest = brick / (uint32_t)dm.nby;
        const uint32_t bx = rest % (uint32_t)dm.nbx, bz = rest / (uint32_t)dm.nbx;
        v.y = (int)(by * 4 + (ib & 3));
        v.x = (int)(bx * 4 + ((ib >> 2) & 3));
        v.z = (int)(bz * 4 + (ib >> 4));
    } else {
        const uint32_t p = (uint32_t)dm.p, q = (uint32_t)dm.q;
        const uint32_t yx = l % (p * q);
        v.y = (int)(yx % p); v.x = (int)(yx / p); v.z = (int)(l / (p * q));
    }
    return true;
}

__device__ __forceinline__ int nxt(int v, int n) { return v + 1 == n ? 0 : v + 1; }
__device__ __forceinline__ int prv(int v, int n) { return v == 0 ? n - 1 : v - 1; }

// Sum of part[k * nb + j] for j in [lo, hi), K arrays, in a fixed order
// (256 threads, strided partial sums, then a fixed tree); every thread gets
// the totals.
template <int K>
__device__ void seg_sums(const double *part, int nb, int lo, int hi, double *out) {
    __shared__ double sh[K][GB];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double *p = part + (int64_t)k * nb;
        double a0 = 0, a1 = 0;
        int j = lo + t;
        for (; j + GB < hi; j += 2 * GB) { a0 += p[j]; a1 += p[j + GB]; }
        if (j < hi) a0 += p[j];
        sh[k][t] = a0 + a1;
    }
    __syncthreads();
    for (int off = GB / 2; off > 0; off >>= 1) {
        if (t < off)
#pragma unroll
            for (int k = 0; k < K; ++k) sh[k][t] += sh[k][t + off];
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = sh[k][0];
    __syncthreads();   // sh is reused by the next call
}

enum Pred { P_ACTIVE = 0, P_TAU = 1, P_ALL = 2, P_OUTER = 3, P_MULT = 4 };

// P_MULT: active, and get_g's A^T-skip decision multiplies
__device__ __forceinline__ bool pred_on(const FCtl &c, int pred) {
    switch (pred) {
    case P_ACTIVE: return c.active != 0;
    case P_MULT: return c.active != 0 && c.mul_T != 0;
    case P_TAU: return c.need_tau != 0;
    case P_OUTER: return c.outer_flag != 0;
    default: return true;
    }
}

// ------------------------------------------------------------------ products
// Epilogue of the block-diagonal products (the fused solver's product
// kernels, spmv_stream.hpp): y = M x (- sub); a row block or group of a
// problem whose predicate is off is skipped (blocks never cross a problem's
// first row, which is a multiple of 256).
struct EpiBatch {
    float *y;
    const float *sub;
    const int32_t *starts;
    int P;
    const FCtl *C;
    int pred;
    const int32_t *gate = nullptr;   // device count; the launch is a no-op while it is 0
    struct Pre { float s; };
    __device__ bool skip() const { return gate && !*gate; }
    __device__ __forceinline__ bool skip_rows(int row) const {
        return !pred_on(C[find_prob(starts, P, row / GB)], pred);
    }
    __device__ __forceinline__ Pre load(int r) const { return {sub ? sub[r] : 0.f}; }
    __device__ __forceinline__ void row(int r, float s, bool active, const Pre &p) const {
        if (active) y[r] = sub ? s - p.s : s;
    }
    __device__ __forceinline__ void finish() const {}
};

// Pass 1 of A x in panel form over the sparse problems: segment partial sums;
// a panel whose first segment belongs to a problem whose predicate is off is
// skipped (a panel holds the segments of one tile, so of one problem).
struct EpiSegBatch {
    float *part;
    const int32_t *seg_starts;   // first segment of each sparse problem (+ total)
    int P;                       // sparse problems
    const FCtl *C;
    int pred;
    const int32_t *gate = nullptr;   // device count; the launch is a no-op while it is 0
    struct Pre {};
    __device__ bool skip() const { return gate && !*gate; }
    __device__ __forceinline__ bool skip_rows(int row) const {
        return !pred_on(C[find_prob(seg_starts, P, row)], pred);
    }
    __device__ __forceinline__ Pre load(int) const { return {}; }
    __device__ __forceinline__ void row(int r, float s, bool active, const Pre &) const {
        if (active) part[r] = s;
    }
    __device__ __forceinline__ void finish() const {}
};

// ------------------------------------------------------------------ stencils
// descend: U_t = max(U_c - tau d, 0) (:690-696); V0/V1 statistics.
__global__ __launch_bounds__(256) void k_b_descend(const BDims *__restrict__ D,
                                                   const int32_t *__restrict__ vst, int P,
                                                   const FCtl *__restrict__ C,
                                                   const float *__restrict__ U0,
                                                   const float *__restrict__ U1,
                                                   const float *__restrict__ d,
                                                   const float *__restrict__ sx,
                                                   const float *__restrict__ sy,
                                                   const float *__restrict__ sz,
                                                   const float *__restrict__ wsx,
                                                   const float *__restrict__ wsy,
                                                   const float *__restrict__ wsz,
                                                   float *__restrict__ Un, double *part) {
    const int k = find_prob(vst, P, blockIdx.x);
    const FCtl &c = C[k];
    if (!c.active) return;
    const BDims dm = D[k];
    const float *Uc = c.cur ? U1 : U0;
    double acc[3] = {0, 0, 0};
    const float tau = (float)c.tau, beta = (float)c.beta;
    const bool nn = c.nonneg, iso = c.iso, old = c.sigw_old;
    // W of the last update_W: old sigma at the old beta after an outer step
    // (3d/tval3_gpu_3d.cpp:484-492 reads the stored W)
    const float ib = 1.f / (old ? (float)c.beta_w : beta);
    const int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x;
    Vox v;
    if (voxel_at(dm, i, v)) {
        const int64_t iy = v.at(nxt(v.y, v.p), v.x, v.z), ix = v.at(v.y, nxt(v.x, v.q), v.z),
                      iz = v.at(v.y, v.x, nxt(v.z, v.r));
        const float u0 = Uc[i], uy = Uc[iy], ux = Uc[ix], uz = Uc[iz];
        auto step = [&](float u, float dv) {
            float t = u - tau * dv;
            return (nn && t < 0.f) ? 0.f : t;
        };
        const float t0 = step(u0, d[i]), ty = step(uy, d[iy]), tx = step(ux, d[ix]),
                    tz = step(uz, d[iz]);
        const float s_x = sx[i], s_y = sy[i], s_z = sz[i];
        float wx, wy, wz;
        shrink3(ux - u0, uy - u0, uz - u0, old ? wsx[i] : s_x, old ? wsy[i] : s_y,
                old ? wsz[i] : s_z, ib, iso, wx, wy, wz);
        const float v0x = (ux - u0) - wx, v0y = (uy - u0) - wy, v0z = (uz - u0) - wz;
        const float v1x = (tx - t0) - wx, v1y = (ty - t0) - wy, v1z = (tz - t0) - wz;
        acc[0] += (double)v1x * v1x + (double)v1y * v1y + (double)v1z * v1z;
        acc[1] += (double)v1x * s_x + (double)v1y * s_y + (double)v1z * s_z;
        acc[2] += (double)v0x * v1x + (double)v0y * v1y + (double)v0z * v1z;
        Un[i] = t0;
    }
    block_sum_store<3>(acc, part, gridDim.x);
}

// update_W + D^T + gradient + BB sums (:965-1029, :896-919, :1017-1067).
// same = 1: the new iterate is the current copy (set-up), else the other one.
// tiled = 1: problems of the B = 8 numbering are left to k_b_update_w_tile.
// Workgroup w covers voxel group gfirst + w; partial sums at that group's
// slot of nbV.
__global__ __launch_bounds__(256) void k_b_update_w(const BDims *__restrict__ D,
                                                    const int32_t *__restrict__ vst, int P,
                                                    const FCtl *__restrict__ C, int same,
                                                    int tiled, int gfirst, int nbV,
                                                    float *__restrict__ U0,
                                                    float *__restrict__ U1,
                                                    const float *__restrict__ T,
                                                    const float *__restrict__ sx,
                                                    const float *__restrict__ sy,
                                                    const float *__restrict__ sz,
                                                    float *__restrict__ g0,
                                                    float *__restrict__ g1,
                                                    const float *__restrict__ gT,
                                                    const float *__restrict__ dts,
                                                    float *__restrict__ d, double *part) {
    const int grp = gfirst + (int)blockIdx.x;
    const int k = find_prob(vst, P, grp);
    const FCtl &c = C[k];
    if (!c.active) return;
    const BDims dm = D[k];
    if (tiled && dm.B == 8) return;
    const float *Uc = c.cur ? U1 : U0;
    // the new iterate: the current copy (same) or the accepted step, read
    // from the trial T (blended with U_c when alpha != 1) and stored here
    const float *Ur = same ? Uc : T;
    const bool blend = !same && c.alpha != 1.0;
    const float al = (float)c.alpha;
    auto U = [&](int64_t j) {
        const float t = Ur[j];
        if (!blend) return t;
        const float u = Uc[j];
        return u + al * (t - u);
    };
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const float beta = (float)c.beta, ib = 1.f / beta, mdb = (float)c.muDbeta;
    const bool iso = c.iso;
    const int64_t i = (int64_t)grp * GB + threadIdx.x;
    Vox v;
    if (voxel_at(dm, i, v)) {
        const int y = v.y, x = v.x, z = v.z;
        const int yp = nxt(y, v.p), ym = prv(y, v.p), xp = nxt(x, v.q), xm = prv(x, v.q),
                  zp = nxt(z, v.r), zm = prv(z, v.r);
        const float u0 = U(i);
        const float uyp = U(v.at(yp, x, z)), uxp = U(v.at(y, xp, z)), uzp = U(v.at(y, x, zp));
        float wx, wy, wz;
        const float s_x = sx[i], s_y = sy[i], s_z = sz[i];
        const float l1 = shrink3(uxp - u0, uyp - u0, uzp - u0, s_x, s_y, s_z, ib, iso, wx,
                                 wy, wz);
        const float vx = (uxp - u0) - wx, vy = (uyp - u0) - wy, vz = (uzp - u0) - wz;
        float vyb, vxb, vzb;
        {
            const int64_t j = v.at(ym, x, z);
            const float uj = U(j);
            const float dyj = u0 - uj, dxj = U(v.at(ym, xp, z)) - uj, dzj = U(v.at(ym, x, zp)) - uj;
            float a, b, cc;
            shrink3(dxj, dyj, dzj, sx[j], sy[j], sz[j], ib, iso, a, b, cc);
            vyb = dyj - b;
        }
        {
            const int64_t j = v.at(y, xm, z);
            const float uj = U(j);
            const float dxj = u0 - uj, dyj = U(v.at(yp, xm, z)) - uj, dzj = U(v.at(y, xm, zp)) - uj;
            float a, b, cc;
            shrink3(dxj, dyj, dzj, sx[j], sy[j], sz[j], ib, iso, a, b, cc);
            vxb = dxj - a;
        }
        {
            const int64_t j = v.at(y, x, zm);
            const float uj = U(j);
            const float dzj = u0 - uj, dxj = U(v.at(y, xp, zm)) - uj, dyj = U(v.at(yp, x, zm)) - uj;
            float a, b, cc;
            shrink3(dxj, dyj, dzj, sx[j], sy[j], sz[j], ib, iso, a, b, cc);
            vzb = dzj - cc;
        }
        const float g2 = vyb - vy + vxb - vx + vzb - vz;            // D^T V, :501
        float gv;
        if (same) {
            gv = (c.cur ? g1 : g0)[i];
        } else {
            gv = gT[i];
            if (blend) {
                const float gc = (c.cur ? g1 : g0)[i];
                gv = gc + al * (gv - gc);
            }
            (c.cur ? U0 : U1)[i] = u0;
            (c.cur ? g0 : g1)[i] = gv;
        }
        const float dn = mdb * gv + g2 + dts[i];                     // :915-918
        const float uc = Uc[i], du = u0 - uc;
        acc[0] += l1;
        acc[1] += (double)vx * vx + (double)vy * vy + (double)vz * vz;
        acc[2] += (double)vx * s_x + (double)vy * s_y + (double)vz * s_z;
        acc[3] += (double)du * du;
        acc[4] += (double)du * (double)(dn - d[i]);
        acc[5] += (double)uc * uc;
        acc[6] += (double)dn * dn;
        d[i] = dn;
    }
    block_sum_store_at<7>(acc, part, nbV, grp);
}

// k_b_update_w on one 8x8x8 tile per workgroup for the problems of the B = 8
// numbering (whose voxel ranges start at multiples of 512): U staged in LDS
// over the tile and a one-voxel shell, V = DU - W computed once for every
// voxel of the tile and of its three lower faces (where D^T reads it), D^T V
// from LDS.  Per voxel the arithmetic is k_b_update_w's, so d is the same;
// the tile's sums go to the first of its two 256-voxel group slots and the
// second gets zeros, so the per-problem sums read them like k_b_update_w's.
__global__ __launch_bounds__(512) void k_b_update_w_tile(const BDims *__restrict__ D,
                                                         const int32_t *__restrict__ vst, int P,
                                                         const FCtl *__restrict__ C, int same,
                                                         float *__restrict__ U0,
                                                         float *__restrict__ U1,
                                                         const float *__restrict__ Ut,
                                                         const float *__restrict__ sx,
                                                         const float *__restrict__ sy,
                                                         const float *__restrict__ sz,
                                                         float *__restrict__ g0,
                                                         float *__restrict__ g1,
                                                         const float *__restrict__ gT,
                                                         const float *__restrict__ dts,
                                                         float *__restrict__ d, double *part,
                                                         int nbV) {
    constexpr int NBOX = kTileBox * kTileBox * kTileBox;
    __shared__ float su[NBOX], svx[NBOX], svy[NBOX], svz[NBOX];
    const int T = blockIdx.x;                       // voxels 512 T .. 512 T + 511
    const int k = find_prob(vst, P, 2 * T);
    const FCtl &c = C[k];
    if (!c.active) return;
    const BDims dm = D[k];
    if (dm.B != 8) return;
    // the new iterate: as in k_b_update_w (same: the current copy, else the
    // trial blended at alpha, stored in the other copy)
    const float *Uc = c.cur ? U1 : U0;
    const float *Ur = same ? Uc : Ut;
    const bool blend = !same && c.alpha != 1.0;
    const float al = (float)c.alpha;
    const float beta = (float)c.beta, ib = 1.f / beta, mdb = (float)c.muDbeta;
    const bool iso = c.iso;
    const int t = threadIdx.x;
    const float *Uk = Ur + dm.v0;                    // the problem's volume
    const int lt = T - (int)(dm.v0 >> 9);            // tile of the problem
    const int nby = dm.nby, nbx = dm.nbx, nbz = dm.r >> 3;
    const int ty = lt % nby, rest = lt / nby, tx = rest % nbx, tz = rest / nbx;
    // box entries t and t + 512, sigma at the tile voxel t and (t < 192) at
    // the lower-face voxel 512 + t; indices clamped instead of guarded
    const int k1 = t + 512 < NBOX ? t + 512 : NBOX - 1;
    int a0, b0, c0;
    tile_pos(t, a0, b0, c0);
    const int q1 = t < 192 ? t : 191;
    const int f = q1 >> 6, j = q1 & 63, fu = j & 7, fv = j >> 3;
    const int a1 = f == 0 ? -1 : fu, b1 = f == 0 ? fu : (f == 1 ? -1 : fv),
              c1 = f == 2 ? -1 : fv;
    const int64_t j0 = tile_gidx(nby, nbx, nbz, ty, tx, tz, box_off(t));
    const int64_t j1 = tile_gidx(nby, nbx, nbz, ty, tx, tz, box_off(k1));
    float ub0 = Uk[j0], ub1 = Uk[j1];
    const int64_t i = ((int64_t)T << 9) + t;
    const int64_t i1 = dm.v0 + tile_gidx(nby, nbx, nbz, ty, tx, tz, tile_off(a1, b1, c1));
    const float s0[3] = {sx[i], sy[i], sz[i]}, s1[3] = {sx[i1], sy[i1], sz[i1]};
    const float dv = dts[i], ucv = Uc[i], dold = d[i];
    float gv = same ? (c.cur ? g1 : g0)[i] : gT[i];
    if (blend) {
        const float c0u = Uc[dm.v0 + j0], c1u = Uc[dm.v0 + j1];
        ub0 = c0u + al * (ub0 - c0u);
        ub1 = c1u + al * (ub1 - c1u);
        const float gc = (c.cur ? g1 : g0)[i];
        gv = gc + al * (gv - gc);
    }
    su[t] = ub0;
    if (t + 512 < NBOX) su[t + 512] = ub1;
    __syncthreads();
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    auto vat = [&](int a, int b, int cc, const float *s, bool tile_voxel) {
        const int sl = box_slot(a, b, cc);
        const float u0 = su[sl], uyp = su[sl + 1], uxp = su[sl + kTileBox],
                    uzp = su[sl + kTileBox * kTileBox];
        float wx, wy, wz;
        const float l1 = shrink3(uxp - u0, uyp - u0, uzp - u0, s[0], s[1], s[2], ib, iso,
                                 wx, wy, wz);
        const float vx = (uxp - u0) - wx, vy = (uyp - u0) - wy, vz = (uzp - u0) - wz;
        svx[sl] = vx;
        svy[sl] = vy;
        svz[sl] = vz;
        if (tile_voxel) {
            acc[0] += l1;
            acc[1] += (double)vx * vx + (double)vy * vy + (double)vz * vz;
            acc[2] += (double)vx * s[0] + (double)vy * s[1] + (double)vz * s[2];
        }
    };
    vat(a0, b0, c0, s0, true);
    if (t < 192) vat(a1, b1, c1, s1, false);
    __syncthreads();
    const int sl = box_slot(a0, b0, c0);
    const float g2 = svy[sl - 1] - svy[sl] + svx[sl - kTileBox] - svx[sl] +
                     svz[sl - kTileBox * kTileBox] - svz[sl];              // D^T V, :501
    const float dn = mdb * gv + g2 + dv;                                    // :915-918
    const float du = su[sl] - ucv;
    acc[3] += (double)du * du;
    acc[4] += (double)du * (double)(dn - dold);
    acc[5] += (double)ucv * ucv;
    acc[6] += (double)dn * dn;
    d[i] = dn;
    if (!same) {
        (c.cur ? U0 : U1)[i] = su[sl];
        (c.cur ? g0 : g1)[i] = gv;
    }
    block_sum_store_at<7, 8>(acc, part, nbV, 2 * T);
    if (t < 7) part[(int64_t)t * nbV + 2 * T + 1] = 0.0;
}

// |D d|^2 for the first-iteration step length (:587-596)
__global__ __launch_bounds__(256) void k_b_dd_norm(const BDims *__restrict__ D,
                                                   const int32_t *__restrict__ vst, int P,
                                                   const FCtl *__restrict__ C,
                                                   const float *__restrict__ d, double *part,
                                                   int ng, const int32_t *gate) {
    if (gate && !*gate) return;
    // ng groups (256 voxels or rays) over a fixed grid (a launch whose gate
    // is off ends after that many workgroups, not ng); the groups whose
    // problem takes this step are found in one pass
    for_active_groups(ng, vst, P, [&](int k) { return C[k].need_tau != 0; }, [&](int grp, int k) {
        const BDims dm = D[k];
        double acc[1] = {0};
        const int64_t i = (int64_t)grp * GB + threadIdx.x;
        Vox v;
        if (voxel_at(dm, i, v)) {
            const float d0 = d[i];
            const float a = d[v.at(nxt(v.y, v.p), v.x, v.z)] - d0,
                        b = d[v.at(v.y, nxt(v.x, v.q), v.z)] - d0,
                        c = d[v.at(v.y, v.x, nxt(v.z, v.r))] - d0;
            acc[0] += (double)a * a + (double)b * b + (double)c * c;
        }
        block_sum_store_at<1>(acc, part, ng, grp);
        __syncthreads();   // the block sum's LDS is rewritten by the next group
    });
}

// update_mlp, sigma part (:1036-1069): keeps the old sigma for the W of the
// next descend and returns lam4 = <sigma_new, V>
__global__ __launch_bounds__(256) void k_b_mlp_sigma(const BDims *__restrict__ D,
                                                     const int32_t *__restrict__ vst, int P,
                                                     const FCtl *__restrict__ C,
                                                     const float *__restrict__ U0,
                                                     const float *__restrict__ U1,
                                                     float *__restrict__ sx, float *__restrict__ sy,
                                                     float *__restrict__ sz, float *__restrict__ ox,
                                                     float *__restrict__ oy, float *__restrict__ oz,
                                                     double *part,
                                                     int ng, const int32_t *gate) {
    if (gate && !*gate) return;
    // ng groups (256 voxels or rays) over a fixed grid (a launch whose gate
    // is off ends after that many workgroups, not ng); the groups whose
    // problem takes this step are found in one pass
    for_active_groups(ng, vst, P, [&](int k) { return C[k].outer_flag != 0; }, [&](int grp, int k) {
        const FCtl &c = C[k];
        const BDims dm = D[k];
        const float *U = c.cur ? U1 : U0;
        double acc[1] = {0};
        const float beta = (float)c.beta, ib = 1.f / beta;
        const int64_t i = (int64_t)grp * GB + threadIdx.x;
        Vox v;
        if (voxel_at(dm, i, v)) {
            const float u0 = U[i];
            const float dy = U[v.at(nxt(v.y, v.p), v.x, v.z)] - u0,
                        dx = U[v.at(v.y, nxt(v.x, v.q), v.z)] - u0,
                        dz = U[v.at(v.y, v.x, nxt(v.z, v.r))] - u0;
            const float a = sx[i], b = sy[i], cc = sz[i];
            float wx, wy, wz;
            shrink3(dx, dy, dz, a, b, cc, ib, c.iso, wx, wy, wz);
            const float vx = dx - wx, vy = dy - wy, vz = dz - wz;
            const float na = a - beta * vx, nb = b - beta * vy, nc = cc - beta * vz;
            ox[i] = a; oy[i] = b; oz[i] = cc;
            sx[i] = na; sy[i] = nb; sz[i] = nc;
            acc[0] += (double)na * vx + (double)nb * vy + (double)nc * vz;
        }
        block_sum_store_at<1>(acc, part, ng, grp);
        __syncthreads();   // the block sum's LDS is rewritten by the next group
    });
}

// outer update of d and DtsAtd (:1239-1246) without a stored g2
__global__ __launch_bounds__(256) void k_b_outer_d(const BDims *__restrict__ D,
                                                   const int32_t *__restrict__ vst, int P,
                                                   const FCtl *__restrict__ C,
                                                   float *__restrict__ d,
                                                   const float *__restrict__ g0,
                                                   const float *__restrict__ g1,
                                                   float *__restrict__ dts, double *part,
                                                   int ng, const int32_t *gate) {
    if (gate && !*gate) return;
    // ng groups (256 voxels or rays) over a fixed grid (a launch whose gate
    // is off ends after that many workgroups, not ng); the groups whose
    // problem takes this step are found in one pass
    for_active_groups(ng, vst, P, [&](int k) { return C[k].outer_flag != 0; }, [&](int grp, int k) {
        const FCtl &c = C[k];
        const BDims dm = D[k];
        const float *g = c.cur ? g1 : g0;
        double acc[1] = {0};
        const float ratio = (float)c.ratio, mo = (float)c.mdb_old, mn = (float)c.muDbeta;
        const int64_t i = (int64_t)grp * GB + threadIdx.x;
        if (i - dm.v0 < dm.n) {
            const float dv = d[i], gv = g[i];
            const float g2 = dv - mo * gv - dts[i];
            const float nd = ratio * dv;
            const float dn = mn * gv + g2 + nd;
            dts[i] = nd;
            d[i] = dn;
            acc[0] += (double)dn * dn;
        }
        block_sum_store_at<1>(acc, part, ng, grp);
        __syncthreads();   // the block sum's LDS is rewritten by the next group
    });
}

// |U - rcdU|^2 then rcdU = U (:1208-1217)
__global__ __launch_bounds__(256) void k_b_relout(const BDims *__restrict__ D,
                                                  const int32_t *__restrict__ vst, int P,
                                                  const FCtl *__restrict__ C,
                                                  const float *__restrict__ U0,
                                                  const float *__restrict__ U1,
                                                  float *__restrict__ rcd, double *part,
                                                  int ng, const int32_t *gate) {
    if (gate && !*gate) return;
    // ng groups (256 voxels or rays) over a fixed grid (a launch whose gate
    // is off ends after that many workgroups, not ng); the groups whose
    // problem takes this step are found in one pass
    for_active_groups(ng, vst, P, [&](int k) { return C[k].outer_flag != 0; }, [&](int grp, int k) {
        const FCtl &c = C[k];
        const BDims dm = D[k];
        const float *U = c.cur ? U1 : U0;
        double acc[1] = {0};
        const int64_t i = (int64_t)grp * GB + threadIdx.x;
        if (i - dm.v0 < dm.n) {
            const float u = U[i], t = u - rcd[i];
            acc[0] += (double)t * t;
            rcd[i] = u;
        }
        block_sum_store_at<1>(acc, part, ng, grp);
        __syncthreads();   // the block sum's LDS is rewritten by the next group
    });
}

// ------------------------------------------------------------------ rays
// residual statistics of the trial: r1 = Au_t - b, r0 = Au_c - b:
// |r1|^2, <delta, r1>, <r0, r1>
__global__ __launch_bounds__(256) void k_b_res(const BDims *__restrict__ D,
                                               const int32_t *__restrict__ rst, int P,
                                               const FCtl *__restrict__ C, int pred,
                                               const float *__restrict__ AuT,
                                               const float *__restrict__ b,
                                               const float *__restrict__ delta,
                                               const float *__restrict__ Au0,
                                               const float *__restrict__ Au1, double *part) {
    const int k = find_prob(rst, P, blockIdx.x);
    const FCtl &c = C[k];
    if (!pred_on(c, pred)) return;
    const BDims dm = D[k];
    const float *Auc = c.cur ? Au1 : Au0;
    double acc[3] = {0, 0, 0};
    const int64_t r = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (r - dm.r0 < dm.m) {
        const float bv = b[r], r1 = AuT[r] - bv, r0 = Auc[r] - bv;
        acc[0] += (double)r1 * r1;
        acc[1] += (double)delta[r] * r1;
        acc[2] += (double)r0 * r1;
    }
    block_sum_store<3>(acc, part, gridDim.x);
}

// A^T-skip heuristic (:398-468), per problem: |Au_t - last Au|^2, |Au_t|^2
__global__ __launch_bounds__(256) void k_b_skip_stats(const BDims *__restrict__ D,
                                                      const int32_t *__restrict__ rst, int P,
                                                      const FCtl *__restrict__ C,
                                                      const float *__restrict__ AuT,
                                                      const float *__restrict__ last,
                                                      double *part) {
    const int k = find_prob(rst, P, blockIdx.x);
    if (!C[k].active) return;
    const BDims dm = D[k];
    double acc[2] = {0, 0};
    const int64_t r = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (r - dm.r0 < dm.m) {
        const float a = AuT[r], dl = a - last[r];
        acc[0] = (double)dl * dl;
        acc[1] = (double)a * a;
    }
    block_sum_store<2>(acc, part, gridDim.x);
}

// after the gated A^T product: a problem that multiplied keeps Au_t as its
// last Au; one that skipped keeps its gradient (g_t = g_c), as get_g does
// (voxel groups first, then ray groups)
__global__ __launch_bounds__(256) void k_b_skip_after(const BDims *__restrict__ D,
                                                      const int32_t *__restrict__ vst,
                                                      const int32_t *__restrict__ rst, int P,
                                                      int nbV, const FCtl *__restrict__ C,
                                                      const float *__restrict__ g0,
                                                      const float *__restrict__ g1,
                                                      float *__restrict__ gT,
                                                      const float *__restrict__ AuT,
                                                      float *__restrict__ last) {
    const bool vox = (int)blockIdx.x < nbV;
    const int w = vox ? (int)blockIdx.x : (int)blockIdx.x - nbV;
    const int k = find_prob(vox ? vst : rst, P, w);
    const FCtl &c = C[k];
    if (!c.active) return;
    const BDims dm = D[k];
    const int64_t i = (int64_t)w * GB + threadIdx.x;
    if (vox) {
        if (!c.mul_T && i - dm.v0 < dm.n) gT[i] = (c.cur ? g1 : g0)[i];
    } else {
        if (c.mul_T && i - dm.r0 < dm.m) last[i] = AuT[i];
    }
}

// |A d|^2 (first-iteration step length)
__global__ __launch_bounds__(256) void k_b_sqnorm(const BDims *__restrict__ D,
                                                  const int32_t *__restrict__ rst, int P,
                                                  const FCtl *__restrict__ C,
                                                  const float *__restrict__ y, double *part,
                                                  int ng, const int32_t *gate) {
    if (gate && !*gate) return;
    // ng groups (256 voxels or rays) over a fixed grid (a launch whose gate
    // is off ends after that many workgroups, not ng); the groups whose
    // problem takes this step are found in one pass
    for_active_groups(ng, rst, P, [&](int k) { return C[k].need_tau != 0; }, [&](int grp, int k) {
        const BDims dm = D[k];
        double acc[1] = {0};
        const int64_t r = (int64_t)grp * GB + threadIdx.x;
        if (r - dm.r0 < dm.m) acc[0] = (double)y[r] * y[r];
        block_sum_store_at<1>(acc, part, ng, grp);
        __syncthreads();   // the block sum's LDS is rewritten by the next group
    });
}

__global__ __launch_bounds__(256) void k_b_mlp_delta(const BDims *__restrict__ D,
                                                     const int32_t *__restrict__ rst, int P,
                                                     const FCtl *__restrict__ C,
                                                     const float *__restrict__ Au0,
                                                     const float *__restrict__ Au1,
                                                     const float *__restrict__ b,
                                                     float *__restrict__ delta, double *part,
                                                     int ng, const int32_t *gate) {
    if (gate && !*gate) return;
    // ng groups (256 voxels or rays) over a fixed grid (a launch whose gate
    // is off ends after that many workgroups, not ng); the groups whose
    // problem takes this step are found in one pass
    for_active_groups(ng, rst, P, [&](int k) { return C[k].outer_flag != 0; }, [&](int grp, int k) {
        const FCtl &c = C[k];
        const BDims dm = D[k];
        const float *Au = c.cur ? Au1 : Au0;
        double acc[1] = {0};
        const float mu = (float)c.mu;
        const int64_t r = (int64_t)grp * GB + threadIdx.x;
        if (r - dm.r0 < dm.m) {
            const float rv = Au[r] - b[r];
            const float dn = delta[r] - mu * rv;
            delta[r] = dn;
            acc[0] += (double)dn * rv;
        }
        block_sum_store_at<1>(acc, part, ng, grp);
        __syncthreads();   // the block sum's LDS is rewritten by the next group
    });
}

// accepted Au: Au_new = Au_c + alpha (Au_t - Au_c), Au_t itself when alpha = 1
// (ray groups; U and g are accepted inside update_W)
__global__ __launch_bounds__(256) void k_b_apply(const BDims *__restrict__ D,
                                                 const int32_t *__restrict__ rst, int P,
                                                 const FCtl *__restrict__ C,
                                                 float *__restrict__ Au0, float *__restrict__ Au1,
                                                 const float *__restrict__ AuT) {
    const int w = (int)blockIdx.x;
    const int k = find_prob(rst, P, w);
    const FCtl &c = C[k];
    if (!c.active) return;
    const BDims dm = D[k];
    const int64_t i = (int64_t)w * GB + threadIdx.x;
    if (i - dm.r0 >= dm.m) return;
    const float *Ac = c.cur ? Au1 : Au0;
    float *An = c.cur ? Au0 : Au1;
    const float t = AuT[i];
    if (c.alpha == 1.0) {
        An[i] = t;
    } else {
        const float a = (float)c.alpha, v = Ac[i];
        An[i] = v + a * (t - v);
    }
}

// ------------------------------------------------------------------ per-problem scalar kernels
// one 256-thread workgroup per problem
__global__ __launch_bounds__(256) void k_b_search(FCtl *C, const int32_t *__restrict__ vst,
                                                  const int32_t *__restrict__ rst, int nbV,
                                                  int nbR, const double *pV, const double *pR) {
    FCtl *c = C + blockIdx.x;
    if (!c->active) return;
    double a[3], b[3];
    seg_sums<3>(pV, nbV, vst[blockIdx.x], vst[blockIdx.x + 1], a);
    seg_sums<3>(pR, nbR, rst[blockIdx.x], rst[blockIdx.x + 1], b);
    if (threadIdx.x == 0) line_search(c, a, b, c->mode);
}

__global__ __launch_bounds__(256) void k_b_iter_end(FCtl *C, const int32_t *__restrict__ vst,
                                                    int nbV, const double *pW, double *hist,
                                                    int64_t hist_stride) {
    FCtl *c = C + blockIdx.x;
    if (!c->active) return;
    double a[7];
    seg_sums<7>(pW, nbV, vst[blockIdx.x], vst[blockIdx.x + 1], a);
    if (threadIdx.x == 0) {
        iter_end(c, a, hist + (int64_t)blockIdx.x * hist_stride);
        c->cur ^= 1;
    }
}

// the skip decision of each active problem; forced (a product) on the
// fallback step (mode 1)
__global__ __launch_bounds__(256) void k_b_skip_decide(FCtl *C, const int32_t *__restrict__ rst,
                                                       int nbR, const double *pK) {
    FCtl *c = C + blockIdx.x;
    if (!c->active) return;
    double a[2];
    seg_sums<2>(pK, nbR, rst[blockIdx.x], rst[blockIdx.x + 1], a);
    if (threadIdx.x == 0) skip_decide(c, a[0], a[1], c->mode);
}

__global__ __launch_bounds__(256) void k_b_tau(FCtl *C, const int32_t *__restrict__ vst,
                                               const int32_t *__restrict__ rst, int nbV, int nbR,
                                               const double *pD, const double *pA,
                                               const int32_t *gate) {
    if (gate && !*gate) return;
    FCtl *c = C + blockIdx.x;
    if (!c->need_tau) return;
    double a[1], b[1];
    seg_sums<1>(pD, nbV, vst[blockIdx.x], vst[blockIdx.x + 1], a);
    seg_sums<1>(pA, nbR, rst[blockIdx.x], rst[blockIdx.x + 1], b);
    if (threadIdx.x == 0) tau_first_step(c, a[0], b[0], c->need_fallback);
}

__global__ __launch_bounds__(256) void k_b_init(FCtl *C, const int32_t *__restrict__ vst,
                                                const int32_t *__restrict__ rst, int nbV, int nbR,
                                                const double *pN, const double *pW) {
    FCtl *c = C + blockIdx.x;
    double a[3], w[7];
    seg_sums<3>(pN, nbR, rst[blockIdx.x], rst[blockIdx.x + 1], a);
    seg_sums<7>(pW, nbV, vst[blockIdx.x], vst[blockIdx.x + 1], w);
    if (threadIdx.x == 0) init_step(c, a, w);
}

// relative change of U over the inner loop; below tol the problem stops
// (:1210-1218)
__global__ __launch_bounds__(256) void k_b_relout_scalar(FCtl *C, const int32_t *__restrict__ vst,
                                                         int nbV, const double *p, double tol,
                                                         const int32_t *gate) {
    if (gate && !*gate) return;
    FCtl *c = C + blockIdx.x;
    if (!c->outer_flag) return;
    double a[1];
    seg_sums<1>(p, nbV, vst[blockIdx.x], vst[blockIdx.x + 1], a);
    if (threadIdx.x == 0) {
        c->relout = sqrt(a[0]) / sqrt(c->nrmup2);
        if (c->relout < tol) {
            c->outer_flag = 0;
            c->done = 1;
        }
    }
}

__global__ __launch_bounds__(256) void k_b_outer_scalars(FCtl *C, const int32_t *__restrict__ vst,
                                                         const int32_t *__restrict__ rst, int nbV,
                                                         int nbR, const double *p4,
                                                         const double *p5,
                                                         const int32_t *gate) {
    if (gate && !*gate) return;
    FCtl *c = C + blockIdx.x;
    if (!c->outer_flag) return;
    double a[1], b[1];
    seg_sums<1>(p4, nbV, vst[blockIdx.x], vst[blockIdx.x + 1], a);
    seg_sums<1>(p5, nbR, rst[blockIdx.x], rst[blockIdx.x + 1], b);
    if (threadIdx.x == 0) outer_step(c, a[0], b[0]);
}

// end of the outer update: d.d for the next inner loop, then the loop
// condition of :1166 (outer iterations and total inner iterations)
__global__ __launch_bounds__(256) void k_b_set_dd(FCtl *C, const int32_t *__restrict__ vst,
                                                  int nbV, const double *p, int maxcnt, int maxit,
                                                  const int32_t *gate) {
    if (gate && !*gate) return;
    FCtl *c = C + blockIdx.x;
    if (!c->outer_flag) return;
    double a[1];
    seg_sums<1>(p, nbV, vst[blockIdx.x], vst[blockIdx.x + 1], a);
    if (threadIdx.x == 0) {
        c->dd = a[0];
        c->outer_flag = 0;
        if (c->outer >= maxcnt || c->total >= maxit) {
            c->done = 1;
            c->active = 0;
            c->need_tau = 0;
        }
    }
}

// ---------------------------------------------------------- rounds on the device
// A round advances every unfinished problem by one step of its own loop:
// the outer update when its inner loop has ended, the first step of an inner
// loop (or of a fallback) when it needs one, then an inner iteration.  The
// counts below gate the launches of the rarer steps (a kernel whose gate is
// 0 returns at once); per problem, each kernel still acts only on the
// problems whose own flags ask for it.  Gate layout: [0] problems in their
// outer update, [1] problems needing a first step, [2] unfinished problems,
// [3 + g] problems of dense group g needing a first step.
constexpr int kGateOuter = 0, kGateTau = 1, kGateAlive = 2, kGateGroup = 3;

// classification at the start of a round (the host loop of :1166-1220 per
// problem): an inner loop that has ended starts the outer update, a
// non-finite objective stops the problem
__global__ __launch_bounds__(256) void k_b_classify(FCtl *C, int P, int32_t *gates) {
    __shared__ int32_t cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int mine = 0;
    for (int k = threadIdx.x; k < P; k += blockDim.x) {
        FCtl &c = C[k];
        if (c.done || c.need_tau || c.active) continue;
        if (c.nonfinite) {
            c.done = 1;
            continue;
        }
        c.outer += 1;
        c.outer_flag = 1;
        ++mine;
    }
    if (mine) atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) gates[kGateOuter] = cnt;
}

// after the outer update: problems, and problems of each dense group, that
// need a first step (grp: first problem of each group, and the end)
__global__ __launch_bounds__(256) void k_b_gate_tau(const FCtl *C, int P,
                                                    const int32_t *__restrict__ grp, int ngroups,
                                                    int32_t *gates) {
    __shared__ int32_t cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int mine = 0;
    for (int k = threadIdx.x; k < P; k += blockDim.x) mine += C[k].need_tau != 0;
    if (mine) atomicAdd(&cnt, mine);
    for (int g = threadIdx.x; g < ngroups; g += blockDim.x) {
        int any = 0;
        for (int k = grp[g]; k < grp[g + 1]; ++k) any |= C[k].need_tau != 0;
        gates[kGateGroup + g] = any;
    }
    __syncthreads();
    if (threadIdx.x == 0) gates[kGateTau] = cnt;
}

// unfinished problems after the last round of a launch
__global__ __launch_bounds__(256) void k_b_alive(const FCtl *C, int P, int32_t *gates) {
    __shared__ int32_t cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int mine = 0;
    for (int k = threadIdx.x; k < P; k += blockDim.x) mine += C[k].done == 0;
    if (mine) atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) gates[kGateAlive] = cnt;
}

// ------------------------------------------------------------------ set-up kernels
__global__ void k_concat_ptr(const int32_t *__restrict__ ptr, int64_t rows, int64_t padded,
                             int64_t off, int32_t *__restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j > padded) return;
    out[j] = (int32_t)(off + (j <= rows ? ptr[j] : ptr[rows]));
}

__global__ void k_concat_ind(const int32_t *__restrict__ ind, int64_t nnz, int64_t off,
                             int32_t *__restrict__ out) {
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < nnz;
         j += (int64_t)gridDim.x * 256)
        out[j] = (int32_t)(ind[j] + off);
}

// out[i] = s in[map[i]] (natural -> internal order, scaled; s in device memory)
__global__ void k_gather_scale(const float *__restrict__ in, const int32_t *__restrict__ map,
                               int64_t n, const float *__restrict__ sp, float *__restrict__ out) {
    const float s = *sp;
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256)
        out[j] = s * in[map ? map[j] : j];
}

// out[map[i]] = in[i] / s (internal -> natural order, unscaled, :1275-1277)
__global__ void k_scatter_unscale(const float *__restrict__ in, const int32_t *__restrict__ map,
                                  int64_t n, const float *__restrict__ sp,
                                  float *__restrict__ out) {
    const float s = *sp;
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256)
        out[map ? map[j] : j] = in[j] / s;
}

// scaleb (:225-266), one workgroup per problem: the entries of least and of
// largest magnitude (the first such in index order, as a sequential scan
// finds them), their distance dif, the factor 0.5 / dif below 0.5 and
// 1.5 / dif above 1.5; *bad set when b holds NaN or Inf
__global__ __launch_bounds__(1024) void k_b_scaleb(const float *const *__restrict__ bptr,
                                                   const BDims *__restrict__ D,
                                                   float *__restrict__ scl,
                                                   int32_t *__restrict__ bad) {
    __shared__ float sa[1024], sb[1024];
    __shared__ int64_t ia[1024], ib[1024];
    __shared__ int32_t snf;
    const int k = blockIdx.x, t = threadIdx.x;
    const float *b = bptr[k];
    const int64_t m = D[k].m;
    float amin = INFINITY, amax = -1.f;
    int64_t imin = INT64_MAX, imax = INT64_MAX;
    int nf = 0;
    for (int64_t j = t; j < m; j += 1024) {
        const float v = b[j], a = fabsf(v);
        nf |= !isfinite(v);
        if (a < amin) { amin = a; imin = j; }
        if (a > amax) { amax = a; imax = j; }
    }
    if (t == 0) snf = 0;
    sa[t] = amin; ia[t] = imin; sb[t] = amax; ib[t] = imax;
    __syncthreads();
    if (nf) atomicOr(&snf, 1);
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) {
            if (sa[t + w] < sa[t] || (sa[t + w] == sa[t] && ia[t + w] < ia[t])) {
                sa[t] = sa[t + w];
                ia[t] = ia[t + w];
            }
            if (sb[t + w] > sb[t] || (sb[t + w] == sb[t] && ib[t + w] < ib[t])) {
                sb[t] = sb[t + w];
                ib[t] = ib[t + w];
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        float s = 1.f;
        if (snf) {
            atomicOr(bad, 1);
        } else if (m > 0) {
            const float dif = fabsf(b[ib[0]] - b[ia[0]]);
            if (dif < 0.5f) s = 0.5f / dif;
            else if (dif > 1.5f) s = 1.5f / dif;
        }
        scl[k] = s;
    }
}

inline int64_t pad256(int64_t v) { return (v + GB - 1) / GB * GB; }
// voxel ranges: multiples of 512, so 8^3 tiles and 512-row panels never
// straddle two problems (a problem solved alone gets the same padding)
inline int64_t pad512(int64_t v) { return (v + 511) / 512 * 512; }

inline unsigned grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}

// ------------------------------------------------------------------ host side
struct Batch {
    tv3_ctx *ctx;
    hipStream_t s;
    const tv3_opts *o;
    int P = 0;
    std::vector<BDims> dims;
    int64_t N = 0, M = 0;            // padded totals
    int nbV = 0, nbR = 0;            // groups
    tv3_dtype vt = TV3_F32;
    SparseRows A, AT;                // block-diagonal CSR of A and of A^T
    int lanes_N = 0, lanes_T = 0;
    bool tiles = false;              // update_W of B = 8 problems on 8^3 tiles
    bool untiled = true;             // problems left to the per-voxel update_W
    int gfirst = 0;                  // first voxel group of those problems
    Panels PT;                       // A^T in panel form (when panels_T != 0)
    int panels_T = 0;                // its row-block size, 0: A^T by the row kernels
    Panels PA;                       // A in two-pass panel form (when panels_A != 0)
    SegSum RS;
    DevBuf<int32_t> segst;           // first segment of each sparse problem (+ total)
    int panels_A = 0;
    int box_A = 0;                   // tiles per panel of that form: 1 or 8 (16^3 voxels)
    int ncsr = 0;                    // sparse problems (the first ncsr)
    DevBuf<BDims> ddims;
    DevBuf<int32_t> vst, rst;
    DevBuf<float> pool;
    float *U[2], *T, *g[2], *gT, *d, *sx, *sy, *sz, *ox, *oy, *oz, *dts, *Atb, *rcd;
    float *Au[2], *AuT, *b, *delta, *last;
    bool skipping = false;           // A^T-skip heuristic on (skipMulRatio != 0)
    DevBuf<double> parts, hist;
    double *pD, *pN, *pW, *pDd, *pAd, *p4, *p5, *pOD, *pRO, *pK;
    int64_t hist_stride = 1;
    DevBuf<FCtl> ctl;
    PinnedBuf<FCtl> hctl;
    hipGraphExec_t graph = nullptr;
    int graph_iters = 8;
    // problems sharing one dense BF16 operator: their vectors are consecutive,
    // so G of them form an n x G (m x G) column-major block and a product is
    // one matrix-core multi-RHS product
    struct DGroup {
        tv3_op *op;
        int k0, G;                 // internal problems k0 .. k0 + G - 1
        int64_t v0, r0, ldn, ldm;  // first ranges and per-problem strides
        int sN = 1, sT = 1;        // K splits of the two products
    };
    std::vector<DGroup> groups;
    DevBuf<float> gpart;           // split-K partials of the group products
    DevBuf<int32_t> grp;           // first problem of each dense group, then the end
    DevBuf<float> scl;             // per problem: the scaleb factor of the current solve
    DevBuf<const float *> bptr;    // per problem: the caller's b of the current solve
    DevBuf<int32_t> bad;           // the current b holds NaN or Inf
    PinnedBuf<int32_t> hbad;
    DevBuf<int32_t> gates;         // device counts gating a round's launches
    PinnedBuf<int32_t> halive;     // unfinished problems, read after each launch

    ~Batch() {
        if (graph) (void)hipGraphExecDestroy(graph);
    }

    void read_ctl() {
        TV3_HIP(hipMemcpyAsync(hctl.p, ctl.p, sizeof(FCtl) * P, hipMemcpyDeviceToHost, s));
        TV3_HIP(hipStreamSynchronize(s));
    }
    void write_ctl() {
        TV3_HIP(hipMemcpyAsync(ctl.p, hctl.p, sizeof(FCtl) * P, hipMemcpyHostToDevice, s));
    }

    void product(const SparseRows &S, const float *x, float *y, const float *sub,
                 const int32_t *starts, int pred, const int32_t *gate = nullptr) {
        launch_product<float>(S, x, EpiBatch{y, sub, starts, P, ctl.p, pred, gate}, s);
    }
    // dense groups: every column is computed (the outputs are scratch, read
    // only for the problems whose predicate holds)
    void group_N(const DGroup &g, const float *x, float *y, int splits,
                 const int32_t *gate = nullptr) {
        for (int c = 0; c < g.G; c += 64)
            gemm_rhs((const uint16_t *)g.op->dense_oN.p, g.op->m, g.op->n, x + g.v0 + c * g.ldn,
                     g.ldn, std::min(64, g.G - c), y + g.r0 + c * g.ldm, g.ldm, nullptr, splits,
                     gpart.p, s, gate);
    }
    void group_T(const DGroup &g, const float *x, float *y, const float *sub, int splits) {
        for (int c = 0; c < g.G; c += 64)
            gemm_rhs((const uint16_t *)g.op->dense_oT.p, g.op->n, g.op->m, x + g.r0 + c * g.ldm,
                     g.ldm, std::min(64, g.G - c), y + g.v0 + c * g.ldn, g.ldn,
                     sub ? sub + g.v0 + c * g.ldn : nullptr, splits, gpart.p, s);
    }
    // gate (optional): device count gating the sparse launches; ggate: one
    // per dense group
    void prod_N(const float *x, float *y, int pred, const int32_t *gate = nullptr,
                const int32_t *ggate = nullptr) {
        if (A.rows) {
            if (panels_A) {
                launch_panel(PA, x, EpiSegBatch{RS.part.p, segst.p, ncsr, ctl.p, pred, gate}, s);
                launch_seg_sum(RS, EpiBatch{y, nullptr, rst.p, P, ctl.p, pred, gate}, s);
            } else {
                product(A, x, y, nullptr, rst.p, pred, gate);
            }
        }
        for (size_t i = 0; i < groups.size(); ++i)
            group_N(groups[i], x, y, groups[i].sN, ggate ? ggate + i : nullptr);
    }
    void prod_T(const float *x, float *y, const float *sub, int pred) {
        if (AT.rows) {
            if (panels_T) launch_panel(PT, x, EpiBatch{y, sub, vst.p, P, ctl.p, pred}, s);
            else product(AT, x, y, sub, vst.p, pred);
        }
        for (const DGroup &g : groups) group_T(g, x, y, sub, g.sT);
    }

    // one inner iteration of every active problem
    void iteration() {
        const dim3 bV(nbV), bR(nbR), bP(P), t(256);
        hipLaunchKernelGGL(k_b_descend, bV, t, 0, s, ddims.p, vst.p, P, ctl.p, U[0], U[1], d, sx,
                           sy, sz, ox, oy, oz, T, pD);
        prod_N(T, AuT, P_ACTIVE);
        hipLaunchKernelGGL(k_b_res, bR, t, 0, s, ddims.p, rst.p, P, ctl.p, (int)P_ACTIVE, AuT, b,
                           delta, Au[0], Au[1], pN);
        if (skipping) {                // get_g's A^T-skip decision (:398-468)
            hipLaunchKernelGGL(k_b_skip_stats, bR, t, 0, s, ddims.p, rst.p, P, ctl.p, AuT, last,
                               pK);
            hipLaunchKernelGGL(k_b_skip_decide, bP, t, 0, s, ctl.p, rst.p, nbR, pK);
            prod_T(AuT, gT, Atb, P_MULT);
            hipLaunchKernelGGL(k_b_skip_after, dim3(nbV + nbR), t, 0, s, ddims.p, vst.p, rst.p, P,
                               nbV, ctl.p, g[0], g[1], gT, AuT, last);
        } else {
            prod_T(AuT, gT, Atb, P_ACTIVE);
        }
        hipLaunchKernelGGL(k_b_search, bP, t, 0, s, ctl.p, vst.p, rst.p, nbV, nbR, pD, pN);
        hipLaunchKernelGGL(k_b_apply, bR, t, 0, s, ddims.p, rst.p, P, ctl.p, Au[0], Au[1], AuT);
        update_w(0);
        hipLaunchKernelGGL(k_b_iter_end, bP, t, 0, s, ctl.p, vst.p, nbV, pW, hist.p,
                           hist_stride);
    }

    // update_W of every active problem (same = 1: of the current copy)
    void update_w(int same) {
        if (tiles)
            hipLaunchKernelGGL(k_b_update_w_tile, dim3(nbV / 2), dim3(512), 0, s, ddims.p, vst.p,
                               P, ctl.p, same, U[0], U[1], T, sx, sy, sz, g[0], g[1], gT, dts, d,
                               pW, nbV);
        if (untiled)
            hipLaunchKernelGGL(k_b_update_w, dim3(nbV - gfirst), dim3(256), 0, s, ddims.p,
                               vst.p, P, ctl.p, same, tiles ? 1 : 0, gfirst, nbV, U[0], U[1], T,
                               sx, sy, sz, g[0], g[1], gT, dts, d, pW);
    }

    // the tile kernel takes the B = 8 problems when all of them start at a
    // multiple of 512 voxels (tiles then never straddle two problems)
    void choose_tiles() {
        bool any8 = false, aligned = true, other = false;
        for (const BDims &bd : dims) {
            if (bd.B == 8) {
                any8 = true;
                aligned = aligned && bd.v0 % 512 == 0 && bd.n % 512 == 0;
            } else {
                other = true;
            }
        }
        tiles = any8 && aligned;
        untiled = !tiles || other;
        gfirst = 0;
        if (tiles && other) {
            gfirst = nbV;
            for (const BDims &bd : dims)
                if (bd.B != 8) gfirst = std::min(gfirst, (int)(bd.v0 / GB));
        }
    }

    // first-iteration step lengths of the problems that need one
    // first step of an inner loop or of a fallback (get_tau, :620-660) for
    // the problems that need one; gated by the device counts
    void tau_pass() {
        const dim3 bV(round_grid(nbV)), bR(round_grid(nbR)), t(256);
        const int32_t *gt = gates.p + kGateTau;
        hipLaunchKernelGGL(k_b_dd_norm, bV, t, 0, s, ddims.p, vst.p, P, ctl.p, d, pDd, nbV, gt);
        prod_N(d, AuT, P_TAU, gt, gates.p + kGateGroup);
        hipLaunchKernelGGL(k_b_sqnorm, bR, t, 0, s, ddims.p, rst.p, P, ctl.p, AuT, pAd, nbR, gt);
        hipLaunchKernelGGL(k_b_tau, dim3(P), t, 0, s, ctl.p, vst.p, rst.p, nbV, nbR, pDd, pAd, gt);
    }

    // grid of the gated group kernels: they loop over their groups
    static unsigned round_grid(int ng) { return (unsigned)std::max(1, std::min(ng, 2048)); }

    // outer update of the problems whose inner loop has ended (:1170-1220),
    // gated by the device count of such problems
    void outer_update() {
        const dim3 bV(round_grid(nbV)), bR(round_grid(nbR)), bP(P), t(256);
        const int32_t *go = gates.p + kGateOuter;
        hipLaunchKernelGGL(k_b_relout, bV, t, 0, s, ddims.p, vst.p, P, ctl.p, U[0], U[1], rcd, pRO,
                           nbV, go);
        hipLaunchKernelGGL(k_b_relout_scalar, bP, t, 0, s, ctl.p, vst.p, nbV, pRO, (double)o->tol,
                           go);
        hipLaunchKernelGGL(k_b_mlp_sigma, bV, t, 0, s, ddims.p, vst.p, P, ctl.p, U[0], U[1], sx, sy,
                           sz, ox, oy, oz, p4, nbV, go);
        hipLaunchKernelGGL(k_b_mlp_delta, bR, t, 0, s, ddims.p, rst.p, P, ctl.p, Au[0], Au[1], b,
                           delta, p5, nbR, go);
        hipLaunchKernelGGL(k_b_outer_scalars, bP, t, 0, s, ctl.p, vst.p, rst.p, nbV, nbR, p4, p5,
                           go);
        hipLaunchKernelGGL(k_b_outer_d, bV, t, 0, s, ddims.p, vst.p, P, ctl.p, d, g[0], g[1], dts,
                           pOD, nbV, go);
        hipLaunchKernelGGL(k_b_set_dd, bP, t, 0, s, ctl.p, vst.p, nbV, pOD, (int)o->maxcnt,
                           (int)o->maxit, go);
    }

    // one round: every unfinished problem takes one step of its own loop
    void round() {
        hipLaunchKernelGGL(k_b_classify, dim3(1), dim3(256), 0, s, ctl.p, P, gates.p);
        outer_update();
        hipLaunchKernelGGL(k_b_gate_tau, dim3(1), dim3(256), 0, s, ctl.p, P, grp.p,
                           (int)groups.size(), gates.p);
        tau_pass();
        iteration();
    }

    // graph_iters rounds per launch, then the count of unfinished problems
    void capture() {
        hipGraph_t gr;
        TV3_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < graph_iters; ++k) round();
        hipLaunchKernelGGL(k_b_alive, dim3(1), dim3(256), 0, s, ctl.p, P, gates.p);
        TV3_HIP(hipStreamEndCapture(s, &gr));
        TV3_HIP(hipGraphInstantiate(&graph, gr, nullptr, nullptr, 0));
        TV3_HIP(hipGraphDestroy(gr));
    }

    template <typename F>
    float time_ms(F &&f) {
        hipEvent_t a, e;
        TV3_HIP(hipEventCreate(&a)); TV3_HIP(hipEventCreate(&e));
        f();
        TV3_HIP(hipEventRecord(a, s));
        for (int r = 0; r < 3; ++r) f();
        TV3_HIP(hipEventRecord(e, s));
        TV3_HIP(hipEventSynchronize(e));
        float ms = 0;
        TV3_HIP(hipEventElapsedTime(&ms, a, e));
        (void)hipEventDestroy(a); (void)hipEventDestroy(e);
        return ms / 3;
    }

    // launch shape of one block-diagonal product: the row-group kernel with
    // L lanes per row (code L) or the row-block kernel with NB-entry blocks
    // (code -NB), timed with every problem on
    int tune_matrix(SparseRows &S, const float *x, float *y, const int32_t *starts,
                    const std::vector<int64_t> &row_starts) {
        float best = 1e30f;
        int code = 8;
        for (int L : {2, 4, 8, 16, 32, 64}) {
            S.blk_nnz = 0;
            S.lanes = L;
            const float t = time_ms([&] { product(S, x, y, nullptr, starts, P_ALL); });
            if (t < best) { best = t; code = L; }
        }
        for (int NB : {1024, 2048, 4096}) {
            build_row_blocks(ctx, S, NB, &row_starts);
            const float t = time_ms([&] { product(S, x, y, nullptr, starts, P_ALL); });
            if (t < best) { best = t; code = -NB; }
        }
        set_code(S, code, row_starts);
        return code;
    }
    void set_code(SparseRows &S, int code, const std::vector<int64_t> &row_starts) {
        if (code > 0) {
            S.blk_nnz = 0;
            S.lanes = code;
        } else if (S.blk_nnz != -code) {
            build_row_blocks(ctx, S, -code, &row_starts);
        }
    }
    std::vector<int64_t> a_starts, at_starts;   // first rows of the problems
    void tune() {
        if (!A.rows) { lanes_N = lanes_T = 8; return; }
        lanes_N = tune_matrix(A, U[0], AuT, rst.p, a_starts);
        lanes_T = tune_matrix(AT, AuT, gT, vst.p, at_starts);
    }

    // K splits of each group's two products, timed (the scratch holds the
    // largest candidate)
    void tune_groups() {
        for (DGroup &g : groups) {
            float bN = 1e30f, bT = 1e30f;
            for (int S = 1; S <= kMaxSplits; S *= 2) {
                if (gemm_rhs_shape_ok(g.op->m, g.op->n, S)) {
                    const float t = time_ms([&] { group_N(g, U[0], AuT, S); });
                    if (t < bN) { bN = t; g.sN = S; }
                }
                if (gemm_rhs_shape_ok(g.op->n, g.op->m, S)) {
                    const float t = time_ms([&] { group_T(g, AuT, gT, nullptr, S); });
                    if (t < bT) { bT = t; g.sT = S; }
                }
            }
        }
    }
    // A^T in panel form: want = 1024 / 2048 (that row-block size), -1
    // (never) or 0 (the faster of both sizes when faster than the tuned row
    // kernel).  Tiles of 512 voxels must not straddle two problems.
    void choose_panels(int want, const std::vector<int64_t> &v0, int ncsr) {
        panels_T = 0;
        if (!AT.rows || want < 0) return;
        bool aligned = true;
        for (int k = 0; k < ncsr; ++k) aligned = aligned && v0[k] % 512 == 0;
        TV3_REQUIRE(want == 0 || want == 1024 || want == 2048 || want == 4096, TV3_E_ARG,
                    "panels_T must be 1024, 2048, 4096, -1 or 0");
        if (!aligned) {
            TV3_REQUIRE(want == 0, TV3_E_UNSUPPORTED,
                        "panel form of A^T needs every sparse problem's voxel count to be a "
                        "multiple of 512");
            return;
        }
        const std::vector<int> blks =
            want ? std::vector<int>{want} : std::vector<int>{1024, 2048, 4096};
        std::vector<Panels> cand;
        if (!build_transposed_panels(ctx, AT, blks, cand)) {
            TV3_REQUIRE(want == 0, TV3_E_UNSUPPORTED,
                        "a tile of this batch crosses too many rays for the panel form");
            return;
        }
        if (want) {
            PT = std::move(cand[0]);
            panels_T = want;
            return;
        }
        float best = time_ms([&] { product(AT, AuT, gT, nullptr, vst.p, P_ALL); });
        for (size_t i = 0; i < cand.size(); ++i) {
            const float t = time_ms([&] {
                launch_panel(cand[i], AuT, EpiBatch{gT, nullptr, vst.p, P, ctl.p, P_ALL}, s);
            });
            if (t < best) { best = t; PT = std::move(cand[i]); panels_T = blks[i]; }
        }
    }
    // A in two-pass panel form (segment sums per panel, then per ray): want as
    // for choose_panels (4096 too for boxes of 8 tiles); box = tiles per panel,
    // 1 (8^3) or 8 (16^3: fewer, longer segments), 0 = timed with want 0 and 1
    // otherwise.  Boxes of 8 tiles need every sparse problem on 8^3 tiles with
    // even tile counts.
    void choose_panels_A(int want, int box, const std::vector<int64_t> &v0) {
        panels_A = 0;
        box_A = 0;
        if (!A.rows || want < 0) return;
        TV3_REQUIRE(box == 0 || box == 1 || box == 8, TV3_E_ARG, "panel_box_A must be 0, 1 or 8");
        if (want != 0 && box == 0) box = 1;
        TV3_REQUIRE(want == 0 || want == 1024 || want == 2048 || (want == 4096 && box == 8),
                    TV3_E_ARG, "panels_A must be 1024, 2048 (4096 with boxes of 8 tiles), -1 or 0");
        std::vector<TileVolume> vols;
        bool boxes_ok = true;
        for (int k = 0; k < ncsr; ++k) {
            const BDims &d = dims[k];
            if (d.B != 8 || d.v0 % 4096 != 0) { boxes_ok = false; break; }
            vols.push_back({d.v0, d.p / 8, d.q / 8, d.r / 8});
            if (vols.back().nby % 2 || vols.back().nbx % 2 || vols.back().nbz % 2) boxes_ok = false;
        }
        TV3_REQUIRE(box != 8 || boxes_ok, TV3_E_UNSUPPORTED,
                    "panels of 8 tiles need every sparse problem on 8^3 tiles with an even tile "
                    "count along every axis");
        const std::vector<int64_t> starts(v0.begin(), v0.begin() + ncsr);
        struct Cand { Panels P; SegSum R; DevBuf<int32_t> st; int blk, box; };
        std::vector<Cand> cand;
        auto add = [&](const std::vector<int> &blks, int bx) {
            std::vector<Panels> cp;
            std::vector<SegSum> cr;
            std::vector<int32_t> st;
            if (!build_forward_panels(ctx, AT, blks, starts, cp, cr, st,
                                      bx == 8 ? &vols : nullptr))
                return;
            for (size_t i = 0; i < blks.size(); ++i) {
                Cand c{std::move(cp[i]), std::move(cr[i]), DevBuf<int32_t>(st.size()), blks[i], bx};
                TV3_HIP(hipMemcpyAsync(c.st.p, st.data(), sizeof(int32_t) * st.size(),
                                       hipMemcpyHostToDevice, s));
                cand.push_back(std::move(c));
            }
        };
        if (want) {
            add({want}, box);
        } else {
            if (box != 8) add({1024, 2048}, 1);
            if (box != 1 && boxes_ok) add({2048, 4096}, 8);
        }
        if (cand.empty()) {
            TV3_REQUIRE(want == 0, TV3_E_UNSUPPORTED,
                        "a tile of this batch crosses too many rays for the panel form");
            return;
        }
        for (Cand &c : cand) c.R.rows = A.rows;      // the sparse problems' rays only
        auto run = [&](Cand &c) {
            launch_panel(c.P, U[0], EpiSegBatch{c.R.part.p, c.st.p, ncsr, ctl.p, P_ALL}, s);
            launch_seg_sum(c.R, EpiBatch{AuT, nullptr, rst.p, P, ctl.p, P_ALL}, s);
        };
        int pick = want ? 0 : -1;
        if (!want) {
            float best = time_ms([&] { product(A, U[0], AuT, nullptr, rst.p, P_ALL); });
            for (size_t i = 0; i < cand.size(); ++i) {
                const float t = time_ms([&] { run(cand[i]); });
                if (t < best) { best = t; pick = (int)i; }
            }
        }
        if (pick >= 0) {
            Cand &c = cand[pick];
            segst = std::move(c.st);
            PA = std::move(c.P);
            RS = std::move(c.R);
            panels_A = c.blk;
            box_A = c.box;
        }
    }
    static constexpr int kMaxSplits = 32;
};

void concat_matrix(tv3_ctx *ctx, hipStream_t s, const std::vector<const SparseRows *> &parts,
                   const std::vector<int64_t> &row0, const std::vector<int64_t> &col0,
                   int64_t rows_total, int64_t cols_total, SparseRows &out) {
    const int P = (int)parts.size();
    int64_t nnz = 0;
    for (auto *S : parts) nnz += S->nnz;
    TV3_REQUIRE(nnz < ((int64_t)1 << 31), TV3_E_SHAPE,
                "the batch holds more than 2^31 - 1 entries (int32 CSR)");
    out.rows = rows_total; out.cols = cols_total; out.nnz = nnz; out.vtype = parts[0]->vtype;
    const size_t vs = dtype_size(out.vtype);
    out.ptr.alloc(rows_total + 1);
    out.ind.alloc(nnz > 0 ? nnz : 1);
    out.val.alloc((nnz > 0 ? nnz : 1) * vs);
    int64_t off = 0;
    for (int k = 0; k < P; ++k) {
        const SparseRows &S = *parts[k];
        const int64_t padded = (k + 1 < P ? row0[k + 1] : rows_total) - row0[k];
        hipLaunchKernelGGL(k_concat_ptr, dim3((unsigned)((padded + 256) / 256)), dim3(256), 0, s,
                           S.ptr.p, S.rows, padded, off, out.ptr.p + row0[k]);
        if (S.nnz) {
            hipLaunchKernelGGL(k_concat_ind, dim3(grid_for(S.nnz)), dim3(256), 0, s, S.ind.p,
                               S.nnz, col0[k], out.ind.p + off);
            TV3_HIP(hipMemcpyAsync(out.val.p + off * vs, S.val.p, S.nnz * vs,
                                   hipMemcpyDeviceToDevice, s));
        }
        off += S.nnz;
    }
    (void)ctx;
    TV3_HIP(hipGetLastError());
}

}  // namespace

// The last batched solve of a context, reused by the next call that has the
// same operators (by serial number: a destroyed operator never matches, and
// the session's buffers stay until the next batched call or the context's
// end), the same volume shapes, the same options captured in the graph
// (tol, maxcnt, maxit, whether the A^T-skip rule runs, graph size) and any
// launch shapes the call fixes.
struct BatchSession : CtxSession {
    std::vector<uint64_t> uids;      // caller's order
    std::vector<int32_t> dims;       // caller's order, three per problem
    double tol = 0;
    int32_t maxcnt = 0, maxit = 0;
    tv3_opts oc;                     // the current call's options (Batch::o points here)
    Batch B;
    std::vector<std::unique_ptr<FusedLayout>> lay;   // internal order
    std::vector<int> orig;           // orig[k]: caller's index of internal problem k
};

static bool session_matches(const BatchSession &S, int P, tv3_op *const *ops,
                            const int32_t *dims, const tv3_opts &o,
                            const tv3_batch_info *binfo) {
    if ((int)S.uids.size() != P) return false;
    for (int k = 0; k < P; ++k)
        if (!ops[k] || ops[k]->uid != S.uids[k]) return false;
    if (!std::equal(S.dims.begin(), S.dims.end(), dims)) return false;
    const Batch &B = S.B;
    if (S.tol != o.tol || S.maxcnt != o.maxcnt || S.maxit != o.maxit) return false;
    if (B.skipping != (o.skipMulRatio != 0.0)) return false;
    if ((o.graph_iters > 0 ? o.graph_iters : 8) != B.graph_iters) return false;
    if (!binfo) return true;
    if (binfo->lanes_N != 0 && binfo->lanes_T != 0 &&
        (binfo->lanes_N != B.lanes_N || binfo->lanes_T != B.lanes_T))
        return false;
    auto panels_ok = [](int want, int have) { return want == 0 || (want < 0 ? 0 : want) == have; };
    if (!panels_ok(binfo->panels_T, B.panels_T) || !panels_ok(binfo->panels_A, B.panels_A))
        return false;
    const int want_box = binfo->panel_box_A != 0 ? binfo->panel_box_A
                         : binfo->panels_A > 0   ? 1
                                                 : 0;
    if (want_box != 0 && binfo->panels_A >= 0 && want_box != B.box_A) return false;
    if (binfo->splits_N > 0 && binfo->splits_T > 0)
        for (const auto &g : B.groups)
            if (g.sN != binfo->splits_N || g.sT != binfo->splits_T) return false;
    return true;
}

// Everything of a batched solve that depends only on its operators, volume
// shapes and the options its graph captures: ranges, the fused numbering of
// the sparse problems and their block-diagonal matrices, dense groups,
// buffers, launch shapes (timed unless given), panel forms, the graph.
static std::unique_ptr<BatchSession> build_session(tv3_ctx *ctx, int P, tv3_op *const *ops_in,
                                                   const int32_t *dims_in, const tv3_opts &o,
                                                   const tv3_batch_info *binfo) {
    std::unique_ptr<BatchSession> SP(new BatchSession());
    BatchSession &S = *SP;
    S.oc = o;
    for (int k = 0; k < P; ++k) S.uids.push_back(ops_in[k] ? ops_in[k]->uid : 0);
    S.dims.assign(dims_in, dims_in + 3 * P);
    S.tol = o.tol; S.maxcnt = o.maxcnt; S.maxit = o.maxit;
    Batch &B = S.B;
    B.ctx = ctx; B.s = ctx->stream; B.o = &S.oc; B.P = P;
    hipStream_t s = B.s;
    B.graph_iters = o.graph_iters > 0 ? o.graph_iters : 8;
    // internal order: sparse problems first, then the problems of each shared
    // dense operator consecutively; orig[k] = caller's index of problem k
    std::vector<int> &orig = S.orig;
    {
        std::vector<int> dense;
        for (int k = 0; k < P; ++k) {
            tv3_op *op = ops_in[k];
            TV3_REQUIRE(op, TV3_E_ARG, "NULL operator in the batch");
            if (op->kind == OpKind::CSR) orig.push_back(k);
            else if (op->kind == OpKind::DENSE) dense.push_back(k);
            else
                throw tv3_error(TV3_E_FORMAT,
                                "batched solves take sparse (CSR) or dense operators");
        }
        std::stable_sort(dense.begin(), dense.end(), [&](int a, int b) {
            return ops_in[a] < ops_in[b];
        });
        orig.insert(orig.end(), dense.begin(), dense.end());
    }
    std::vector<tv3_op *> ops(P);
    std::vector<int32_t> dims3v(3 * P);
    for (int k = 0; k < P; ++k) {
        ops[k] = ops_in[orig[k]];
        for (int a = 0; a < 3; ++a) dims3v[3 * k + a] = dims_in[3 * orig[k] + a];
    }
    const int32_t *dims3 = dims3v.data();
    int ncsr = 0;
    while (ncsr < P && ops[ncsr]->kind == OpKind::CSR) ++ncsr;
    // ranges
    B.dims.resize(P);
    std::vector<int32_t> vst(P + 1), rst(P + 1);
    std::vector<int64_t> v0(P), r0(P);
    int64_t N = 0, M = 0, Ncsr = 0, Mcsr = 0;
    for (int k = 0; k < P; ++k) {
        tv3_op *op = ops[k];
        if (k < ncsr)
            TV3_REQUIRE(op->A.vtype == ops[0]->A.vtype, TV3_E_FORMAT,
                        "every sparse operator of a batch must have the same weight type");
        const int32_t *dk = dims3 + 3 * k;
        TV3_REQUIRE(dk[0] > 0 && dk[1] > 0 && dk[2] > 0, TV3_E_ARG,
                    "dims must be three positive sizes");
        const int64_t n = (int64_t)dk[0] * dk[1] * dk[2];
        TV3_REQUIRE(n == op->n, TV3_E_SHAPE, "prod(dims) must equal the columns of A");
        TV3_REQUIRE(n < ((int64_t)1 << 31) && op->m < ((int64_t)1 << 31), TV3_E_SHAPE,
                    "problem too large for the batched solver");
        BDims &bd = B.dims[k];
        bd.p = dk[0]; bd.q = dk[1]; bd.r = dk[2];
        bd.B = 1; bd.nby = 0; bd.nbx = 0;
        bd.v0 = N; bd.n = n; bd.r0 = M; bd.m = op->m;
        v0[k] = N; r0[k] = M;
        vst[k] = (int32_t)(N / GB); rst[k] = (int32_t)(M / GB);
        N += pad512(n);
        M += pad256(std::max<int64_t>(op->m, 1));
        if (k + 1 == ncsr) { Ncsr = N; Mcsr = M; }
    }
    TV3_REQUIRE(N / GB < ((int64_t)1 << 31) && M / GB < ((int64_t)1 << 31), TV3_E_SHAPE,
                "batch too large");
    vst[P] = (int32_t)(N / GB); rst[P] = (int32_t)(M / GB);
    B.N = N; B.M = M; B.nbV = (int)(N / GB); B.nbR = (int)(M / GB);
    B.ddims.alloc(P);
    B.vst.alloc(P + 1);
    B.rst.alloc(P + 1);
    TV3_HIP(hipMemcpyAsync(B.vst.p, vst.data(), sizeof(int32_t) * (P + 1), hipMemcpyHostToDevice, s));
    TV3_HIP(hipMemcpyAsync(B.rst.p, rst.data(), sizeof(int32_t) * (P + 1), hipMemcpyHostToDevice, s));
    // sparse problems in the fused solver's numbering (8^3 tiles of 4^3 bricks
    // where the dims allow, rays in a locality order), then the block-diagonal
    // matrices over the sparse prefix; the renumbered copies are released
    // once concatenated.  Dense problems keep the natural order.
    std::vector<std::unique_ptr<FusedLayout>> &lay = S.lay;
    lay.resize(P);
    if (ncsr > 0) {
        std::vector<const SparseRows *> pa(ncsr), pt(ncsr);
        std::vector<int64_t> v0c(v0.begin(), v0.begin() + ncsr), r0c(r0.begin(), r0.begin() + ncsr);
        for (int k = 0; k < ncsr; ++k) {
            lay[k].reset(new FusedLayout());
            renumber_operator(ops[k], dims3 + 3 * k, lay[k].get());
            BDims &bd = B.dims[k];
            bd.B = lay[k]->B; bd.nby = lay[k]->nby; bd.nbx = lay[k]->nbx;
            pa[k] = &lay[k]->A;
            pt[k] = &lay[k]->AT;
        }
        concat_matrix(ctx, s, pa, r0c, v0c, Mcsr, N, B.A);
        concat_matrix(ctx, s, pt, v0c, r0c, Ncsr, M, B.AT);
        B.a_starts = r0c;
        B.at_starts = v0c;
        TV3_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < ncsr; ++k) { lay[k]->A = SparseRows(); lay[k]->AT = SparseRows(); }
    }
    TV3_HIP(hipMemcpyAsync(B.ddims.p, B.dims.data(), sizeof(BDims) * P, hipMemcpyHostToDevice, s));
    B.choose_tiles();
    // dense groups
    size_t gscratch = 0;
    for (int k = ncsr; k < P;) {
        int e = k;
        while (e < P && ops[e] == ops[k]) ++e;
        tv3_op *op = ops[k];
        TV3_REQUIRE(op->dtype == TV3_BF16, TV3_E_FORMAT,
                    "dense operators of a batch must hold BF16 values (matrix-core products)");
        TV3_REQUIRE(op->m % 256 == 0 && op->n % 256 == 0 && op->lda % 8 == 0, TV3_E_SHAPE,
                    "dense operators of a batch need rows and columns multiples of 256");
        dense_operand_order(op, s);
        Batch::DGroup g;
        g.op = op; g.k0 = k; g.G = e - k;
        g.v0 = v0[k]; g.r0 = r0[k]; g.ldn = pad512(op->n); g.ldm = pad256(op->m);
        B.groups.push_back(g);
        gscratch = std::max({gscratch, gemm_rhs_scratch(op->m, Batch::kMaxSplits),
                             gemm_rhs_scratch(op->n, Batch::kMaxSplits)});
        k = e;
    }
    if (gscratch) B.gpart.alloc(gscratch);
    // vectors
    const int nv = 2 + 1 + 2 + 1 + 1 + 3 + 3 + 1 + 1 + 1, mv = 2 + 1 + 1 + 1 + 1;
    B.pool.alloc((size_t)nv * N + (size_t)mv * M);
    TV3_HIP(hipMemsetAsync(B.pool.p, 0, B.pool.bytes(), s));
    {
        float *p = B.pool.p;
        float **nvp[] = {&B.U[0], &B.U[1], &B.T, &B.g[0], &B.g[1], &B.gT, &B.d, &B.sx, &B.sy,
                         &B.sz, &B.ox, &B.oy, &B.oz, &B.dts, &B.Atb, &B.rcd};
        for (float **q : nvp) { *q = p; p += N; }
        float **mvp[] = {&B.Au[0], &B.Au[1], &B.AuT, &B.b, &B.delta, &B.last};
        for (float **q : mvp) { *q = p; p += M; }
    }
    {
        const size_t nbV = B.nbV, nbR = B.nbR;
        B.parts.alloc(3 * nbV + 3 * nbR + 7 * nbV + nbV + nbR + nbV + nbR + nbV + nbV +
                      2 * nbR);
        double *q = B.parts.p;
        B.pD = q; q += 3 * nbV;
        B.pN = q; q += 3 * nbR;
        B.pW = q; q += 7 * nbV;
        B.pDd = q; q += nbV;
        B.pAd = q; q += nbR;
        B.p4 = q; q += nbV;
        B.p5 = q; q += nbR;
        B.pOD = q; q += nbV;
        B.pRO = q; q += nbV;
        B.pK = q; q += 2 * nbR;
    }
    B.hist_stride = (int64_t)std::max(o.maxit, 1) * HF;
    B.hist.alloc((size_t)P * B.hist_stride);
    B.ctl.alloc(P);
    B.hctl.alloc(P);
    {   // round gates and the dense groups' problem ranges
        std::vector<int32_t> gk;
        for (size_t i = 0; i < B.groups.size(); ++i) {
            TV3_REQUIRE(i == 0 || B.groups[i].k0 == B.groups[i - 1].k0 + B.groups[i - 1].G,
                        TV3_E_INTERNAL, "dense groups are consecutive");
            gk.push_back(B.groups[i].k0);
        }
        gk.push_back(B.groups.empty() ? P : B.groups.back().k0 + B.groups.back().G);
        B.grp.alloc(gk.size());
        TV3_HIP(hipMemcpyAsync(B.grp.p, gk.data(), sizeof(int32_t) * gk.size(),
                               hipMemcpyHostToDevice, s));
        B.gates.alloc(kGateGroup + B.groups.size());
        TV3_HIP(hipMemsetAsync(B.gates.p, 0, B.gates.bytes(), s));
        B.halive.alloc(1);
    }

    // launch shapes of the two products (every problem active), unless given
    if (binfo && binfo->lanes_N != 0 && binfo->lanes_T != 0) {
        B.lanes_N = binfo->lanes_N;
        B.lanes_T = binfo->lanes_T;
        if (B.A.rows) {
            B.set_code(B.A, B.lanes_N, B.a_starts);
            B.set_code(B.AT, B.lanes_T, B.at_starts);
        }
    } else {
        B.tune();
    }
    B.choose_panels(binfo ? binfo->panels_T : 0, v0, ncsr);
    B.ncsr = ncsr;
    B.choose_panels_A(binfo ? binfo->panels_A : 0, binfo ? binfo->panel_box_A : 0, v0);
    if (binfo && binfo->splits_N > 0 && binfo->splits_T > 0) {
        for (auto &g : B.groups) {
            TV3_REQUIRE(gemm_rhs_shape_ok(g.op->m, g.op->n, binfo->splits_N) &&
                            gemm_rhs_shape_ok(g.op->n, g.op->m, binfo->splits_T) &&
                            binfo->splits_N <= Batch::kMaxSplits &&
                            binfo->splits_T <= Batch::kMaxSplits,
                        TV3_E_ARG, "K splits must divide the dense products' K into 64-multiples");
            g.sN = binfo->splits_N;
            g.sT = binfo->splits_T;
        }
    } else {
        B.tune_groups();
    }
    B.skipping = o.skipMulRatio != 0.0;
    B.scl.alloc(P);
    B.bptr.alloc(P);
    B.bad.alloc(1);
    B.hbad.alloc(1);
    B.capture();
    return SP;
}

void solve_batched(tv3_ctx *ctx, int P, tv3_op *const *ops_in, const int32_t *dims_in,
                   const tv3_opts &o, const float *const *b_in, float *const *U_in,
                   const float *const *Ut_in, tv3_info *infos, tv3_history *hists,
                   tv3_batch_info *binfo) {
    auto t_host0 = std::chrono::steady_clock::now();
    // the context's last batch when this call has its operators, shapes and
    // captured options; otherwise it is freed first and this one built
    BatchSession *S = dynamic_cast<BatchSession *>(ctx->batch_session.get());
    const bool reused = S && session_matches(*S, P, ops_in, dims_in, o, binfo);
    if (!reused) {
        ctx->batch_session.reset();
        std::unique_ptr<BatchSession> ns = build_session(ctx, P, ops_in, dims_in, o, binfo);
        S = ns.get();
        ctx->batch_session = std::move(ns);
    }
    S->oc = o;
    Batch &B = S->B;
    hipStream_t s = B.s;
    const std::vector<int> &orig = S->orig;
    const std::vector<std::unique_ptr<FusedLayout>> &lay = S->lay;
    const int64_t N = B.N, M = B.M;
    std::vector<const float *> b_dev(P), Ut_dev(P);
    std::vector<float *> U_dev(P);
    for (int k = 0; k < P; ++k) {
        b_dev[k] = b_in[orig[k]];
        U_dev[k] = U_in[orig[k]];
        Ut_dev[k] = Ut_in ? Ut_in[orig[k]] : nullptr;
    }
    hipEvent_t ev0, ev1, evl0, evl1;
    TV3_HIP(hipEventCreate(&ev0)); TV3_HIP(hipEventCreate(&ev1));
    TV3_HIP(hipEventCreate(&evl0)); TV3_HIP(hipEventCreate(&evl1));

    TV3_HIP(hipMemsetAsync(B.pool.p, 0, B.pool.bytes(), s));
    TV3_HIP(hipEventRecord(ev0, s));

    // scaleb (:225-266) per problem on the device; b gathered into the
    // internal order and scaled by the factor found there
    TV3_HIP(hipMemcpyAsync(B.bptr.p, b_dev.data(), sizeof(const float *) * P,
                           hipMemcpyHostToDevice, s));
    TV3_HIP(hipMemsetAsync(B.bad.p, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_b_scaleb, dim3(P), dim3(1024), 0, s, B.bptr.p, B.ddims.p, B.scl.p,
                       B.bad.p);
    TV3_HIP(hipMemcpyAsync(B.hbad.p, B.bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    TV3_HIP(hipStreamSynchronize(s));
    TV3_REQUIRE(*B.hbad.p == 0, TV3_E_NONFINITE, "b holds NaN or Inf");
    for (int k = 0; k < P; ++k) {
        const int64_t m = B.dims[k].m;
        hipLaunchKernelGGL(k_gather_scale, dim3(grid_for(m)), dim3(256), 0, s, b_dev[k],
                           lay[k] ? lay[k]->r_int2nat.p : nullptr, m, B.scl.p + k,
                           B.b + B.dims[k].r0);
    }

    FCtl *H = B.hctl.p;
    std::memset(H, 0, sizeof(FCtl) * P);
    for (int k = 0; k < P; ++k) {
        FCtl &c = H[k];
        c.mu = std::min(o.mu0, o.mu); c.beta = std::min(o.beta0, o.beta);
        c.muDbeta = c.mu / c.beta;
        c.c = o.c; c.gamma = o.gamma; c.rate_gam = o.rate_gam; c.tol_inn = o.tol_inn;
        c.gam_opt = o.gam; c.gam = o.gam; c.rate_cnt = o.rate_cnt; c.mu_max = o.mu;
        c.beta_max = o.beta; c.maxit = o.maxit; c.nonneg = o.nonneg;
        c.iso = o.tv_mode == TV3_TV_ISO; c.hist_cap = o.maxit; c.active = 1; c.alpha = 1;
        c.skip_ratio = o.skipMulRatio; c.max_rel = o.maxRelativeChange;
        c.n_init = (int32_t)std::min<uint32_t>(o.numInitIterations, INT32_MAX);
    }
    B.write_ctl();

    // initialisation (:1128-1153), every problem at once
    B.prod_T(B.b, B.Atb, nullptr, P_ALL);
    for (int k = 0; k < P; ++k) {
        const BDims &bd = B.dims[k];
        if (o.warm_start)
            hipLaunchKernelGGL(k_gather_scale, dim3(grid_for(bd.n)), dim3(256), 0, s, U_dev[k],
                               lay[k] ? lay[k]->v_int2nat.p : nullptr, bd.n, B.scl.p + k,
                               B.U[0] + bd.v0);
    }
    if (!o.warm_start)
        TV3_HIP(hipMemcpyAsync(B.U[0], B.Atb, sizeof(float) * N, hipMemcpyDeviceToDevice, s));
    TV3_HIP(hipMemcpyAsync(B.rcd, B.U[0], sizeof(float) * N, hipMemcpyDeviceToDevice, s));
    B.prod_N(B.U[0], B.Au[0], P_ALL);
    hipLaunchKernelGGL(k_b_res, dim3(B.nbR), dim3(256), 0, s, B.ddims.p, B.rst.p, P, B.ctl.p,
                       (int)P_ALL, B.Au[0], B.b, B.delta, B.Au[0], B.Au[1], B.pN);
    B.prod_T(B.Au[0], B.g[0], B.Atb, P_ALL);
    if (B.skipping)                  // the initial get_g multiplies (forced): last Au
        TV3_HIP(hipMemcpyAsync(B.last, B.Au[0], sizeof(float) * M, hipMemcpyDeviceToDevice, s));
    B.update_w(1);
    hipLaunchKernelGGL(k_b_init, dim3(P), dim3(256), 0, s, B.ctl.p, B.vst.p, B.rst.p, B.nbV,
                       B.nbR, B.pN, B.pW);
    TV3_HIP(hipGetLastError());
    TV3_HIP(hipEventRecord(evl0, s));
    if (o.maxcnt <= 0 || o.maxit <= 0) {     // the outer loop of :1166 does not run
        B.read_ctl();
        for (int k = 0; k < P; ++k) { H[k].done = 1; H[k].active = 0; H[k].need_tau = 0; }
        B.write_ctl();
    }

    // every round of a launch advances each unfinished problem by one step
    // (outer update, first step, inner iteration: one problem's next step
    // never waits for another's); after a launch the host reads one count.
    // A problem takes at most maxit inner iterations, maxcnt outer updates
    // and one first step per inner iteration, so the rounds are bounded.
    int graph_launches = 0;
    const int64_t max_rounds = 2 * ((int64_t)std::max(o.maxit, 0) + 1) *
                                   ((int64_t)std::max(o.maxcnt, 0) + 2) + 16;
    for (;;) {
        TV3_REQUIRE((int64_t)graph_launches * B.graph_iters <= max_rounds, TV3_E_INTERNAL,
                    "batched solve did not finish within its iteration limits");
        TV3_HIP(hipGraphLaunch(B.graph, s));
        ++graph_launches;
        TV3_HIP(hipMemcpyAsync(B.halive.p, B.gates.p + kGateAlive, sizeof(int32_t),
                               hipMemcpyDeviceToHost, s));
        TV3_HIP(hipStreamSynchronize(s));
        if (*B.halive.p == 0) break;
    }
    B.read_ctl();
    TV3_HIP(hipEventRecord(evl1, s));
    for (int k = 0; k < P; ++k) {
        const BDims &bd = B.dims[k];
        const float *Uk = H[k].cur ? B.U[1] : B.U[0];
        hipLaunchKernelGGL(k_scatter_unscale, dim3(grid_for(bd.n)), dim3(256), 0, s, Uk + bd.v0,
                           lay[k] ? lay[k]->v_int2nat.p : nullptr, bd.n, B.scl.p + k,
                           U_dev[k]);
    }
    TV3_HIP(hipEventRecord(ev1, s));
    TV3_HIP(hipStreamSynchronize(s));
    float ms = 0, msl = 0;
    TV3_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    TV3_HIP(hipEventElapsedTime(&msl, evl0, evl1));
    int64_t tot = 0;
    std::vector<double> hh;
    for (int k = 0; k < P; ++k) {
        const FCtl &c = H[k];
        tv3_info &inf = infos[orig[k]];
        std::memset(&inf, 0, sizeof(inf));
        inf.total_iters = c.total;
        inf.outer_iters = c.outer;
        inf.rel_chg = c.relout;
        inf.secs = ms / 1000.0;
        inf.loop_secs = msl / 1000.0;
        inf.backtracks = c.backtracks;
        inf.fallbacks = c.fallbacks;
        inf.n_matvec_N = 1 + (int64_t)c.total + c.fallbacks + c.n_first;
        // with skipping: A^T b, the initial product and the products get_g took
        inf.n_matvec_T = B.skipping ? 2 + (int64_t)c.n_T : 2 + (int64_t)c.total + c.fallbacks;
        inf.status = c.nonfinite ? TV3_E_NONFINITE : TV3_OK;
        tot += c.total;
        if (Ut_dev[k]) {
            auto e = diff_norm<float>(ctx, B.dims[k].n, Ut_dev[k], U_dev[k], s);
            inf.rel_error = std::sqrt(e[0]) / std::sqrt(e[1]);
        }
        if (hists && hists[orig[k]].cap > 0) {
            tv3_history &h = hists[orig[k]];
            const int cnt = std::min(c.total, h.cap);
            hh.assign((size_t)std::max(cnt, 1) * HF, 0.0);
            if (cnt > 0)
                TV3_HIP(hipMemcpy(hh.data(), B.hist.p + (int64_t)k * B.hist_stride,
                                  sizeof(double) * HF * cnt, hipMemcpyDeviceToHost));
            for (int j = 0; j < cnt; ++j) {
                const double *q = &hh[(size_t)j * HF];
                h.f[j] = q[0]; h.lam1[j] = q[1]; h.lam2[j] = q[2]; h.lam3[j] = q[3];
                h.lam4[j] = q[4]; h.lam5[j] = q[5]; h.tau[j] = q[6]; h.alpha[j] = q[7];
                h.relchg[j] = q[8]; h.mu[j] = q[9]; h.beta[j] = q[10]; h.C[j] = q[11];
                h.cnt[j] = (int32_t)q[12]; h.outer[j] = (int32_t)q[13];
                if (h.mulT) h.mulT[j] = (int32_t)q[14];
                if (h.f_trial) h.f_trial[j] = q[15];
            }
            h.count = cnt;
        }
    }
    if (binfo) {
        std::memset(binfo, 0, sizeof(*binfo));
        binfo->lanes_N = B.lanes_N;
        binfo->lanes_T = B.lanes_T;
        binfo->dense_groups = (int32_t)B.groups.size();
        if (!B.groups.empty()) {
            binfo->splits_N = B.groups[0].sN;
            binfo->splits_T = B.groups[0].sT;
        }
        binfo->panels_T = B.panels_T;
        binfo->panels_A = B.panels_A;
        binfo->panel_box_A = B.box_A;
        binfo->graph_iters = B.graph_iters;
        binfo->graph_launches = graph_launches;
        binfo->secs = ms / 1000.0;
        binfo->loop_secs = msl / 1000.0;
        binfo->host_secs =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t_host0).count();
        binfo->total_iters = tot;
        binfo->nnz = B.A.nnz;
        for (const auto &g : B.groups) binfo->dense_values += g.op->m * g.op->n;
        binfo->voxels = N;
        binfo->rays = M;
    }
    (void)hipEventDestroy(ev0); (void)hipEventDestroy(ev1);
    (void)hipEventDestroy(evl0); (void)hipEventDestroy(evl1);
}

}  // namespace tv3
