// solver_fused.hip -- TVAL3 with fused kernels, device-side FP64 scalars and
// hipGraph-captured inner iterations.
//
// Algorithm: 3d/tval3_gpu_3d.cpp:1088-1345 (same formulas, same order of
// decisions).  What changes is how the work is laid out:
//
//  * one steady inner iteration is seven steps, one launch each, except
//    that A x takes two with the panel form of A and the search runs inside
//    the A^T launch when A^T has its panel form:
//      descend   U_t = max(U - tau d, 0) and, from U and U_t on the fly,
//                DU, W = shrink(DU - sigma/beta), V0 = DU - W, V1 = DU_t - W
//                with |V1|^2, <V1,sigma>, <V0,V1>                  (:671-701)
//      A x       Au_t = A U_t, reductions |r1|^2, <delta,r1>, <r0,r1>
//                of r = Au - b                                    (:382, :500)
//      A^T x     g_t = A^T Au_t - Atb                            (:423, :565)
//      search    the nonmonotone Armijo test (:839-890) on f(alpha): U, Au,
//                g, DU are affine in alpha (:717-737), so lam2/lam3 are
//                quadratic and lam4/lam5 linear in alpha and the up to five
//                backtracking sweeps of update_g become scalar work; it
//                reads only the partials of descend and A x
//      apply     U, g, Au = prev + alpha (trial - prev), skipped if alpha = 1
//      update_W  W = shrink(DU - sigma/beta), V = DU - W, g2 = D^T V,
//                d = g2 + mu/beta g + DtsAtd, and the Barzilai-Borwein sums
//                ss = |U - Up|^2, sy = <U - Up, d - d_prev> (equal to the
//                reference's <uup, dg2 + mu/beta dg> inside an inner loop,
//                where mu/beta and DtsAtd are fixed)          (:965-1029, :896)
//      end       lam1, lam2, lam4, f, Q, C, RelChg, tau, stop flags (:1176-1202)
//  * W, g2, Ux..Uz and their previous copies are never stored: W and DU are
//    recomputed from U and sigma where needed (the D^T of get_g, :536, is
//    dead in the reference: g2 is overwritten in update_W before any read);
//  * every kernel starts by reading a device control block and returns at
//    once when the inner loop has ended, so K iterations are captured in one
//    graph and the host reads the control block once per graph;
//  * the first iteration of an inner loop (tau from |d|^2, |Dd|^2, |Ad|^2,
//    :583-625), the cnt == 5 fallback (:876-883) and the outer update
//    (:1206-1249) run as ordinary launches.
//  * vectors are FP32 in the layout of layout.hip, products accumulate in
//    FP32, all inner products / norms accumulate in FP64 (block partials
//    reduced in a fixed order), all scalars are FP64.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "panels.hpp"
#include "tiles.hpp"
#include "tval3_step.hpp"
#include "vec.hpp"

namespace tv3 {

namespace {

struct Lay {
    int p, q, r, nby, nbx;
    int nbz;             // tiles along z (B = 8)
    int64_t pq;
};

template <int B>
__device__ __forceinline__ int64_t vidx(const Lay &L, int y, int x, int z) {
    return brick_index(B, L.pq, L.p, L.nby, L.nbx, y, x, z);
}

// Voxel groups of 256 threads (one wavefront per brick where there are
// bricks): B = 8 -> the four bricks of one z-half of a tile (256 consecutive
// internal voxels); B = 4 -> four bricks consecutive in y; B = 1 -> 64 (y) x 4
// (x) voxels of one z-plane.
// Workgroups of the 1-D grid walk contiguous chunks of groups, consecutive
// chunks on one XCD.
struct Groups {
    int gy, gx, gz;      // groups along y, x, z
    int64_t count;
};

template <int B>
static Groups groups_of(const Lay &L) {
    Groups g;
    if (B == 8) { g.gy = L.nby; g.gx = L.nbx; g.gz = L.r / 4; }
    else if (B == 4) { g.gy = (L.nby + 3) / 4; g.gx = L.nbx; g.gz = L.r / 4; }
    else { g.gy = (L.p + 63) / 64; g.gx = (L.q + 3) / 4; g.gz = L.r; }
    g.count = (int64_t)g.gy * g.gx * g.gz;
    return g;
}

template <int B>
__device__ __forceinline__ bool voxel_of(const Lay &L, const Groups &G, int grp, int &y, int &x,
                                         int &z) {
    const int t = threadIdx.x;
    const int gy = grp % G.gy, rest = grp / G.gy;
    const int gx = rest % G.gx, gz = rest / G.gx;
    if (B == 8) {
        const int lane = t & 63;
        y = gy * 8 + ((t >> 4) & 4) + (lane & 3);
        x = gx * 8 + ((t >> 5) & 4) + ((lane >> 2) & 3);
        z = gz * 4 + (lane >> 4);
        return true;
    }
    if (B == 4) {
        const int by = gy * 4 + (t >> 6), lane = t & 63;
        y = by * 4 + (lane & 3);
        x = gx * 4 + ((lane >> 2) & 3);
        z = gz * 4 + (lane >> 4);
        return by < L.nby;
    }
    y = gy * 64 + (t & 63);
    x = gx * 4 + (t >> 6);
    z = gz;
    return y < L.p && x < L.q;
}

constexpr int kMaxStencilGrid = 1 << 30;

static unsigned stencil_grid(const Groups &G) {
    return (unsigned)(G.count < kMaxStencilGrid ? G.count : kMaxStencilGrid);
}

// range of groups of this workgroup (group counts below 2^31, checked at
// set-up: 32-bit index arithmetic)
__device__ __forceinline__ void group_range(int64_t count, int &g0, int &g1) {
    const int Gd = gridDim.x, cnt = (int)count, chunk = (cnt + Gd - 1) / Gd;
    const int w = xcd_remap(blockIdx.x, Gd);
    g0 = w * chunk;
    g1 = min(cnt, g0 + chunk);
}

__device__ __forceinline__ int nxt(int v, int n) { return v + 1 == n ? 0 : v + 1; }
__device__ __forceinline__ int prv(int v, int n) { return v == 0 ? n - 1 : v - 1; }

// ------------------------------------------------------------------ stencils
// descend: U_t = max(U_c - tau d, 0) (:690-696); V0/V1 statistics.
template <int B>
__global__ __launch_bounds__(256) void k_descend(Lay L, Groups G, const float *__restrict__ Uc,
                                                 const float *__restrict__ d,
                                                 const float *__restrict__ sx,
                                                 const float *__restrict__ sy,
                                                 const float *__restrict__ sz,
                                                 const float *__restrict__ wsx,
                                                 const float *__restrict__ wsy,
                                                 const float *__restrict__ wsz,
                                                 float *__restrict__ Un, const FCtl *C,
                                                 double *part) {
    // control scalars and the data loads before the graph's stop flag is
    // tested (the loads are in flight while it arrives)
    const int active = C->active;
    double acc[3] = {0, 0, 0};
    const float tau = (float)C->tau, beta = (float)C->beta, ib = 1.f / beta;
    const bool nn = C->nonneg, iso = C->iso;
    const bool old = C->sigw_old;
    // after an outer step W is still the one update_W built from the old
    // sigma at the old beta (3d/tval3_gpu_3d.cpp:484-492 reads the stored W)
    const float ibw = old ? 1.f / (float)C->beta_w : ib;
    int g0, g1;
    group_range(G.count, g0, g1);
    // with B = 8 every voxel of a group is valid and the grid has one group
    // per workgroup: the flag is tested in the first pass, after its loads
    // (tied to grp == g0 so that it stays there); other B test it first
    if constexpr (B != 8) {
        if (!active) return;
    }
    for (int grp = g0; grp < g1; ++grp) {
        int y, x, z;
        if (!voxel_of<B>(L, G, grp, y, x, z)) continue;
        const int64_t i = vidx<B>(L, y, x, z);
        const int64_t iy = vidx<B>(L, nxt(y, L.p), x, z), ix = vidx<B>(L, y, nxt(x, L.q), z),
                      iz = vidx<B>(L, y, x, nxt(z, L.r));
        const float u0 = Uc[i], uy = Uc[iy], ux = Uc[ix], uz = Uc[iz];
        const float d0 = d[i], dy = d[iy], dx = d[ix], dz = d[iz];
        const float s_x = sx[i], s_y = sy[i], s_z = sz[i];
        const float w_x = old ? wsx[i] : s_x, w_y = old ? wsy[i] : s_y, w_z = old ? wsz[i] : s_z;
        if (B == 8 && grp == g0 && !active) {
            keep_issued(u0, uy, ux, uz, d0, dy, dx, dz, s_x, s_y, s_z, w_x, w_y, w_z);
            return;
        }
        auto step = [&](float u, float dv) {
            float t = u - tau * dv;
            return (nn && t < 0.f) ? 0.f : t;
        };
        const float t0 = step(u0, d0), ty = step(uy, dy), tx = step(ux, dx), tz = step(uz, dz);
        float wx, wy, wz;
        shrink3(ux - u0, uy - u0, uz - u0, w_x, w_y, w_z, ibw, iso, wx, wy, wz);
        const float v0x = (ux - u0) - wx, v0y = (uy - u0) - wy, v0z = (uz - u0) - wz;
        const float v1x = (tx - t0) - wx, v1y = (ty - t0) - wy, v1z = (tz - t0) - wz;
        acc[0] += (double)v1x * v1x + (double)v1y * v1y + (double)v1z * v1z;
        acc[1] += (double)v1x * s_x + (double)v1y * s_y + (double)v1z * s_z;
        acc[2] += (double)v0x * v1x + (double)v0y * v1y + (double)v0z * v1z;
        Un[i] = t0;
    }
    if (!active) return;
    block_sum_store<3>(acc, part, gridDim.x);
}

// update_W + D^T + gradient + BB sums (:965-1029, :896-919, :1017-1067)
template <int B>
__global__ __launch_bounds__(256) void k_update_w(Lay L, Groups G, const float *__restrict__ U,
                                                  const float *__restrict__ Uc,
                                                  const float *__restrict__ sx,
                                                  const float *__restrict__ sy,
                                                  const float *__restrict__ sz,
                                                  const float *__restrict__ g,
                                                  const float *__restrict__ dts,
                                                  float *__restrict__ d, const FCtl *C,
                                                  double *part) {
    if (!C->active || C->need_fallback) return;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const float beta = (float)C->beta, ib = 1.f / beta, mdb = (float)C->muDbeta;
    const bool iso = C->iso;
    int g0, g1;
    group_range(G.count, g0, g1);
    for (int grp = g0; grp < g1; ++grp) {
        int y, x, z;
        if (!voxel_of<B>(L, G, grp, y, x, z)) continue;
        const int yp = nxt(y, L.p), ym = prv(y, L.p), xp = nxt(x, L.q), xm = prv(x, L.q),
                  zp = nxt(z, L.r), zm = prv(z, L.r);
        const int64_t i = vidx<B>(L, y, x, z);
        const float u0 = U[i];
        const float uyp = U[vidx<B>(L, yp, x, z)], uxp = U[vidx<B>(L, y, xp, z)],
                    uzp = U[vidx<B>(L, y, x, zp)];
        // V at i
        float wx, wy, wz;
        const float s_x = sx[i], s_y = sy[i], s_z = sz[i];
        const float l1 = shrink3(uxp - u0, uyp - u0, uzp - u0, s_x, s_y, s_z, ib, iso, wx,
                                 wy, wz);
        const float vx = (uxp - u0) - wx, vy = (uyp - u0) - wy, vz = (uzp - u0) - wz;
        // V_a at i - e_a (needs the full DU there)
        float vyb, vxb, vzb;
        {
            const int64_t j = vidx<B>(L, ym, x, z);
            const float uj = U[j];
            const float dyj = u0 - uj, dxj = U[vidx<B>(L, ym, xp, z)] - uj,
                        dzj = U[vidx<B>(L, ym, x, zp)] - uj;
            float a, b, c;
            shrink3(dxj, dyj, dzj, sx[j], sy[j], sz[j], ib, iso, a, b, c);
            vyb = dyj - b;
        }
        {
            const int64_t j = vidx<B>(L, y, xm, z);
            const float uj = U[j];
            const float dxj = u0 - uj, dyj = U[vidx<B>(L, yp, xm, z)] - uj,
                        dzj = U[vidx<B>(L, y, xm, zp)] - uj;
            float a, b, c;
            shrink3(dxj, dyj, dzj, sx[j], sy[j], sz[j], ib, iso, a, b, c);
            vxb = dxj - a;
        }
        {
            const int64_t j = vidx<B>(L, y, x, zm);
            const float uj = U[j];
            const float dzj = u0 - uj, dxj = U[vidx<B>(L, y, xp, zm)] - uj,
                        dyj = U[vidx<B>(L, yp, x, zm)] - uj;
            float a, b, c;
            shrink3(dxj, dyj, dzj, sx[j], sy[j], sz[j], ib, iso, a, b, c);
            vzb = dzj - c;
        }
        const float g2 = vyb - vy + vxb - vx + vzb - vz;            // D^T V, :501
        const float dn = mdb * g[i] + g2 + dts[i];                   // :915-918
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
    block_sum_store<7>(acc, part, gridDim.x);
}

// The same on one 8x8x8 tile per workgroup of the B = 8 numbering (512
// threads, one voxel each): U is staged in LDS over the tile and a one-voxel
// shell ([-1, 8]^3), V = DU - W is computed once for every voxel of the tile
// and of its three lower faces (where D^T reads it), and D^T V reads V from
// LDS.  Per voxel the arithmetic is k_update_w's, so d is identical.
// internal index of that voxel for the tile at tile coordinates (ty, tx, tz)
__device__ __forceinline__ int64_t tile_gidx(const Lay &L, int ty, int tx, int tz,
                                             const TileOff &w) {
    return tile_gidx(L.nby, L.nbx, L.nbz, ty, tx, tz, w);
}

__global__ __launch_bounds__(512) void k_update_w_tile(Lay L, int ntiles, int tpw,
                                                       const float *__restrict__ U,
                                                       const float *__restrict__ Uc,
                                                       const float *__restrict__ sx,
                                                       const float *__restrict__ sy,
                                                       const float *__restrict__ sz,
                                                       const float *__restrict__ g,
                                                       const float *__restrict__ dts,
                                                       float *__restrict__ d, const FCtl *C,
                                                       double *part) {
    constexpr int NBOX = kTileBox * kTileBox * kTileBox;
    __shared__ float su[NBOX], svx[NBOX], svy[NBOX], svz[NBOX];
    // control scalars, then the first tile's data loads, then the graph's
    // stop flag: the loads are in flight while the flag arrives
    const int active = C->active && !C->need_fallback;
    const float beta = (float)C->beta, ib = 1.f / beta, mdb = (float)C->muDbeta;
    const bool iso = C->iso;
    const int t = threadIdx.x;
    // a workgroup owns tpw consecutive tiles (ordered y, x, z: consecutive
    // tiles are y neighbours, one tile's upper shell is the next one's face),
    // consecutive workgroups on one XCD
    const int tile0 = xcd_remap(blockIdx.x, gridDim.x) * tpw;
    const int tile1 = min(ntiles, tile0 + tpw);
    // per thread, for every tile: the U box ([-1, 8]^3) entries t and
    // k1 = t + 512, sigma at the tile voxel q0 = t and at the lower-face voxel
    // q1 = 512 + t (t < 192), and g, dts, Uc, d of voxel t; indices clamped
    // instead of guarded
    const int k1 = t + 512 < NBOX ? t + 512 : NBOX - 1;
    int a0, b0, c0;
    tile_pos(t, a0, b0, c0);
    const int q1 = t < 192 ? t : 191;                 // face item 512 + q1
    const int f = q1 >> 6, j = q1 & 63, fu = j & 7, fv = j >> 3;
    const int a1 = f == 0 ? -1 : fu, b1 = f == 0 ? fu : (f == 1 ? -1 : fv),
              c1 = f == 2 ? -1 : fv;
    const TileOff w0 = box_off(t), w1 = box_off(k1), wf = tile_off(a1, b1, c1);
    struct Loads { float ub[2], s0[3], s1[3], gv, dv, ucv, dold; };
    auto load = [&](int tile, Loads &r) {
        const int ty = tile % L.nby, rest = tile / L.nby;
        const int tx = rest % L.nbx, tz = rest / L.nbx;
        r.ub[0] = U[tile_gidx(L, ty, tx, tz, w0)];
        r.ub[1] = U[tile_gidx(L, ty, tx, tz, w1)];
        const int64_t i = ((int64_t)tile << 9) + t, gi1 = tile_gidx(L, ty, tx, tz, wf);
        r.s0[0] = sx[i]; r.s0[1] = sy[i]; r.s0[2] = sz[i];
        r.s1[0] = sx[gi1]; r.s1[1] = sy[gi1]; r.s1[2] = sz[gi1];
        r.gv = g[i]; r.dv = dts[i]; r.ucv = Uc[i]; r.dold = d[i];
    };
    Loads cur;
    load(tile0 < tile1 ? tile0 : ntiles - 1, cur);
    if (!active || tile0 >= tile1) {
        keep_issued(cur.ub[0], cur.ub[1], cur.s0[0], cur.s0[1], cur.s0[2], cur.s1[0], cur.s1[1],
                    cur.s1[2], cur.gv, cur.dv, cur.ucv, cur.dold);
        if (!active) return;
    }
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    // V at a voxel of the tile or (t < 192) of its lower faces y = -1,
    // x = -1, z = -1, where D^T reads it
    auto vat = [&](int a, int b, int c, const float *s, bool tile_voxel) {
        const int sl = box_slot(a, b, c);
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
    for (int tile = tile0; tile < tile1; ++tile) {
        su[t] = cur.ub[0];
        if (t + 512 < NBOX) su[t + 512] = cur.ub[1];
        // the next tile's loads fly while this one is computed
        Loads nxt;
        const bool more = tile + 1 < tile1;
        if (more) load(tile + 1, nxt);
        __syncthreads();
        vat(a0, b0, c0, cur.s0, true);
        if (t < 192) vat(a1, b1, c1, cur.s1, false);
        __syncthreads();
        {
            const int sl = box_slot(a0, b0, c0);
            const float g2 = svy[sl - 1] - svy[sl] + svx[sl - kTileBox] - svx[sl] +
                             svz[sl - kTileBox * kTileBox] - svz[sl];      // D^T V, :501
            const float dn = mdb * cur.gv + g2 + cur.dv;                    // :915-918
            const float du = su[sl] - cur.ucv;
            acc[3] += (double)du * du;
            acc[4] += (double)du * (double)(dn - cur.dold);
            acc[5] += (double)cur.ucv * cur.ucv;
            acc[6] += (double)dn * dn;
            d[((int64_t)tile << 9) + t] = dn;
        }
        if (more) {
            __syncthreads();        // su and sv are rewritten for the next tile
            cur = nxt;
        }
    }
    block_sum_store<7, 8>(acc, part, gridDim.x);
}

// descend on tiles of B = 8 (:690-696): U_t = max(U_c - tau d, 0) and the
// line-search statistics of k_descend.  U_c and d are staged in LDS on the
// tile and its upper shell ([0, 8]^3: the +1 neighbours D takes); a
// workgroup walks tpw consecutive tiles with the next tile's loads in flight.
constexpr int kDBox = 9;

__global__ __launch_bounds__(512) void k_descend_tile(Lay L, int ntiles, int tpw,
                                                      const float *__restrict__ Uc,
                                                      const float *__restrict__ d,
                                                      const float *__restrict__ sx,
                                                      const float *__restrict__ sy,
                                                      const float *__restrict__ sz,
                                                      const float *__restrict__ wsx,
                                                      const float *__restrict__ wsy,
                                                      const float *__restrict__ wsz,
                                                      float *__restrict__ Un, const FCtl *C,
                                                      double *part) {
    constexpr int NB = kDBox * kDBox * kDBox;        // 729
    constexpr int NSHELL = NB - 512;                  // 217 upper-shell voxels
    __shared__ float su[NB], sd[NB];
    const int active = C->active;
    const float tau = (float)C->tau, beta = (float)C->beta;
    const bool nn = C->nonneg, iso = C->iso, old = C->sigw_old;
    // W of the last update_W: old sigma at the old beta after an outer step
    const float ibw = 1.f / (old ? (float)C->beta_w : beta);
    const int t = threadIdx.x;
    const int tile0 = xcd_remap(blockIdx.x, gridDim.x) * tpw;
    const int tile1 = min(ntiles, tile0 + tpw);
    int a0, b0, c0;
    tile_pos(t, a0, b0, c0);
    const int sl = (c0 * kDBox + b0) * kDBox + a0;
    // shell voxel e = t (t < 217): the points of [0, 8]^3 with c = 8 (81),
    // then b = 8, c < 8 (72), then a = 8, b < 8, c < 8 (64)
    const int e = t < NSHELL ? t : NSHELL - 1;
    int ae, be, ce;
    if (e < 81) { ce = 8; be = e / 9; ae = e % 9; }
    else if (e < 153) { be = 8; ce = (e - 81) / 9; ae = (e - 81) % 9; }
    else { ae = 8; ce = (e - 153) >> 3; be = (e - 153) & 7; }
    const int sle = (ce * kDBox + be) * kDBox + ae;
    const TileOff we = tile_off(ae, be, ce);
    struct Loads { float u, dd, ue, de, s[3], w[3]; };
    auto load = [&](int tile, Loads &r) {
        const int ty = tile % L.nby, rest = tile / L.nby;
        const int tx = rest % L.nbx, tz = rest / L.nbx;
        const int64_t i = ((int64_t)tile << 9) + t, ie = tile_gidx(L, ty, tx, tz, we);
        r.u = Uc[i]; r.dd = d[i]; r.ue = Uc[ie]; r.de = d[ie];
        r.s[0] = sx[i]; r.s[1] = sy[i]; r.s[2] = sz[i];
        if (old) { r.w[0] = wsx[i]; r.w[1] = wsy[i]; r.w[2] = wsz[i]; }
        else { r.w[0] = r.s[0]; r.w[1] = r.s[1]; r.w[2] = r.s[2]; }
    };
    Loads cur;
    load(tile0 < tile1 ? tile0 : ntiles - 1, cur);
    if (!active || tile0 >= tile1) {
        keep_issued(cur.u, cur.dd, cur.ue, cur.de, cur.s[0], cur.s[1], cur.s[2], cur.w[0],
                    cur.w[1], cur.w[2]);
        if (!active) return;
    }
    double acc[3] = {0, 0, 0};
    auto step = [&](float u, float dv) {
        const float r = u - tau * dv;
        return (nn && r < 0.f) ? 0.f : r;
    };
    for (int tile = tile0; tile < tile1; ++tile) {
        su[sl] = cur.u;
        sd[sl] = cur.dd;
        if (t < NSHELL) {
            su[sle] = cur.ue;
            sd[sle] = cur.de;
        }
        Loads nxt;
        const bool more = tile + 1 < tile1;
        if (more) load(tile + 1, nxt);
        __syncthreads();
        const float u0 = su[sl], uy = su[sl + 1], ux = su[sl + kDBox], uz = su[sl + kDBox * kDBox];
        const float t0 = step(u0, sd[sl]), ty = step(uy, sd[sl + 1]),
                    tx = step(ux, sd[sl + kDBox]), tz = step(uz, sd[sl + kDBox * kDBox]);
        float wx, wy, wz;
        shrink3(ux - u0, uy - u0, uz - u0, cur.w[0], cur.w[1], cur.w[2], ibw, iso, wx, wy, wz);
        const float v0x = (ux - u0) - wx, v0y = (uy - u0) - wy, v0z = (uz - u0) - wz;
        const float v1x = (tx - t0) - wx, v1y = (ty - t0) - wy, v1z = (tz - t0) - wz;
        acc[0] += (double)v1x * v1x + (double)v1y * v1y + (double)v1z * v1z;
        acc[1] += (double)v1x * cur.s[0] + (double)v1y * cur.s[1] + (double)v1z * cur.s[2];
        acc[2] += (double)v0x * v1x + (double)v0y * v1y + (double)v0z * v1z;
        Un[((int64_t)tile << 9) + t] = t0;
        if (more) {
            __syncthreads();        // su and sd are rewritten for the next tile
            cur = nxt;
        }
    }
    block_sum_store<3, 8>(acc, part, gridDim.x);
}

// |D d|^2 for the first-iteration step length (:587-596)
template <int B>
__global__ __launch_bounds__(256) void k_dd_norm(Lay L, Groups G, const float *__restrict__ d,
                                                 double *part) {
    double acc[1] = {0};
    int g0, g1;
    group_range(G.count, g0, g1);
    for (int grp = g0; grp < g1; ++grp) {
        int y, x, z;
        if (!voxel_of<B>(L, G, grp, y, x, z)) continue;
        const int64_t i = vidx<B>(L, y, x, z);
        const float d0 = d[i];
        const float a = d[vidx<B>(L, nxt(y, L.p), x, z)] - d0,
                    b = d[vidx<B>(L, y, nxt(x, L.q), z)] - d0,
                    c = d[vidx<B>(L, y, x, nxt(z, L.r))] - d0;
        acc[0] += (double)a * a + (double)b * b + (double)c * c;
    }
    block_sum_store<1>(acc, part, gridDim.x);
}

// update_mlp, sigma part (:1036-1069): keeps the old sigma for the W of the
// next descend and returns lam4 = <sigma_new, V>
template <int B>
__global__ __launch_bounds__(256) void k_mlp_sigma(Lay L, Groups G, const float *__restrict__ U,
                                                   float *__restrict__ sx, float *__restrict__ sy,
                                                   float *__restrict__ sz, float *__restrict__ ox,
                                                   float *__restrict__ oy, float *__restrict__ oz,
                                                   const FCtl *C, double *part) {
    double acc[1] = {0};
    const float beta = (float)C->beta, ib = 1.f / beta;
    int g0, g1;
    group_range(G.count, g0, g1);
    for (int grp = g0; grp < g1; ++grp) {
        int y, x, z;
        if (!voxel_of<B>(L, G, grp, y, x, z)) continue;
        const int64_t i = vidx<B>(L, y, x, z);
        const float u0 = U[i];
        const float dy = U[vidx<B>(L, nxt(y, L.p), x, z)] - u0,
                    dx = U[vidx<B>(L, y, nxt(x, L.q), z)] - u0,
                    dz = U[vidx<B>(L, y, x, nxt(z, L.r))] - u0;
        const float a = sx[i], b = sy[i], c = sz[i];
        float wx, wy, wz;
        shrink3(dx, dy, dz, a, b, c, ib, C->iso, wx, wy, wz);
        const float vx = dx - wx, vy = dy - wy, vz = dz - wz;
        const float na = a - beta * vx, nb = b - beta * vy, nc = c - beta * vz;
        ox[i] = a; oy[i] = b; oz[i] = c;
        sx[i] = na; sy[i] = nb; sz[i] = nc;
        acc[0] += (double)na * vx + (double)nb * vy + (double)nc * vz;
    }
    block_sum_store<1>(acc, part, gridDim.x);
}

// ------------------------------------------------------------------ m-vectors
__global__ __launch_bounds__(256) void k_mlp_delta(int64_t m, const float *__restrict__ Au,
                                                   const float *__restrict__ b,
                                                   float *__restrict__ delta, const FCtl *C,
                                                   double *part) {
    double acc[1] = {0};
    const float mu = (float)C->mu;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256) {
        const float r = Au[i] - b[i];
        const float dn = delta[i] - mu * r;
        delta[i] = dn;
        acc[0] += (double)dn * r;
    }
    block_sum_store<1>(acc, part, gridDim.x);
}

__global__ __launch_bounds__(256) void k_apply_alpha(int64_t n, int64_t m,
                                                     const float *__restrict__ Uc,
                                                     float *__restrict__ Un,
                                                     const float *__restrict__ gc,
                                                     float *__restrict__ gn,
                                                     const float *__restrict__ Auc,
                                                     float *__restrict__ Aun, const FCtl *C) {
    if (!C->active || C->need_fallback || C->alpha == 1.0) return;
    const float a = (float)C->alpha;
    const int64_t mx = n > m ? n : m, G = (int64_t)gridDim.x * 256;
#pragma unroll 4
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < mx; i += G) {
        if (i < n) {
            const float u = Uc[i], g = gc[i];
            Un[i] = u + a * (Un[i] - u);
            gn[i] = g + a * (gn[i] - g);
        }
        if (i < m) {
            const float v = Auc[i];
            Aun[i] = v + a * (Aun[i] - v);
        }
    }
}

// outer update of d and DtsAtd (:1239-1246) without a stored g2
__global__ __launch_bounds__(256) void k_outer_d(int64_t n, float *__restrict__ d,
                                                 const float *__restrict__ g,
                                                 float *__restrict__ dts, const FCtl *C,
                                                 double *part) {
    double acc[1] = {0};
    const float ratio = (float)C->ratio, mo = (float)C->mdb_old, mn = (float)C->muDbeta;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float dv = d[i], gv = g[i];
        const float g2 = dv - mo * gv - dts[i];
        const float nd = ratio * dv;
        const float dn = mn * gv + g2 + nd;
        dts[i] = nd;
        d[i] = dn;
        acc[0] += (double)dn * dn;
    }
    block_sum_store<1>(acc, part, gridDim.x);
}

// |U - rcdU|^2 then rcdU = U (:1208-1217)
__global__ __launch_bounds__(256) void k_relout(int64_t n, const float *__restrict__ U,
                                                float *__restrict__ rcd, double *part) {
    double acc[1] = {0};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float u = U[i], t = u - rcd[i];
        acc[0] += (double)t * t;
        rcd[i] = u;
    }
    block_sum_store<1>(acc, part, gridDim.x);
}

// ------------------------------------------------------------------ product epilogues
struct EpiRes {          // Au_n = s; r1 = s - b; sums |r1|^2, <delta,r1>, <r0,r1>
    float *Aun;
    const float *b, *delta, *Auc;
    const FCtl *C;
    double *part;
    mutable double acc[3];
    struct Pre { float b, d, a; };   // b, delta, Auc of the row
    __device__ bool skip() const { return !C->active; }
    __device__ __forceinline__ Pre load(int r) const { return {b[r], delta[r], Auc[r]}; }
    __device__ __forceinline__ void row(int r, float s, bool active, const Pre &p) const {
        if (active) {
            Aun[r] = s;
            const float r1 = s - p.b, r0 = p.a - p.b;
            acc[0] += (double)r1 * r1;
            acc[1] += (double)p.d * r1;
            acc[2] += (double)r0 * r1;
        }
    }
    __device__ __forceinline__ void finish() const { block_sum_store<3>(acc, part, gridDim.x); }
};

struct EpiNorm {         // |s|^2 only (A d for the first-iteration step)
    const FCtl *C;
    double *part;
    mutable double acc[1];
    struct Pre {};
    __device__ bool skip() const { return false; }
    __device__ __forceinline__ Pre load(int) const { return {}; }
    __device__ __forceinline__ void row(int, float s, bool active, const Pre &) const {
        if (active) acc[0] += (double)s * s;
    }
    __device__ __forceinline__ void finish() const { block_sum_store<1>(acc, part, gridDim.x); }
};

struct EpiGrad {         // g_n = s - Atb
    float *g;
    const float *Atb;
    const FCtl *C;
    int gated;                       // 1: only when get_g's skip decision multiplies
    struct Pre { float a; };         // Atb of the row
    __device__ bool skip() const { return !C->active || (gated && !C->mul_T); }
    __device__ __forceinline__ Pre load(int r) const { return {Atb[r]}; }
    __device__ __forceinline__ void row(int r, float s, bool active, const Pre &p) const {
        if (active) g[r] = s - p.a;
    }
    __device__ __forceinline__ void finish() const {}
};

// A^T-skip heuristic (:398-468): |Au_n - last Au|^2 and |Au_n|^2
__global__ __launch_bounds__(256) void k_skip_stats(int64_t m, const float *__restrict__ Aun,
                                                    const float *__restrict__ last,
                                                    const FCtl *C, double *part) {
    if (!C->active) return;
    double acc[2] = {0, 0};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * 256) {
        const float a = Aun[i], dl = a - last[i];
        acc[0] += (double)dl * dl;
        acc[1] += (double)a * a;
    }
    block_sum_store<2>(acc, part, gridDim.x);
}

// after the gated A^T product: a product makes Au_n the last Au; a skipped
// one leaves g as it was (g_n = g_c), as the reference's get_g does
__global__ __launch_bounds__(256) void k_skip_after(int64_t n, int64_t m,
                                                    const float *__restrict__ gc,
                                                    float *__restrict__ gn,
                                                    const float *__restrict__ Aun,
                                                    float *__restrict__ last, const FCtl *C) {
    if (!C->active) return;
    const int64_t G = (int64_t)gridDim.x * 256;
    if (C->mul_T) {
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += G) last[i] = Aun[i];
    } else {
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += G) gn[i] = gc[i];
    }
}

// ------------------------------------------------------------------ scalar kernels
// Sum K partial arrays of nb entries each in a fixed order (1024 threads, four
// independent load chains per thread); result in out[K] on thread 0.
template <int K>
__device__ void reduce_parts(const double *part, int nb, double *out) {
    __shared__ double sh[K][1024];
    const int t = threadIdx.x;
    for (int k = 0; k < K; ++k) {
        const double *p = part + (int64_t)k * nb;
        double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        int j = t;
        for (; j + 3 * 1024 < nb; j += 4 * 1024) {
            a0 += p[j]; a1 += p[j + 1024]; a2 += p[j + 2048]; a3 += p[j + 3072];
        }
        for (; j < nb; j += 1024) a0 += p[j];
        sh[k][t] = (a0 + a1) + (a2 + a3);
    }
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (t < off)
            for (int k = 0; k < K; ++k) sh[k][t] += sh[k][t + off];
        __syncthreads();
    }
    if (t == 0)
        for (int k = 0; k < K; ++k) out[k] = sh[k][0];
}

// Totals of KA partial arrays of nbA entries and KB arrays of nbB entries on
// one workgroup of 1024 threads, every load issued up front: thread t takes
// entries t + 1024 u (u < 4) of each array per pass of 4096 entries (one
// pass for nb <= 4096, one memory round trip), then wavefront sums (DPP) and
// the 16 wavefront totals on wavefront 0, in a fixed order.  Every lane of
// wavefront 0 returns the K totals.
template <int KA, int KB>
__device__ __forceinline__ void reduce_wg(const double *pa, int nbA, const double *pb, int nbB,
                                          double *out) {
    constexpr int K = KA + KB;
    __shared__ double sh[K][16];
    const int t = threadIdx.x;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.0;
    const int nbm = nbA > nbB ? nbA : nbB;
    for (int j0 = 0; j0 < nbm; j0 += 4096) {
        double x[4][K];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + u * 1024 + t;
#pragma unroll
            for (int k = 0; k < KA; ++k) x[u][k] = j < nbA ? pa[(int64_t)k * nbA + j] : 0.0;
#pragma unroll
            for (int k = 0; k < KB; ++k) x[u][KA + k] = j < nbB ? pb[(int64_t)k * nbB + j] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += (x[0][k] + x[1][k]) + (x[2][k] + x[3][k]);
    }
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double w = wave_sum(v[k]);
        if (lane == 0) sh[k][wave] = w;
    }
    __syncthreads();
    if (t < 64) {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = wave_sum(lane < 16 ? sh[k][lane] : 0.0);
    }
}

// Armijo line search from the block partials of descend (pa) and of the A
// product's epilogue (pb) directly, one workgroup
__global__ __launch_bounds__(1024) void k_line_search_parts(FCtl *C, const double *pa, int nbA,
                                                            const double *pb, int nbB, int mode) {
    if (!C->active) return;
    double s[6];
    reduce_wg<3, 3>(pa, nbA, pb, nbB, s);
    if (threadIdx.x == 0) line_search(C, s, s + 3, mode);
}

// The sums of reduce_wg (1024 threads) with one 256-thread workgroup:
// thread t plays threads t + 256 q, q = 0..3, of the larger workgroup with
// the same loads and the same order of additions (virtual wave 4 q + wave),
// so both give the same bits.  One array set (K sums) at a time, so few
// registers are live: the kernel that hosts this keeps its occupancy.
template <int K>
__device__ __forceinline__ void reduce_256_part(const double *p, int nb, int nbm, int q,
                                                double (*sh)[16]) {
    const int t = threadIdx.x, tt = t + 256 * q;
    const int lane = t & 63, wave = t >> 6;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.0;
    for (int j0 = 0; j0 < nbm; j0 += 4096) {
        double x[4][K];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + u * 1024 + tt;
#pragma unroll
            for (int k = 0; k < K; ++k) x[u][k] = j < nb ? p[(int64_t)k * nb + j] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += (x[0][k] + x[1][k]) + (x[2][k] + x[3][k]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double w = wave_sum(v[k]);
        if (lane == 0) sh[k][4 * q + wave] = w;
    }
}

template <int KA, int KB>
__device__ __forceinline__ void reduce_wg_256(const double *pa, int nbA, const double *pb,
                                              int nbB, double *out) {
    constexpr int K = KA + KB;
    __shared__ double sh[K][16];
    const int t = threadIdx.x, lane = t & 63;
    const int nbm = nbA > nbB ? nbA : nbB;
#pragma unroll 1
    for (int h = 0; h < 8; ++h) {
        const int q = h >> 1;
        if (h & 1) reduce_256_part<KB>(pb, nbB, nbm, q, sh + KA);
        else reduce_256_part<KA>(pa, nbA, nbm, q, sh);
    }
    __syncthreads();
    if (t < 64) {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = wave_sum(lane < 16 ? sh[k][lane] : 0.0);
    }
}

// The gradient epilogue of the A^T panel product with the line search of the
// iteration in one extra workgroup (block 0) of the same launch: the search
// reads only the partials of descend and of the A product, complete when the
// launch starts, and leaves active as it is (the panel workgroups read it);
// the kernels after it take need_fallback as inactive.
struct EpiGradLS : EpiGrad {
    static constexpr int kExtra = 1;
    FCtl *Cw;
    const double *pa, *pb;
    int nbA, nbB, mode;
    __device__ void extra() const {
        if (!Cw->active) return;
        double s[6];
        reduce_wg_256<3, 3>(pa, nbA, pb, nbB, s);
        if (threadIdx.x == 0) line_search(Cw, s, s + 3, mode, true);
    }
};

// end of an inner iteration from update_W's block partials directly
__global__ __launch_bounds__(1024) void k_iter_end_parts(FCtl *C, const double *p, int nb,
                                                         double *hist) {
    if (!C->active) return;
    if (C->need_fallback) {          // the line search failed (left active set)
        __syncthreads();             // every thread has read the flags
        if (threadIdx.x == 0) C->active = 0;
        return;
    }
    double s[7];
    reduce_wg<7, 0>(p, nb, nullptr, 0, s);
    if (threadIdx.x == 0) iter_end(C, s, hist);
}

// first-iteration step length tau = |dd / (|Dd|^2 + mu/beta |Ad|^2)| (:583-625);
// fallback = 1 also shrinks gam (:877)
__global__ void k_tau_first(FCtl *C, const double *pD, int nbD, const double *pA, int nbA,
                            int fallback) {
    double a[1], b[1];
    reduce_parts<1>(pD, nbD, a);
    reduce_parts<1>(pA, nbA, b);
    if (threadIdx.x == 0) tau_first_step(C, a[0], b[0], fallback);
}

// after the initial get_g / update (:1146-1153)
__global__ void k_init_scalars(FCtl *C, const double *pN, int nbN, const double *pW, int nbW) {
    double a[3], b[7];
    reduce_parts<3>(pN, nbN, a);
    reduce_parts<7>(pW, nbW, b);
    if (threadIdx.x == 0) init_step(C, a, b);
}

__global__ void k_skip_decide(FCtl *C, const double *p, int nb, int force) {
    if (!C->active) return;
    double a[2];
    reduce_parts<2>(p, nb, a);
    if (threadIdx.x == 0) skip_decide(C, a[0], a[1], force);
}

__global__ void k_relout_scalar(FCtl *C, const double *p, int nb) {
    double a[1];
    reduce_parts<1>(p, nb, a);
    if (threadIdx.x == 0) C->relout = sqrt(a[0]) / sqrt(C->nrmup2);
}

// update_mlp bookkeeping and continuation (:1034-1083, :1222-1249)
__global__ void k_outer_scalars(FCtl *C, const double *p4, int nb4, const double *p5, int nb5) {
    double a[1], b[1];
    reduce_parts<1>(p4, nb4, a);
    reduce_parts<1>(p5, nb5, b);
    if (threadIdx.x == 0) outer_step(C, a[0], b[0]);
}

__global__ void k_set_dd(FCtl *C, const double *p, int nb) {
    double a[1];
    reduce_parts<1>(p, nb, a);
    if (threadIdx.x == 0) C->dd = a[0];
}

// ------------------------------------------------------------------ scaleb
// scaleb_gpu (3d/tval3_gpu_3d.cpp:225-266) on the device: the entries of
// smallest and largest magnitude, the first such entry on ties as in the
// reference's forward scan, and whether b holds NaN or Inf.  Block partials,
// then one workgroup; (|b|, index) pairs form a total order, so the result
// does not depend on the grid.
struct BRange {
    float amin, amax;
    int64_t imin, imax;
    int32_t bad;
};

__device__ __forceinline__ void brange_merge(BRange &a, const BRange &b) {
    if (b.amin < a.amin || (b.amin == a.amin && b.imin < a.imin)) { a.amin = b.amin; a.imin = b.imin; }
    if (b.amax > a.amax || (b.amax == a.amax && b.imax < a.imax)) { a.amax = b.amax; a.imax = b.imax; }
    a.bad |= b.bad;
}

__device__ __forceinline__ BRange brange_block(BRange v) {
    __shared__ BRange sh[256];
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) brange_merge(sh[t], sh[t + off]);
        __syncthreads();
    }
    return sh[0];
}

constexpr int kScalebBlocks = 512;

__global__ __launch_bounds__(256) void k_scaleb_parts(int64_t m, const float *__restrict__ b,
                                                      BRange *part) {
    BRange v{INFINITY, -1.f, INT64_MAX, INT64_MAX, 0};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * 256) {
        const float x = b[i], a = fabsf(x);
        if (!isfinite(x)) { v.bad = 1; continue; }
        if (a < v.amin) { v.amin = a; v.imin = i; }     // i grows: first index kept
        if (a > v.amax) { v.amax = a; v.imax = i; }
    }
    const BRange r = brange_block(v);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// range of the selected entries scaled to 0.5 (below 0.5) or 1.5 (above 1.5)
__global__ __launch_bounds__(256) void k_scaleb_final(const float *__restrict__ b,
                                                      const BRange *part, int nb, FCtl *C) {
    BRange v{INFINITY, -1.f, INT64_MAX, INT64_MAX, 0};
    for (int j = threadIdx.x; j < nb; j += 256) brange_merge(v, part[j]);
    const BRange r = brange_block(v);
    if (threadIdx.x == 0) {
        C->b_bad = r.bad;
        float s = 1.f;
        if (!r.bad) {
            const float dif = fabsf(b[r.imax] - b[r.imin]);
            if (dif < 0.5f) s = 0.5f / dif;
            else if (dif > 1.5f) s = 1.5f / dif;
        }
        C->scl = s;
    }
}

// y = scl x (scaling b, :1122-1124; a warm-start U) and y = x / scl (the
// result, :1275-1277) with scl from the control block
__global__ __launch_bounds__(256) void k_scale_ctl(int64_t n, float *__restrict__ y,
                                                   const float *__restrict__ x, const FCtl *C,
                                                   int divide) {
    const float s = (float)C->scl;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        y[i] = divide ? x[i] / s : s * x[i];
}

static unsigned ew_grid(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 8192));
}

// ------------------------------------------------------------------ host side
// workgroups of the launch that runs the A x epilogue (block partial sums)
int product_blocks(const FusedLayout &F) {
    if (F.panels_A || F.dyn) return seg_sum_blocks(F.RS);
    const SparseRows &S = F.A;
    if (S.blk_nnz) return (int)S.nblk;
    const int L = S.lanes ? S.lanes : 32;
    return (int)((S.rows * L + 255) / 256);
}

// The solver's state for one layout, kept in FusedLayout::session between
// solves: buffers, control block, partials and the two instantiated graphs
// are made by the first solve and reused by later ones with the same
// context, A^T-skip setting, graph size and tile setting.
template <int B>
struct Fused : FusedSession {
    tv3_ctx *ctx = nullptr;
    FusedLayout *F = nullptr;
    const tv3_opts *o = nullptr;
    hipStream_t s = nullptr;
    // what the captured graphs depend on
    bool k_skip = false;
    int k_graph_iters = 0, k_tiles = 0;
    size_t hist_cap = 0;             // history rows the buffer holds
    bool captured = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    DevBuf<BRange> brange;
    Lay L;
    int64_t n, m;
    int nbS, nbN, nbNd, nbM, nbW, nbD;
    int ntiles = 0, tpw = 1, tpwD = 1;   // 8x8x8 tiles, tiles per update_W / descend workgroup
    Groups G;
    DevBuf<float> pool;
    float *U[2], *g[2], *Au[2], *d, *sx, *sy, *sz, *ox, *oy, *oz, *dts, *Atb, *rcd, *b, *delta,
        *lastAu, *tmp;
    bool skipping = false;           // A^T-skip heuristic on (skipMulRatio != 0)
    DevBuf<float> tmpbuf;
    DevBuf<double> parts, hist;
    double *pD, *pN, *pW, *pDd, *pAd, *pS4, *pM5, *pOD, *pRO, *pK;
    DevBuf<FCtl> ctl;
    PinnedBuf<FCtl> hctl;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_iters = 8;

    void alloc() {
        auto pad = [](int64_t k) { return (k + 63) / 64 * 64; };
        const int nv = 2 + 2 + 1 + 3 + 3 + 1 + 1 + 1 + 1, mv = 2 + 2 + 1 + 1;
        pool.alloc((size_t)nv * pad(n) + (size_t)mv * pad(m));
        TV3_HIP(hipMemsetAsync(pool.p, 0, pool.bytes(), s));
        float *p = pool.p;
        float **nvp[] = {&U[0], &U[1], &g[0], &g[1], &d, &sx, &sy, &sz, &ox, &oy, &oz, &dts, &Atb,
                         &rcd};
        for (float **q : nvp) { *q = p; p += pad(n); }
        float **mvp[] = {&Au[0], &Au[1], &b, &delta, &lastAu};
        for (float **q : mvp) { *q = p; p += pad(m); }
        tmpbuf.alloc(n > m ? n : m);
        tmp = tmpbuf.p;
        G = groups_of<B>(L);
        TV3_REQUIRE(G.count < (int64_t(1) << 31), TV3_E_SHAPE,
                    "volume too large for the fused solver's voxel groups");
        nbS = (int)stencil_grid(G);
        nbN = product_blocks(*F);
        // update_W with B = 8: tpw consecutive 8x8x8 tiles per workgroup, as
        // many workgroups as are resident at once (one round, nbW <= nbS)
        ntiles = (int)(n / 512);
        if (B == 8) {
            auto tiles_per_wg = [&](const void *kernel) {
                int per_cu = 0;
                TV3_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 512, 0));
                const int resident = std::max(1, per_cu) * std::max(1, ctx->num_cus);
                return std::max(1, (ntiles + resident - 1) / resident);
            };
            tpw = tiles_per_wg(reinterpret_cast<const void *>(k_update_w_tile));
            tpwD = tiles_per_wg(reinterpret_cast<const void *>(k_descend_tile));
            if (o->tiles_per_wg > 0) tpw = tpwD = o->tiles_per_wg;
        }
        nbW = B == 8 ? (ntiles + tpw - 1) / tpw : nbS;
        nbD = B == 8 ? (ntiles + tpwD - 1) / tpwD : nbS;
        nbM = 2048;
        const size_t tot = 3 * (size_t)nbS + 3 * (size_t)nbN + 7 * (size_t)nbS + (size_t)nbS +
                           (size_t)nbN + (size_t)nbS + nbM + nbM + nbM + 2 * (size_t)nbM;
        parts.alloc(tot);
        double *q = parts.p;
        pD = q; q += 3 * (size_t)nbS;
        pN = q; q += 3 * (size_t)nbN;
        pW = q; q += 7 * (size_t)nbS;
        pDd = q; q += nbS;
        pAd = q; q += nbN;
        pS4 = q; q += nbS;
        pM5 = q; q += nbM;
        pOD = q; q += nbM;
        pRO = q; q += nbM;
        pK = q; q += 2 * (size_t)nbM;
        ctl.alloc(1);
        hctl.alloc(1);
        brange.alloc(kScalebBlocks);
        for (auto &e : ev) TV3_HIP(hipEventCreate(&e));
    }

    bool matches(const tv3_ctx *c, bool skip, int gi, int tiles) const {
        return c == ctx && c->stream == s && skip == k_skip && gi == k_graph_iters &&
               tiles == k_tiles;
    }

    // history rows for maxit iterations; a new buffer invalidates the graphs
    void ensure_hist(int maxit) {
        const size_t rows = (size_t)std::max(maxit, 1);
        if (rows <= hist_cap) return;
        drop_graphs();
        hist.alloc(rows * HF);
        hist_cap = rows;
    }

    void drop_graphs() {
        for (auto &gx : graph)
            if (gx) { (void)hipGraphExecDestroy(gx); gx = nullptr; }
        captured = false;
    }

    void read_ctl() {
        TV3_HIP(hipMemcpyAsync(hctl.p, ctl.p, sizeof(FCtl), hipMemcpyDeviceToHost, s));
        TV3_HIP(hipStreamSynchronize(s));
    }
    void write_ctl() {
        TV3_HIP(hipMemcpyAsync(ctl.p, hctl.p, sizeof(FCtl), hipMemcpyHostToDevice, s));
    }

    // one inner iteration from U[c] into U[1-c]; mode 1 = accept (fallback)
    void iteration(int c, int mode) {
        const int n_ = 1 - c;
        const dim3 gS(nbS);
        if (B == 8)
            hipLaunchKernelGGL(k_descend_tile, dim3(nbD), dim3(512), 0, s, L, ntiles, tpwD, U[c], d,
                               sx, sy, sz, ox, oy, oz, U[n_], ctl.p, pD);
        else
            hipLaunchKernelGGL(k_descend<B>, gS, dim3(256), 0, s, L, G, U[c], d, sx, sy, sz, ox,
                               oy, oz, U[n_], ctl.p, pD);
        launch_forward<float>(*F, U[n_], EpiRes{Au[n_], b, delta, Au[c], ctl.p, pN, {0, 0, 0}}, s,
                              &ctl.p->active);
        // A^T Au_n - Atb; with the panel form of A^T the line search runs in
        // an extra workgroup of that launch
        auto gradient = [&](int gated) {
            const EpiGrad e{g[n_], Atb, ctl.p, gated};
            if (F->panels_T) launch_panel(F->PT, Au[n_], EpiGradLS{e, ctl.p, pD, pN, nbD, nbN, mode}, s);
            else launch_transposed<float>(*F, Au[n_], e, s);
        };
        if (skipping) {                // get_g's A^T-skip decision (:398-468)
            hipLaunchKernelGGL(k_skip_stats, dim3(nbM), dim3(256), 0, s, m, Au[n_], lastAu, ctl.p,
                               pK);
            hipLaunchKernelGGL(k_skip_decide, dim3(1), dim3(1024), 0, s, ctl.p, pK, nbM, mode);
            gradient(1);
            hipLaunchKernelGGL(k_skip_after, dim3(nbM), dim3(256), 0, s, n, m, g[c], g[n_], Au[n_],
                               lastAu, ctl.p);
        } else {
            gradient(0);
        }
        if (!F->panels_T)
            hipLaunchKernelGGL(k_line_search_parts, dim3(1), dim3(1024), 0, s, ctl.p, pD, nbD, pN,
                               nbN, mode);
        hipLaunchKernelGGL(k_apply_alpha, dim3(512), dim3(256), 0, s, n, m, U[c], U[n_], g[c],
                           g[n_], Au[c], Au[n_], ctl.p);
        update_w(U[n_], U[c], g[n_]);
        hipLaunchKernelGGL(k_iter_end_parts, dim3(1), dim3(1024), 0, s, ctl.p, pW, nbW, hist.p);
    }

    void update_w(const float *Un, const float *Ucur, const float *gn) {
        if (B == 8)
            hipLaunchKernelGGL(k_update_w_tile, dim3(nbW), dim3(512), 0, s, L, ntiles, tpw, Un,
                               Ucur, sx, sy, sz, gn, dts, d, ctl.p, pW);
        else
            hipLaunchKernelGGL(k_update_w<B>, dim3(nbS), dim3(256), 0, s, L, G, Un, Ucur, sx, sy,
                               sz, gn, dts, d, ctl.p, pW);
    }

    void tau_first(int fallback) {
        hipLaunchKernelGGL(k_dd_norm<B>, dim3(nbS), dim3(256), 0, s, L, G, d, pDd);
        launch_forward<float>(*F, d, EpiNorm{ctl.p, pAd, {0}}, s);
        hipLaunchKernelGGL(k_tau_first, dim3(1), dim3(1024), 0, s, ctl.p, pDd, nbS, pAd, nbN,
                           fallback);
    }

    void capture(int c) {
        hipGraph_t gr;
        TV3_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < graph_iters; ++k) iteration((c + k) & 1, 0);
        TV3_HIP(hipStreamEndCapture(s, &gr));
        TV3_HIP(hipGraphInstantiate(&graph[c], gr, nullptr, nullptr, 0));
        TV3_HIP(hipGraphDestroy(gr));
    }

    ~Fused() override {
        if (s) (void)hipStreamSynchronize(s);
        drop_graphs();
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// The session of this layout for these options: the kept one when it
// matches, else a new one (buffers allocated, graphs captured on first use).
template <int B>
Fused<B> &fused_session(tv3_ctx *ctx, FusedLayout &FL, const tv3_opts &o) {
    const bool skip = o.skipMulRatio != 0.0;
    const int gi = o.graph_iters > 0 ? o.graph_iters : 8;
    auto *cur = dynamic_cast<Fused<B> *>(FL.session.get());
    if (cur && cur->matches(ctx, skip, gi, o.tiles_per_wg)) {
        cur->o = &o;
        return *cur;
    }
    FL.session.reset();
    std::unique_ptr<Fused<B>> S(new Fused<B>());
    S->ctx = ctx; S->F = &FL; S->o = &o; S->s = ctx->stream;
    S->L.p = FL.dims[0]; S->L.q = FL.dims[1]; S->L.r = FL.dims[2];
    S->L.nby = FL.nby; S->L.nbx = FL.nbx; S->L.nbz = S->L.r / 8;
    S->L.pq = (int64_t)S->L.p * S->L.q;
    S->n = (int64_t)S->L.p * S->L.q * S->L.r;
    S->m = FL.A.rows;
    S->graph_iters = gi;
    S->skipping = skip;
    S->k_skip = skip; S->k_graph_iters = gi; S->k_tiles = o.tiles_per_wg;
    S->alloc();
    Fused<B> &r = *S;
    FL.session = std::move(S);
    return r;
}

template <int B>
void run_fused(tv3_ctx *ctx, FusedLayout &FL, const tv3_opts &o, const float *b_dev, float *U_out,
               const float *Ut_dev, tv3_info *info, tv3_history *hist) {
    Fused<B> &S = fused_session<B>(ctx, FL, o);
    hipStream_t s = S.s;
    const int64_t n = S.n, m = S.m;
    S.ensure_hist(o.maxit);
    hipEvent_t ev0 = S.ev[0], ev1 = S.ev[1], evl0 = S.ev[2], evl1 = S.ev[3];
    TV3_HIP(hipEventRecord(ev0, s));

    FCtl &H = *S.hctl.p;
    std::memset(&H, 0, sizeof(FCtl));
    H.mu = std::min(o.mu0, o.mu); H.beta = std::min(o.beta0, o.beta); H.muDbeta = H.mu / H.beta;
    H.c = o.c; H.gamma = o.gamma; H.rate_gam = o.rate_gam; H.tol_inn = o.tol_inn;
    H.gam_opt = o.gam; H.gam = o.gam; H.rate_cnt = o.rate_cnt; H.mu_max = o.mu; H.beta_max = o.beta;
    H.maxit = o.maxit; H.nonneg = o.nonneg; H.iso = o.tv_mode == TV3_TV_ISO;
    H.hist_cap = o.maxit; H.active = 1; H.alpha = 1; H.scl = 1;
    H.skip_ratio = o.skipMulRatio; H.max_rel = o.maxRelativeChange;
    H.n_init = (int32_t)std::min<uint32_t>(o.numInitIterations, INT32_MAX);
    S.write_ctl();

    // multipliers and the DtsAtd term start at zero (:1128-1139): sigma, the
    // old sigma and DtsAtd are consecutive in the pool
    TV3_HIP(hipMemsetAsync(S.sx, 0, (size_t)(S.Atb - S.sx) * sizeof(float), s));
    TV3_HIP(hipMemsetAsync(S.delta, 0, sizeof(float) * m, s));

    // scaleb (:225-266) on the device; b into the internal ray order
    hipLaunchKernelGGL(k_scaleb_parts, dim3(kScalebBlocks), dim3(256), 0, s, m, b_dev,
                       S.brange.p);
    hipLaunchKernelGGL(k_scaleb_final, dim3(1), dim3(256), 0, s, b_dev, S.brange.p,
                       kScalebBlocks, S.ctl.p);
    permute_gather(b_dev, FL.r_int2nat.p, S.tmp, m, s);
    hipLaunchKernelGGL(k_scale_ctl, dim3(ew_grid(m)), dim3(256), 0, s, m, S.b, S.tmp, S.ctl.p, 0);

    // initialisation (:1128-1153)
    launch_transposed<float>(FL, S.b, EpiStore<float>{S.Atb}, s);
    if (o.warm_start) {
        permute_gather(U_out, FL.v_int2nat.p, S.tmp, n, s);
        hipLaunchKernelGGL(k_scale_ctl, dim3(ew_grid(n)), dim3(256), 0, s, n, S.U[0], S.tmp,
                           S.ctl.p, 0);
    } else {
        TV3_HIP(hipMemcpyAsync(S.U[0], S.Atb, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    }
    TV3_HIP(hipMemcpyAsync(S.rcd, S.U[0], sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    launch_forward<float>(FL, S.U[0],
                          EpiRes{S.Au[0], S.b, S.delta, S.Au[0], S.ctl.p, S.pN, {0, 0, 0}}, s);
    launch_transposed<float>(FL, S.Au[0], EpiGrad{S.g[0], S.Atb, S.ctl.p, 0}, s);
    if (S.skipping)                  // the initial get_g multiplies (forced): last Au
        TV3_HIP(hipMemcpyAsync(S.lastAu, S.Au[0], sizeof(float) * m, hipMemcpyDeviceToDevice, s));
    S.update_w(S.U[0], S.U[0], S.g[0]);
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(1024), 0, s, S.ctl.p, S.pN, S.nbN, S.pW, S.nbW);
    TV3_HIP(hipGetLastError());

    if (!S.captured) {
        S.capture(0);
        S.capture(1);
        S.captured = true;
    }
    TV3_HIP(hipEventRecord(evl0, s));

    // U[c] holds the accepted iterate; every completed iteration flips c
    int c = 0, outer = 0;
    S.read_ctl();
    TV3_REQUIRE(!H.b_bad, TV3_E_NONFINITE, "b holds NaN or Inf");
    while (outer < o.maxcnt && H.total < o.maxit) {
        int prev = H.total;
        S.tau_first(0);              // first iteration of the inner loop
        S.iteration(c, 0);
        S.read_ctl();
        for (;;) {
            const int done = H.total - prev;
            if (done & 1) c ^= 1;
            prev = H.total;
            if (H.need_fallback) {   // cnt == 5: restart from U[c] (:876-883)
                S.tau_first(1);
                S.iteration(c, 1);
                S.read_ctl();
                continue;
            }
            if (H.inner_done || H.nonfinite || H.total >= o.maxit) break;
            // a whole graph, or the last iterations maxit leaves launched
            // one by one (a graph would run its tail as no-op launches)
            const int left = o.maxit - H.total;
            if (left >= S.graph_iters) {
                TV3_HIP(hipGraphLaunch(S.graph[c], s));
            } else {
                for (int k = 0; k < left; ++k) S.iteration((c + k) & 1, 0);
            }
            S.read_ctl();
        }
        if (H.nonfinite) break;
        ++outer;
        hipLaunchKernelGGL(k_relout, dim3(S.nbM), dim3(256), 0, s, n, S.U[c], S.rcd, S.pRO);
        hipLaunchKernelGGL(k_relout_scalar, dim3(1), dim3(1024), 0, s, S.ctl.p, S.pRO, S.nbM);
        S.read_ctl();
        if (H.relout < o.tol) break;
        // the loop ends here: the multiplier and penalty updates would not be
        // read (the result, counts and rel_chg are final), so they are skipped
        if (outer >= o.maxcnt || H.total >= o.maxit) break;
        hipLaunchKernelGGL(k_mlp_sigma<B>, dim3(S.nbS), dim3(256), 0, s, S.L, S.G, S.U[c], S.sx,
                           S.sy, S.sz, S.ox, S.oy, S.oz, S.ctl.p, S.pS4);
        hipLaunchKernelGGL(k_mlp_delta, dim3(S.nbM), dim3(256), 0, s, m, S.Au[c], S.b, S.delta,
                           S.ctl.p, S.pM5);
        hipLaunchKernelGGL(k_outer_scalars, dim3(1), dim3(1024), 0, s, S.ctl.p, S.pS4, S.nbS, S.pM5,
                           S.nbM);
        hipLaunchKernelGGL(k_outer_d, dim3(S.nbM), dim3(256), 0, s, n, S.d, S.g[c], S.dts, S.ctl.p,
                           S.pOD);
        hipLaunchKernelGGL(k_set_dd, dim3(1), dim3(1024), 0, s, S.ctl.p, S.pOD, S.nbM);
        S.read_ctl();
    }
    TV3_HIP(hipEventRecord(evl1, s));

    // U (natural order) = U_int / scl (:1275-1277)
    hipLaunchKernelGGL(k_scale_ctl, dim3(ew_grid(n)), dim3(256), 0, s, n, S.tmp, S.U[c], S.ctl.p,
                       1);
    permute_scatter(S.tmp, FL.v_int2nat.p, U_out, n, s);
    TV3_HIP(hipEventRecord(ev1, s));
    S.read_ctl();
    float ms = 0, msl = 0;
    TV3_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    TV3_HIP(hipEventElapsedTime(&msl, evl0, evl1));
    info->total_iters = H.total;
    info->outer_iters = outer;
    info->rel_chg = H.relout;
    info->secs = ms / 1000.0;
    info->loop_secs = msl / 1000.0;
    info->backtracks = H.backtracks;
    info->fallbacks = H.fallbacks;
    // setup A U; per completed iteration and per rejected trial (fallback)
    // one A; one A d per first-iteration step length (inner loops started
    // and fallbacks, counted where tau is formed)
    info->n_matvec_N = 1 + (int64_t)H.total + H.fallbacks + H.n_first;
    // with skipping: A^T b, the initial product and the products get_g took
    info->n_matvec_T = S.skipping ? 2 + (int64_t)H.n_T : 2 + (int64_t)H.total + H.fallbacks;
    info->status = H.nonfinite ? TV3_E_NONFINITE : TV3_OK;
    info->rel_error = 0;
    if (Ut_dev) {
        auto e = diff_norm<float>(ctx, n, Ut_dev, U_out, s);
        info->rel_error = std::sqrt(e[0]) / std::sqrt(e[1]);
    }
    if (hist && hist->cap > 0) {
        const int cnt = std::min(H.total, hist->cap);
        std::vector<double> hh((size_t)std::max(cnt, 1) * HF);
        if (cnt > 0)
            TV3_HIP(hipMemcpy(hh.data(), S.hist.p, sizeof(double) * HF * cnt, hipMemcpyDeviceToHost));
        for (int k = 0; k < cnt; ++k) {
            const double *h = &hh[(size_t)k * HF];
            hist->f[k] = h[0]; hist->lam1[k] = h[1]; hist->lam2[k] = h[2]; hist->lam3[k] = h[3];
            hist->lam4[k] = h[4]; hist->lam5[k] = h[5]; hist->tau[k] = h[6]; hist->alpha[k] = h[7];
            hist->relchg[k] = h[8]; hist->mu[k] = h[9]; hist->beta[k] = h[10]; hist->C[k] = h[11];
            hist->cnt[k] = (int32_t)h[12]; hist->outer[k] = (int32_t)h[13];
            if (hist->mulT) hist->mulT[k] = (int32_t)h[14];
            if (hist->f_trial) hist->f_trial[k] = h[15];
        }
        hist->count = cnt;
    }
}

}  // namespace

void solve_fused(tv3_ctx *ctx, tv3_op *op, const tv3_opts &o, const int32_t dims[3],
                 const float *b_dev, float *U_dev, const float *Ut_dev, tv3_info *info,
                 tv3_history *hist) {
    FusedLayout &FL = fused_layout(op, dims);
    if (FL.B == 8) run_fused<8>(ctx, FL, o, b_dev, U_dev, Ut_dev, info, hist);
    else if (FL.B == 4) run_fused<4>(ctx, FL, o, b_dev, U_dev, Ut_dev, info, hist);
    else run_fused<1>(ctx, FL, o, b_dev, U_dev, Ut_dev, info, hist);
}

}  // namespace tv3
