// dyntile.hip -- the tile layout of a matrix-free Siddon operator (DynTiles):
// built once when the operator is created, from the traversal's own
// sparsity pattern; the products (panels.hpp: k_dyn_tile_N, k_dyn_tile_T)
// recompute every weight.
//
// Boxes: for A x, blocks of 2 x 2 x 2 of the fused solver's 8^3 tiles
// (16^3 voxels) where every tile count is even, else pairs of tiles
// consecutive in y (or single tiles, or 4^3 bricks when B = 4); for A^T r,
// the 8^3 tiles, or
// their 4^3 bricks where a tile holds more ray entries than the A^T product
// keeps in LDS (and every brick when B = 4).  A box is convex, so a ray's
// entries inside it are consecutive in traversal order: one segment.
// Segments are numbered box by box, inside a box by decreasing length in
// buckets, ascending ray in a bucket (sort_segments); an A^T box's entries
// are its segments' entries in that order, and every voxel keeps the
// positions of its entries there (ray order).
#include <algorithm>
#include <numeric>

// build-time tuning of the A^T boxes: workgroup size, and the entries one
// box keeps in LDS where a workgroup may hold 160 KiB (15360 otherwise).
// 512 threads and 19456 entries (76 KiB) let two workgroups share a CU, one
// walking while the other sums: measured at 128^3 with 262144 rays, 238 us
// against 302 us for one 1024-thread workgroup holding 32768 entries.
#ifndef TV3_DYN_T_THREADS
#define TV3_DYN_T_THREADS 512
#endif
#ifndef TV3_DYN_EMAX_BIG
#define TV3_DYN_EMAX_BIG 19456
#endif
// work items of the A x boxes of 16^3 voxels: this many times the segments
// and entries of a two-tile box's item (each item stages its box's 4,096
// voxels).  Measured at 128^3: A 85-87 us with 1, 77 with 2, 73 with 4, 77
// with 6, 83 with 8 (two-tile boxes: 81-83 us).
#ifndef TV3_DYN_CHUNK16
#define TV3_DYN_CHUNK16 4
#endif

#include "panels.hpp"

namespace tv3 {

namespace {

std::vector<int32_t> download_i32(tv3_ctx *ctx, const DevBuf<int32_t> &d, int64_t count) {
    std::vector<int32_t> h(count > 0 ? count : 1);
    if (count > 0)
        TV3_HIP(hipMemcpyAsync(h.data(), d.p, sizeof(int32_t) * count, hipMemcpyDeviceToHost,
                               ctx->stream));
    TV3_HIP(hipStreamSynchronize(ctx->stream));
    return h;
}

template <typename T>
void upload_vec(tv3_ctx *ctx, DevBuf<T> &d, const std::vector<T> &h) {
    d.alloc(h.empty() ? 1 : h.size());
    if (!h.empty())
        TV3_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice,
                               ctx->stream));
}

}  // namespace

struct RayGrid {
    float org[3], h[3];
    int N[3];
};

// traversal record of every internal ray (the layout dyn_seg_walk reads:
// A[3], L, B[3], amin, amax, 3 unused)
__global__ void k_ray_pre(int64_t m, const float *__restrict__ ends, RayGrid G,
                          float *__restrict__ rec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    float p0[3], p1[3];
    for (int a = 0; a < 3; ++a) { p0[a] = ends[6 * i + a]; p1[a] = ends[6 * i + 3 + a]; }
    SidRay<float> s;
    sid_setup(s, p0, p1, G.org, G.h, G.N);
    float *o = rec + 12 * i;
    for (int a = 0; a < 3; ++a) { o[a] = s.A[a]; o[4 + a] = s.B[a]; }
    o[3] = s.L; o[7] = s.amin; o[8] = s.amax; o[9] = 0.f; o[10] = 0.f; o[11] = 0.f;
}

namespace {

// Segments of the boxes box_of (per internal brick) along the traversal CSR:
// numbered box by box, rays ascending inside a box; fills X (host side) and
// returns each entry's segment and the segment lengths.
struct SegPlan {
    std::vector<int32_t> seg_box_ptr, seg_ray, seg_len, seg_of_entry;
};

SegPlan plan_segments(const std::vector<int32_t> &ptr, const std::vector<int32_t> &ind,
                      const std::vector<int32_t> &box_of, int64_t npan, int64_t m) {
    SegPlan S;
    S.seg_box_ptr.assign(npan + 1, 0);
    for (int64_t r = 0; r < m; ++r) {
        int32_t last = -1;
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const int32_t b = box_of[ind[k] >> 6];
            if (b != last) ++S.seg_box_ptr[b + 1];
            last = b;
        }
    }
    for (int64_t b = 0; b < npan; ++b) S.seg_box_ptr[b + 1] += S.seg_box_ptr[b];
    const int64_t nseg = S.seg_box_ptr[npan];
    const int64_t nnz = ptr[m];
    S.seg_ray.assign(nseg > 0 ? nseg : 1, 0);
    S.seg_len.assign(nseg > 0 ? nseg : 1, 0);
    S.seg_of_entry.assign(nnz > 0 ? nnz : 1, 0);
    std::vector<int32_t> fill(S.seg_box_ptr.begin(), S.seg_box_ptr.end() - 1);
    for (int64_t r = 0; r < m; ++r) {
        int32_t last = -1, cur = -1;
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const int32_t b = box_of[ind[k] >> 6];
            if (b != last) {
                TV3_REQUIRE(fill[b] == S.seg_box_ptr[b] || S.seg_ray[fill[b] - 1] != (int32_t)r,
                            TV3_E_INTERNAL, "a ray enters a box twice");
                cur = fill[b]++;
                S.seg_ray[cur] = (int32_t)r;
            }
            ++S.seg_len[cur];
            S.seg_of_entry[k] = cur;
            last = b;
        }
    }
    return S;
}

// boxes: tile u (B = 8) whole or as its 8 bricks, or brick u (B = 4)
void add_boxes(std::vector<DynPan> &pan, std::vector<int32_t> &box_of, int B, int64_t u, int nby,
               int nbx, bool split) {
    const int ty = (int)(u % nby), tx = (int)((u / nby) % nbx),
              tz = (int)(u / ((int64_t)nby * nbx));
    if (B == 8 && !split) {
        pan.push_back(DynPan{(int32_t)(u << 9), 8 * tx, 8 * ty, 8 * tz, 8, 8, 0, 0, 0});
        for (int k = 0; k < 8; ++k) box_of[(u << 3) + k] = (int32_t)pan.size() - 1;
    } else if (B == 8) {
        for (int sb = 0; sb < 8; ++sb) {   // sub-brick bits: y (1), x (2), z (4)
            pan.push_back(DynPan{(int32_t)((u << 9) + (sb << 6)), 8 * tx + 4 * ((sb >> 1) & 1),
                                 8 * ty + 4 * (sb & 1), 8 * tz + 4 * ((sb >> 2) & 1), 4, 4, 0, 0,
                                 0});
            box_of[(u << 3) + sb] = (int32_t)pan.size() - 1;
        }
    } else {
        pan.push_back(DynPan{(int32_t)(u << 6), 4 * tx, 4 * ty, 4 * tz, 4, 4, 0, 0, 0});
        box_of[u] = (int32_t)pan.size() - 1;
    }
}

// the A x box of tiles u and u + 1, consecutive in y (ty even): 16 x 8 x 8
// voxels, contiguous in the internal order
void add_pair_box(std::vector<DynPan> &pan, std::vector<int32_t> &box_of, int64_t u, int nby,
                  int nbx) {
    const int ty = (int)(u % nby), tx = (int)((u / nby) % nbx),
              tz = (int)(u / ((int64_t)nby * nbx));
    pan.push_back(DynPan{(int32_t)(u << 9), 8 * tx, 8 * ty, 8 * tz, 8, 16, 0, 0, 0});
    for (int k = 0; k < 16; ++k) box_of[(u << 3) + k] = (int32_t)pan.size() - 1;
}

// the A x box of the 2 x 2 x 2 tiles from tile u (tile coordinates all even):
// 16^3 voxels in 8 tiles, at u, u + 1 (y), u + nby (x), u + nby nbx (z) and
// their sums
void add_box16(std::vector<DynPan> &pan, std::vector<int32_t> &box_of, int64_t u, int nby,
               int nbx) {
    const int ty = (int)(u % nby), tx = (int)((u / nby) % nbx),
              tz = (int)(u / ((int64_t)nby * nbx));
    pan.push_back(DynPan{(int32_t)(u << 9), 8 * tx, 8 * ty, 8 * tz, 16, 16, 0, 0, 0});
    for (int sb = 0; sb < 8; ++sb) {
        const int64_t t = u + (sb & 1) + ((sb >> 1) & 1) * (int64_t)nby +
                          (sb >> 2) * (int64_t)nby * nbx;
        for (int k = 0; k < 8; ++k) box_of[(t << 3) + k] = (int32_t)pan.size() - 1;
    }
}

// box-local coordinates of internal voxel v in box P (numbering B = 8 or 4,
// nby / nbx tiles or bricks along y / x)
void box_local(int32_t v, const DynPan &P, int B, int nby, int nbx, int &x, int &y, int &z) {
    const int sh = B == 8 ? 9 : 6, e = B == 8 ? 8 : 4;
    const int64_t u = v >> sh;
    int lx, ly, lz;
    dyn_local_xyz(v & ((1 << sh) - 1), lx, ly, lz);
    const int uy = (int)(u % nby), ux = (int)((u / nby) % nbx), uz = (int)(u / ((int64_t)nby * nbx));
    x = e * ux + lx - P.bx;
    y = e * uy + ly - P.by;
    z = e * uz + lz - P.bz;
}

// Inside every box, segments by decreasing length in buckets of q entries,
// ascending ray inside a bucket: the threads of a wavefront walk segments of
// about the same length, and neighbouring threads read the records of
// neighbouring rays (rays are numbered in bundles that cross the same
// boxes).  Measured at 128^3 (A x, A^T r in us): q = 1: 92, 180; q = 4: 84,
// 179-182; q = 8: 80.6, 182-183; ray order alone: 88, 184.
void sort_segments(SegPlan &S, int64_t npan, int q) {
    const int64_t nseg = S.seg_box_ptr[npan];
    std::vector<int32_t> newid(nseg > 0 ? nseg : 1), ids;
    for (int64_t b = 0; b < npan; ++b) {
        const int32_t p0 = S.seg_box_ptr[b], p1 = S.seg_box_ptr[b + 1];
        ids.resize(p1 - p0);
        std::iota(ids.begin(), ids.end(), p0);
        std::stable_sort(ids.begin(), ids.end(), [&](int32_t a, int32_t c) {
            return S.seg_len[a] / q > S.seg_len[c] / q;
        });
        for (int32_t k = 0; k < p1 - p0; ++k) newid[ids[k]] = p0 + k;
    }
    std::vector<int32_t> ray(S.seg_ray.size()), len(S.seg_len.size());
    for (int64_t sg = 0; sg < nseg; ++sg) {
        ray[newid[sg]] = S.seg_ray[sg];
        len[newid[sg]] = S.seg_len[sg];
    }
    S.seg_ray.swap(ray);
    S.seg_len.swap(len);
    for (auto &e : S.seg_of_entry) e = newid[e];
}

// box-local coordinates of each segment's first voxel (x | y << 4 | z << 8)
std::vector<uint16_t> segment_starts(const std::vector<int32_t> &ptr,
                                     const std::vector<int32_t> &ind, const SegPlan &S,
                                     const std::vector<int32_t> &box_of,
                                     const std::vector<DynPan> &pan, int64_t m, int B, int nby,
                                     int nbx) {
    const int64_t nseg = S.seg_box_ptr.back();
    std::vector<uint16_t> c0(nseg > 0 ? nseg : 1, 0);
    for (int64_t r = 0; r < m; ++r) {
        int32_t last = -1;
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const int32_t sg = S.seg_of_entry[k];
            if (sg == last) continue;
            last = sg;
            const DynPan &P = pan[box_of[ind[k] >> 6]];
            int x, y, z;
            box_local(ind[k], P, B, nby, nbx, x, y, z);
            TV3_REQUIRE(x >= 0 && y >= 0 && z >= 0 && x < P.bs && y < P.ey && z < P.bs,
                        TV3_E_INTERNAL, "a segment starts outside its box");
            c0[sg] = (uint16_t)(x | (y << 4) | (z << 8));
        }
    }
    return c0;
}

// upload one direction's boxes (segment ranges, entry offsets, first voxels,
// work order)
void finish_boxes(tv3_ctx *ctx, std::vector<DynPan> &pan, const SegPlan &S,
                  const std::vector<uint16_t> &c0, bool offsets, DynBoxes &X,
                  std::vector<int32_t> *eoff_out) {
    const int64_t npan = (int64_t)pan.size();
    std::vector<int32_t> eoff(S.seg_ray.size(), 0);
    X.emax = 0;
    X.smax = 0;
    for (int64_t b = 0; b < npan; ++b) {
        int32_t e = 0;
        for (int32_t sg = S.seg_box_ptr[b]; sg < S.seg_box_ptr[b + 1]; ++sg) {
            eoff[sg] = e;
            e += S.seg_len[sg];
        }
        pan[b].s0 = S.seg_box_ptr[b];
        pan[b].s1 = S.seg_box_ptr[b + 1];
        pan[b].ne = e;
        X.emax = std::max(X.emax, e);
        X.smax = std::max(X.smax, S.seg_box_ptr[b + 1] - S.seg_box_ptr[b]);
    }
    if (!offsets) {
        // A x: a box's segments in chunks of at most 512 segments and about
        // 6,144 entries (TV3_DYN_CHUNK16 times that in boxes of 16^3 voxels),
        // one workgroup each (it stages the box's voxels, and each segment's
        // partial sum is its own), so that the heaviest boxes do not run long
        // chains after the others have ended
        std::vector<DynPan> items;
        for (const DynPan &P : pan) {
            const int scale = P.bs == 16 ? TV3_DYN_CHUNK16 : 1;
            int32_t sg = P.s0;
            while (sg < P.s1) {
                DynPan c = P;
                c.s0 = sg;
                int32_t e = 0;
                while (sg < P.s1 && sg - c.s0 < 512 * scale && e < 6144 * scale)
                    e += S.seg_len[sg++];
                c.s1 = sg;
                c.ne = e;
                items.push_back(c);
            }
        }
        pan.swap(items);
    }
    const int64_t nitem = (int64_t)pan.size();
    std::vector<int32_t> bord(nitem);
    std::iota(bord.begin(), bord.end(), 0);
    std::stable_sort(bord.begin(), bord.end(), [&](int32_t a, int32_t b) {
        return pan[a].ne + 16 * (pan[a].s1 - pan[a].s0) > pan[b].ne + 16 * (pan[b].s1 - pan[b].s0);
    });
    X.npan = nitem;
    X.nseg = S.seg_box_ptr[npan];
    upload_vec(ctx, X.pan, pan);
    upload_vec(ctx, X.order, bord);
    upload_vec(ctx, X.seg_ray, S.seg_ray);
    upload_vec(ctx, X.seg_c0, c0);
    if (offsets) upload_vec(ctx, X.seg_eoff, eoff);
    if (eoff_out) *eoff_out = std::move(eoff);
}

}  // namespace

std::unique_ptr<FusedLayout> build_dyn_layout(tv3_op *op) {
    const DynGeom &G = op->geo;
    if (G.mode != TV3_WEIGHT_SIDDON) return nullptr;
    const int32_t dims[3] = {G.rv_y, G.rv_x, G.rv_z};
    if (dims[0] % 4 || dims[1] % 4 || dims[2] % 4) return nullptr;
    if (G.rv_x > 1024 || G.rv_y > 1024 || G.rv_z > 1024) return nullptr;   // 10-bit indices
    tv3_ctx *ctx = op->ctx;
    std::unique_ptr<FusedLayout> L(new FusedLayout());
    const std::vector<int32_t> ray_nat = geometry_ray_order(G);
    {
        SparseRows A;
        dynamic_materialize(op, TV3_F32, A);
        renumber_rows(ctx, A, ray_nat, dims, L.get());
    }
    const int64_t m = L->A.rows, n = L->AT.rows, nnz = L->A.nnz;
    TV3_REQUIRE(nnz < (int64_t(1) << 31), TV3_E_UNSUPPORTED, "more than 2^31 ray-voxel entries");
    // internal CSR of the traversal: rows = internal rays, entries in walk order
    const std::vector<int32_t> ptr = download_i32(ctx, L->A.ptr, m + 1);
    const std::vector<int32_t> ind = download_i32(ctx, L->A.ind, nnz);
    std::vector<int32_t> order = ray_nat;
    if ((int64_t)order.size() != m) {
        order.resize(m);
        std::iota(order.begin(), order.end(), 0);
    }
    std::unique_ptr<DynTiles> D(new DynTiles());
    D->nnz = nnz;
    for (int a = 0; a < 3; ++a) { D->org[a] = G.origin[a]; D->h[a] = G.voxel[a]; }
    D->N[0] = G.rv_x; D->N[1] = G.rv_y; D->N[2] = G.rv_z;
    // A^T boxes keep their entries and rays in LDS: the larger allocations of
    // devices that allow 160 KiB per workgroup hold whole tiles
    const size_t lds = std::min<size_t>(ctx->lds_per_block, size_t(160) << 10);
    D->emax_cap = lds >= (size_t(150) << 10) ? TV3_DYN_EMAX_BIG : 15360;
    D->threads_T = TV3_DYN_T_THREADS;
    const int B = L->B, nby = L->nby, nbx = L->nbx;   // tiles (B = 8) or bricks (B = 4)
    const int64_t nunit = n / (B == 8 ? 512 : 64);
    const int ushift = B == 8 ? 9 : 6;
    std::vector<int32_t> unit_ent(nunit, 0);
    for (int64_t k = 0; k < nnz; ++k) ++unit_ent[ind[k] >> ushift];
    // forward boxes: whole units; backward: split where the LDS limits ask
    std::vector<DynPan> panF, panT;
    std::vector<int32_t> boxF((size_t)(n >> 6), -1), boxT((size_t)(n >> 6), -1);
    // A x: blocks of 2 x 2 x 2 tiles where every tile count is even (9.1
    // segments per ray at 128^3 against 14.9 for pairs: 73 us against 81-83),
    // else tiles in pairs
    // along y where the tile count along y is even (half the segments, so
    // half the partial sums of the second pass), worked in chunks
    // (finish_boxes).  Measured at 128^3: pairs 92.6 us against 100-102 us
    // for single tiles, 104 us for single tiles in chunks.
    const int nbz = B == 8 ? dims[2] / 8 : 0;
    const bool box16 = B == 8 && nby % 2 == 0 && nbx % 2 == 0 && nbz % 2 == 0;
    const bool pairs = B == 8 && nby % 2 == 0;
    for (int64_t u = 0; u < nunit; ++u) {
        const int64_t ty = u % nby, tx = (u / nby) % nbx, tz = u / ((int64_t)nby * nbx);
        if (box16) {
            if (ty % 2 == 0 && tx % 2 == 0 && tz % 2 == 0) add_box16(panF, boxF, u, nby, nbx);
        } else if (!pairs) {
            add_boxes(panF, boxF, B, u, nby, nbx, false);
        } else if (u % 2 == 0) {
            add_pair_box(panF, boxF, u, nby, nbx);
        }
        add_boxes(panT, boxT, B, u, nby, nbx, unit_ent[u] > D->emax_cap);
    }
    if (B == 8) {
        D->tile_dx = nby * 512;
        D->tile_dz = nby * nbx * 512;
    }
    SegPlan SF = plan_segments(ptr, ind, boxF, (int64_t)panF.size(), m);
    SegPlan ST = plan_segments(ptr, ind, boxT, (int64_t)panT.size(), m);
    sort_segments(SF, (int64_t)panF.size(), 8);
    sort_segments(ST, (int64_t)panT.size(), 4);
    finish_boxes(ctx, panF, SF, segment_starts(ptr, ind, SF, boxF, panF, m, B, nby, nbx), false,
                 D->fwd, nullptr);
    std::vector<int32_t> eoffT;
    finish_boxes(ctx, panT, ST, segment_starts(ptr, ind, ST, boxT, panT, m, B, nby, nbx), true,
                 D->bwd, &eoffT);
    TV3_REQUIRE(D->bwd.emax <= D->emax_cap, TV3_E_UNSUPPORTED,
                "a 4^3 brick of the volume holds more ray entries than the tile products keep in "
                "LDS");
    // positions of every voxel's entries in its backward box; each voxel's
    // list in ascending ray order (the fixed order of its sum)
    std::vector<uint32_t> vcnt(n + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) ++vcnt[ind[k] + 1];
    std::vector<uint32_t> pptr(n + 1, 0);
    for (int64_t v = 0; v < n; ++v) pptr[v + 1] = pptr[v] + ((vcnt[v + 1] + 3) & ~3u);
    TV3_REQUIRE(pptr[n] < (1u << 31), TV3_E_UNSUPPORTED, "too many padded entries");
    std::vector<uint16_t> pos(pptr[n] > 0 ? pptr[n] : 4, 0);
    std::vector<uint32_t> vfill(pptr.begin(), pptr.end() - 1);
    // the j-th entry of segment sg sits at eoff[sg] + j of its box
    for (int64_t r = 0; r < m; ++r) {
        int32_t j = 0, last = -1;
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const int32_t sg = ST.seg_of_entry[k];
            j = sg == last ? j + 1 : 0;
            last = sg;
            pos[vfill[ind[k]]++] = (uint16_t)(eoffT[sg] + j);
        }
    }
    for (const DynPan &P : panT) {                    // padding: the box's zero slot ne
        const int nv = P.ey * P.bs * P.bs;
        for (int32_t v = P.v0; v < P.v0 + nv; ++v)
            for (uint32_t q = vfill[v]; q < pptr[v + 1]; ++q) pos[q] = (uint16_t)P.ne;
    }
    upload_vec(ctx, D->pptr, pptr);
    upload_vec(ctx, D->pos, pos);
    // the second pass of A x: each ray's forward segments, ascending
    const int64_t nsegF = D->fwd.nseg;
    std::vector<int32_t> rptr(m + 1, 0), rseg(nsegF > 0 ? nsegF : 1);
    for (int64_t sg = 0; sg < nsegF; ++sg) ++rptr[SF.seg_ray[sg] + 1];
    for (int64_t r = 0; r < m; ++r) rptr[r + 1] += rptr[r];
    std::vector<int32_t> rfill(rptr.begin(), rptr.end() - 1);
    for (int64_t sg = 0; sg < nsegF; ++sg) rseg[rfill[SF.seg_ray[sg]]++] = (int32_t)sg;
    SegSum &R = L->RS;
    R.rows = m;
    R.nseg = nsegF;
    const double per_ray = m > 0 ? (double)nsegF / (double)m : 0.0;
    R.lanes = per_ray > 24 ? 16 : per_ray > 12 ? 8 : 4;
    upload_vec(ctx, R.ptr, rptr);
    upload_vec(ctx, R.seg, rseg);
    R.part.alloc(nsegF > 0 ? nsegF : 1);
    // ray end points in internal order -> traversal records on the device
    std::vector<float> ends(6 * (m > 0 ? m : 1), 0.f);
    for (int64_t i = 0; i < m; ++i) {
        const int32_t r = order[i], e = G.path_h[2 * r], c = G.path_h[2 * r + 1];
        for (int a = 0; a < 3; ++a) {
            ends[6 * i + a] = G.epos_h[3 * e + a];
            ends[6 * i + 3 + a] = G.rpos_h[3 * c + a];
        }
    }
    {
        DevBuf<float> dends;
        upload_vec(ctx, dends, ends);
        D->ray.alloc(12 * (m > 0 ? m : 1));
        RayGrid rg;
        for (int a = 0; a < 3; ++a) { rg.org[a] = D->org[a]; rg.h[a] = D->h[a]; rg.N[a] = D->N[a]; }
        if (m > 0)
            hipLaunchKernelGGL(k_ray_pre, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                               ctx->stream, m, dends.p, rg, D->ray.p);
        TV3_HIP(hipGetLastError());
        TV3_HIP(hipStreamSynchronize(ctx->stream));
    }
    const int64_t big = n > m ? n : m;
    D->xin.alloc(big > 0 ? big : 1);
    D->yout.alloc(big > 0 ? big : 1);
    TV3_HIP(hipStreamSynchronize(ctx->stream));
    // the matrices themselves are not kept: sizes only
    for (SparseRows *S : {&L->A, &L->AT}) {
        S->ptr.release(); S->ind.release(); S->val.release(); S->desc.release();
        S->nblk = 0; S->blk_nnz = 0;
    }
    L->npan = D->bwd.npan;
    L->nseg = D->bwd.nseg;
    L->max_seg = D->bwd.smax;
    L->dyn = std::move(D);
    return L;
}

bool dyn_tile_apply(tv3_op *op, bool trans, const float *x, float *y, hipStream_t s) {
    FusedLayout *F = op->fused.get();
    if (!F || !F->dyn) return false;
    DynTiles &D = *F->dyn;
    const int64_t n = op->n, m = op->m;
    if (!trans) {
        permute_gather(x, F->v_int2nat.p, D.xin.p, n, s);
        launch_dyn_N<float>(D, D.xin.p, F->RS.part.p, nullptr, s);
        launch_seg_sum(F->RS, EpiStoreIf{D.yout.p, nullptr}, s);
        permute_scatter(D.yout.p, F->r_int2nat.p, y, m, s);
    } else {
        permute_gather(x, F->r_int2nat.p, D.xin.p, m, s);
        launch_dyn_T<float>(D, D.xin.p, EpiStoreIf{D.yout.p, nullptr}, s);
        permute_scatter(D.yout.p, F->v_int2nat.p, y, n, s);
    }
    TV3_HIP(hipGetLastError());
    return true;
}

}  // namespace tv3
