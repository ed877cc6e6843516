// dyntile.hip -- the tile layout of a matrix-free Siddon operator (DynTiles):
// built once when the operator is created, from the traversal's own
// sparsity pattern; the products (panels.hpp: k_dyn_tile_N, k_dyn_tile_T)
// recompute every weight.
//
// Boxes: the 8^3 tiles of the fused solver's voxel numbering (B = 8), or
// their 4^3 bricks where a tile holds more than kDynEmax entries or
// kDynSmax rays (and every brick when B = 4).  A box is convex, so a ray's
// entries inside it are consecutive in traversal order: one segment.
// Segments are numbered box by box, rays ascending inside a box; a box's
// entries are its segments' entries in that order, and every voxel keeps
// the positions of its entries there (ray order).
#include <algorithm>
#include <numeric>

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

std::unique_ptr<FusedLayout> build_dyn_layout(tv3_op *op) {
    const DynGeom &G = op->geo;
    if (G.mode != TV3_WEIGHT_SIDDON) return nullptr;
    const int32_t dims[3] = {G.rv_y, G.rv_x, G.rv_z};
    if (dims[0] % 4 || dims[1] % 4 || dims[2] % 4) return nullptr;
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
    const int B = L->B, nby = L->nby, nbx = L->nbx;   // tiles (B = 8) or bricks (B = 4)
    const int64_t nunit = n / (B == 8 ? 512 : 64);
    // boxes: one per tile, or its 8 bricks where the tile is too full
    std::vector<int32_t> unit_ent(nunit, 0), unit_seg(nunit, 0);
    const int ushift = B == 8 ? 9 : 6;
    for (int64_t r = 0; r < m; ++r) {
        int32_t last = -1;
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const int32_t u = ind[k] >> ushift;
            ++unit_ent[u];
            if (u != last) ++unit_seg[u];
            last = u;
        }
    }
    std::vector<DynPan> pan;
    std::vector<int32_t> box_of_brick((size_t)(n >> 6), -1);
    for (int64_t u = 0; u < nunit; ++u) {
        int ty = 0, tx = 0, tz = 0;
        ty = (int)(u % nby);
        tx = (int)((u / nby) % nbx);
        tz = (int)(u / ((int64_t)nby * nbx));
        if (B == 8 && unit_ent[u] <= kDynEmax && unit_seg[u] <= kDynSmax) {
            pan.push_back(DynPan{(int32_t)(u << 9), 8 * tx, 8 * ty, 8 * tz, 8, 0, 0, 0});
            for (int k = 0; k < 8; ++k) box_of_brick[(u << 3) + k] = (int32_t)pan.size() - 1;
        } else if (B == 8) {
            for (int sb = 0; sb < 8; ++sb) {   // sub-brick bits: y (1), x (2), z (4)
                pan.push_back(DynPan{(int32_t)((u << 9) + (sb << 6)), 8 * tx + 4 * ((sb >> 1) & 1),
                                     8 * ty + 4 * (sb & 1), 8 * tz + 4 * ((sb >> 2) & 1), 4, 0, 0,
                                     0});
                box_of_brick[(u << 3) + sb] = (int32_t)pan.size() - 1;
            }
        } else {
            pan.push_back(DynPan{(int32_t)(u << 6), 4 * tx, 4 * ty, 4 * tz, 4, 0, 0, 0});
            box_of_brick[u] = (int32_t)pan.size() - 1;
        }
    }
    const int64_t npan = (int64_t)pan.size();
    // segments: runs of one box along each ray (a box is entered once)
    std::vector<int32_t> nseg_box(npan + 1, 0);
    for (int64_t r = 0; r < m; ++r) {
        int32_t last = -1;
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const int32_t b = box_of_brick[ind[k] >> 6];
            if (b != last) ++nseg_box[b + 1];
            last = b;
        }
    }
    for (int64_t b = 0; b < npan; ++b) nseg_box[b + 1] += nseg_box[b];
    const int64_t nseg = nseg_box[npan];
    std::vector<int32_t> seg_ray(nseg > 0 ? nseg : 1), seg_len(nseg > 0 ? nseg : 1, 0),
        fill(nseg_box.begin(), nseg_box.end() - 1), seg_of_entry(nnz > 0 ? nnz : 1);
    std::vector<int32_t> rseg_cnt(m + 1, 0);
    for (int64_t r = 0; r < m; ++r) {
        int32_t last = -1, cur = -1;
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const int32_t b = box_of_brick[ind[k] >> 6];
            if (b != last) {
                TV3_REQUIRE(fill[b] == nseg_box[b] || seg_ray[fill[b] - 1] != (int32_t)r,
                            TV3_E_INTERNAL, "a ray enters a box twice");
                cur = fill[b]++;
                seg_ray[cur] = (int32_t)r;
                ++rseg_cnt[r + 1];
            }
            ++seg_len[cur];
            seg_of_entry[k] = cur;
            last = b;
        }
    }
    // entry offsets inside each box; positions of every voxel's entries
    std::vector<int32_t> seg_eoff(nseg > 0 ? nseg : 1, 0);
    int32_t emax = 0, smax = 0;
    for (int64_t b = 0; b < npan; ++b) {
        int32_t e = 0;
        for (int32_t sg = nseg_box[b]; sg < nseg_box[b + 1]; ++sg) {
            seg_eoff[sg] = e;
            e += seg_len[sg];
        }
        pan[b].s0 = nseg_box[b];
        pan[b].s1 = nseg_box[b + 1];
        pan[b].ne = e;
        emax = std::max(emax, e);
        smax = std::max(smax, nseg_box[b + 1] - nseg_box[b]);
    }
    TV3_REQUIRE(emax <= kDynEmax && smax <= kDynSmax, TV3_E_UNSUPPORTED,
                "a 4^3 brick of the volume holds more ray entries than the tile products keep in "
                "LDS");
    std::vector<uint32_t> vcnt(n + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) ++vcnt[ind[k] + 1];
    std::vector<uint32_t> pptr(n + 1, 0);
    for (int64_t v = 0; v < n; ++v) pptr[v + 1] = pptr[v] + ((vcnt[v + 1] + 3) & ~3u);
    TV3_REQUIRE(pptr[n] < (1u << 31), TV3_E_UNSUPPORTED, "too many padded entries");
    std::vector<uint16_t> pos(pptr[n] > 0 ? pptr[n] : 4, 0);
    std::vector<uint32_t> vfill(pptr.begin(), pptr.end() - 1);
    // the j-th entry of segment sg sits at seg_eoff[sg] + j of its box; rays
    // ascending, so every voxel's list is in ray order
    for (int64_t r = 0; r < m; ++r) {
        int32_t j = 0, last = -1;
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const int32_t sg = seg_of_entry[k];
            j = sg == last ? j + 1 : 0;
            last = sg;
            pos[vfill[ind[k]]++] = (uint16_t)(seg_eoff[sg] + j);
        }
    }
    for (int64_t b = 0; b < npan; ++b) {              // padding: the box's zero slot ne
        const int nv = pan[b].bs * pan[b].bs * pan[b].bs;
        for (int32_t v = pan[b].v0; v < pan[b].v0 + nv; ++v)
            for (uint32_t q = vfill[v]; q < pptr[v + 1]; ++q) pos[q] = (uint16_t)pan[b].ne;
    }
    // the second pass of A x: each ray's segments, ascending
    for (int64_t r = 0; r < m; ++r) rseg_cnt[r + 1] += rseg_cnt[r];
    std::vector<int32_t> rseg(nseg > 0 ? nseg : 1), rfill(rseg_cnt.begin(), rseg_cnt.end() - 1);
    for (int64_t sg = 0; sg < nseg; ++sg) rseg[rfill[seg_ray[sg]]++] = (int32_t)sg;
    // ray end points in internal order
    std::vector<float> rays(6 * (m > 0 ? m : 1), 0.f);
    for (int64_t i = 0; i < m; ++i) {
        const int32_t r = order[i], e = G.path_h[2 * r], c = G.path_h[2 * r + 1];
        for (int a = 0; a < 3; ++a) {
            rays[6 * i + a] = G.epos_h[3 * e + a];
            rays[6 * i + 3 + a] = G.rpos_h[3 * c + a];
        }
    }
    // boxes in decreasing work
    std::vector<int32_t> bord(npan);
    std::iota(bord.begin(), bord.end(), 0);
    std::stable_sort(bord.begin(), bord.end(), [&](int32_t a, int32_t b) {
        return pan[a].ne + 16 * (pan[a].s1 - pan[a].s0) > pan[b].ne + 16 * (pan[b].s1 - pan[b].s0);
    });
    std::unique_ptr<DynTiles> D(new DynTiles());
    D->npan = npan; D->nseg = nseg; D->nnz = nnz; D->emax = emax; D->smax = smax;
    for (int a = 0; a < 3; ++a) { D->org[a] = G.origin[a]; D->h[a] = G.voxel[a]; }
    D->N[0] = G.rv_x; D->N[1] = G.rv_y; D->N[2] = G.rv_z;
    upload_vec(ctx, D->pan, pan);
    upload_vec(ctx, D->order, bord);
    upload_vec(ctx, D->seg_ray, seg_ray);
    upload_vec(ctx, D->seg_eoff, seg_eoff);
    upload_vec(ctx, D->rays, rays);
    upload_vec(ctx, D->pptr, pptr);
    upload_vec(ctx, D->pos, pos);
    const int64_t big = n > m ? n : m;
    D->xin.alloc(big > 0 ? big : 1);
    D->yout.alloc(big > 0 ? big : 1);
    SegSum &R = L->RS;
    R.rows = m;
    R.nseg = nseg;
    const double per_ray = m > 0 ? (double)nseg / (double)m : 0.0;
    R.lanes = per_ray > 24 ? 16 : per_ray > 12 ? 8 : 4;
    upload_vec(ctx, R.ptr, rseg_cnt);
    upload_vec(ctx, R.seg, rseg);
    R.part.alloc(nseg > 0 ? nseg : 1);
    TV3_HIP(hipStreamSynchronize(ctx->stream));
    // the matrices themselves are not kept: sizes only
    for (SparseRows *S : {&L->A, &L->AT}) {
        S->ptr.release(); S->ind.release(); S->val.release(); S->desc.release();
        S->nblk = 0; S->blk_nnz = 0;
    }
    L->npan = npan;
    L->nseg = nseg;
    L->max_seg = smax;
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
