// layout.hip -- the fused solver's internal numbering of voxels and rays.
//
// Scattered gathers of x (A x) and of the residual (A^T r) are bound by the
// L1/L2 request rate, so the solver works in a numbering that puts
// neighbours in the same cache lines: voxels in 8x8x8 tiles of 4x4x4 bricks
// (brick_index) and rays ordered by emitter and receiver position.  A^T and
// A get panel forms whose sources are staged in LDS (panels.hpp).
// Inputs and outputs are permuted at the boundary of a solve only.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "panels.hpp"

namespace tv3 {

static std::vector<int32_t> spatial_rank(const std::vector<float> &pos) {
    const int64_t k = (int64_t)pos.size() / 3;
    double c[3] = {0, 0, 0};
    for (int64_t i = 0; i < k; ++i)
        for (int a = 0; a < 3; ++a) c[a] += pos[3 * i + a] / (double)k;
    const int bands = std::max(1, (int)std::lround(std::sqrt((double)k)));
    std::vector<double> th(k), ph(k);
    std::vector<int> bd(k);
    for (int64_t i = 0; i < k; ++i) {
        double x = pos[3 * i] - c[0], y = pos[3 * i + 1] - c[1], z = pos[3 * i + 2] - c[2];
        double rr = std::sqrt(x * x + y * y + z * z);
        th[i] = rr > 0 ? std::acos(std::max(-1.0, std::min(1.0, z / rr))) : 0.0;
        ph[i] = std::atan2(y, x);
        bd[i] = std::min(bands - 1, (int)(th[i] / M_PI * bands));
    }
    std::vector<int32_t> ord(k);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
        if (bd[a] != bd[b]) return bd[a] < bd[b];
        double ka = (bd[a] & 1) ? -ph[a] : ph[a], kb = (bd[b] & 1) ? -ph[b] : ph[b];
        return ka < kb;
    });
    std::vector<int32_t> rank(k);
    for (int64_t i = 0; i < k; ++i) rank[ord[i]] = (int32_t)i;
    return rank;
}

// Rays in bundles: 4 emitters x 4 receivers neighbouring in spatial rank
// form a thin tube whose 16 rays cross nearly the same tiles, so a tile's
// rays (ascending internal number) come in runs that share cache lines, in
// A^T's staging of the residual and in the per-ray sums of A's segment
// partials.  At 128^3 with 256 x 1024 rays this cuts the 64-byte lines those
// gathers touch from 0.39 to 0.29 (A^T) and from 0.34 to 0.23 (A) per
// segment against ordering by emitter, then receiver.
std::vector<int32_t> geometry_ray_order(const DynGeom &g) {
    std::vector<int32_t> order;
    const int64_t P = g.num_paths;
    if (P == 0 || g.epos_h.empty() || g.rpos_h.empty() || (int64_t)g.path_h.size() != 2 * P)
        return order;
    std::vector<int32_t> er = spatial_rank(g.epos_h), rr = spatial_rank(g.rpos_h);
    constexpr int kBundle = 4;
    order.resize(P);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        const int32_t ea = g.path_h[2 * a], eb = g.path_h[2 * b];
        const int32_t ma = ea / g.E, mb = eb / g.E;
        if (ma != mb) return ma < mb;
        const int32_t xa = er[ea], xb = er[eb];
        const int32_t ya = rr[g.path_h[2 * a + 1]], yb = rr[g.path_h[2 * b + 1]];
        if (xa / kBundle != xb / kBundle) return xa / kBundle < xb / kBundle;
        if (ya / kBundle != yb / kBundle) return ya / kBundle < yb / kBundle;
        if (xa != xb) return xa < xb;
        return ya < yb;
    });
    return order;
}

__global__ void k_vox_maps(int p, int q, int r, int B, int nby, int nbx, int32_t *nat2int,
                           int32_t *int2nat) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)p * q * r;
    if (v >= n) return;
    const int y = (int)(v % p), x = (int)((v / p) % q), z = (int)(v / ((int64_t)p * q));
    const int64_t i = brick_index(B, (int64_t)p * q, p, nby, nbx, y, x, z);
    nat2int[v] = (int32_t)i;
    int2nat[i] = (int32_t)v;
}

__global__ void k_perm_rowlen(int64_t m, const int32_t *rmap, const int32_t *ptr, int32_t *len) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int32_t r = rmap[i];
    len[i] = ptr[r + 1] - ptr[r];
}

template <int VB>
__global__ void k_perm_rows(int64_t m, const int32_t *rmap, const int32_t *ptr,
                            const int32_t *ind, const uint8_t *val, const int32_t *nptr,
                            const int32_t *cmap, int32_t *nind, uint8_t *nval) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int32_t r = rmap[i];
    const int b = ptr[r], e = ptr[r + 1], o = nptr[i];
    for (int k = 0; k < e - b; ++k) {
        nind[o + k] = cmap[ind[b + k]];
#pragma unroll
        for (int t = 0; t < VB; ++t) nval[(int64_t)(o + k) * VB + t] = val[(int64_t)(b + k) * VB + t];
    }
}

__global__ void k_gather(const float *in, const int32_t *map, float *out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[map[i]];
}
__global__ void k_scatter_f(const float *in, const int32_t *map, float *out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[map[i]] = in[i];
}

void permute_gather(const float *in, const int32_t *map, float *out, int64_t n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, map,
                           out, n);
}
void permute_scatter(const float *in, const int32_t *map, float *out, int64_t n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_scatter_f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in,
                           map, out, n);
}

// Largest source list of a panel: its values take at most 48 KiB of LDS,
// within the 64 KiB a workgroup may hold.
constexpr int kPanelSegCap = 12288;

// Host copy of a CSR (here: the CSC of A, rows = internal voxels).
struct HostCSR {
    int64_t rows = 0, cols = 0, nnz = 0;
    size_t vb = 0;
    tv3_dtype vtype = TV3_F32;
    std::vector<int32_t> ptr, ind;
    std::vector<uint8_t> val;
};

static HostCSR download_csr(tv3_ctx *ctx, const SparseRows &S) {
    HostCSR h;
    h.rows = S.rows; h.cols = S.cols; h.nnz = S.nnz; h.vtype = S.vtype;
    h.vb = dtype_size(S.vtype);
    hipStream_t s = ctx->stream;
    h.ptr.resize(S.rows + 1);
    h.ind.resize(S.nnz > 0 ? S.nnz : 1);
    h.val.resize((S.nnz > 0 ? S.nnz : 1) * h.vb);
    TV3_HIP(hipMemcpyAsync(h.ptr.data(), S.ptr.p, sizeof(int32_t) * (S.rows + 1),
                           hipMemcpyDeviceToHost, s));
    if (S.nnz > 0) {
        TV3_HIP(hipMemcpyAsync(h.ind.data(), S.ind.p, sizeof(int32_t) * S.nnz,
                               hipMemcpyDeviceToHost, s));
        TV3_HIP(hipMemcpyAsync(h.val.data(), S.val.p, h.vb * S.nnz, hipMemcpyDeviceToHost, s));
    }
    TV3_HIP(hipStreamSynchronize(s));
    return h;
}

// Tiles of the voxel rows of A^T: panel p = rows [pan_row[p], pan_row[p+1])
// (a unit of `unit` rows, split at brick boundaries where it holds more than
// pan_nnz entries or more than seg_cap distinct rays: kPanelSegCap, what LDS
// holds, or 4096 for 12-bit local indices) and its sorted distinct rays
// seg_ray[pan_seg[p] .. pan_seg[p+1]).
struct PanelPlan {
    std::vector<int32_t> pan_row{0}, pan_seg{0}, seg_ray;
    int64_t max_seg = 0;
};

static bool plan_panels(const HostCSR &h, int unit, int pan_nnz, PanelPlan &pl,
                        int64_t seg_cap = kPanelSegCap) {
    const int64_t n = h.rows, m = h.cols;
    const std::vector<int32_t> &ptr = h.ptr, &ind = h.ind;
    const int64_t sub = unit >= 64 ? 64 : unit;   // split granularity: one brick
    std::vector<int32_t> stamp(m > 0 ? m : 1, -1), rays, fresh;
    int32_t cur = 0;
    int64_t cur_nnz = 0;
    auto close = [&](int64_t row_end) {
        std::sort(rays.begin(), rays.end());
        pl.seg_ray.insert(pl.seg_ray.end(), rays.begin(), rays.end());
        pl.max_seg = std::max<int64_t>(pl.max_seg, (int64_t)rays.size());
        pl.pan_seg.push_back((int32_t)pl.seg_ray.size());
        pl.pan_row.push_back((int32_t)row_end);
        rays.clear();
        cur_nnz = 0;
        ++cur;
    };
    for (int64_t u0 = 0; u0 < n; u0 += unit) {
        const int64_t u1 = std::min<int64_t>(n, u0 + unit);
        for (int64_t c0 = u0; c0 < u1; c0 += sub) {
            const int64_t c1 = std::min<int64_t>(u1, c0 + sub);
            const int64_t cn = (int64_t)ptr[c1] - ptr[c0];
            for (int pass = 0; pass < 2; ++pass) {
                fresh.clear();
                for (int64_t k = ptr[c0]; k < ptr[c1]; ++k) {
                    const int32_t r = ind[k];
                    if (stamp[r] != cur) { stamp[r] = cur; fresh.push_back(r); }
                }
                const bool empty = (int64_t)pl.pan_row.back() == c0;
                if (empty || (cur_nnz + cn <= pan_nnz &&
                              (int64_t)(rays.size() + fresh.size()) <= seg_cap))
                    break;
                for (int32_t r : fresh) stamp[r] = -1;   // the chunk opens a new panel
                close(c0);
            }
            if ((int64_t)(rays.size() + fresh.size()) > seg_cap) return false;
            rays.insert(rays.end(), fresh.begin(), fresh.end());
            cur_nnz += cn;
        }
        close(u1);
    }
    return true;
}

// Row blocks per work item of a panel launch (one workgroup each): A's
// panels in chunks of 6 (their sources, a contiguous voxel range, are cheap
// to stage again); A^T's whole (each chunk would gather the panel's rays
// again).  Measured at 128^3 (A, A^T in us): whole 66.0, 56.1; chunks of 3
// 65.0, 64.3; of 6 64.7, 58.7; of 9 67.7, 56.2.
#ifndef TV3_PANEL_CHUNK_A
#define TV3_PANEL_CHUNK_A 6
#endif
constexpr int kPanelChunkA = TV3_PANEL_CHUNK_A, kPanelChunkT = 1 << 30;
// A's packed 16^3 panels (rows in pairs): chunk_blocks 0 = the fewest
// blocks per work item whose items all fit one round of resident
// workgroups.  A second, nearly empty round or an under-filled first one
// costs more than the x staging of smaller items: measured at 128^3 with row
// blocks of 2048 / 4096 entries (8 / 6 resident workgroups per CU), chunks
// of 3 60.0 / 53.7 us, 6 56.0 / 57.2, 9 61.7 / 50.1, 12 51.7 / 52.6,
// 16 54.7 / 56.4; the rule picks 12 / 9.
constexpr int kPanelChunkA12 = 0;

// Workgroups of the packed A kernel (rows in groups of G = 2 or 4) resident
// at once on the GPU, with its dynamic LDS for max_src staged sources (the
// FP16 and BF16 instances differ only in the weight decode).
static int64_t packed_A_slots(tv3_ctx *ctx, int blk_nnz, int G, int max_src) {
    const size_t lds = sizeof(float) * (size_t)std::max(max_src, 1);
    const void *k =
        G == 4 ? (blk_nnz == 2048
                      ? reinterpret_cast<const void *>(&k_panel12<2048, 4, __half, EpiStoreIf>)
                      : reinterpret_cast<const void *>(&k_panel12<4096, 4, __half, EpiStoreIf>))
               : (blk_nnz == 2048
                      ? reinterpret_cast<const void *>(&k_panel12<2048, 2, __half, EpiStoreIf>)
                      : reinterpret_cast<const void *>(&k_panel12<4096, 2, __half, EpiStoreIf>));
    int per_cu = 0;
    TV3_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, lds));
    return (int64_t)std::max(1, per_cu) * std::max(1, ctx->num_cus);
}

// Entry groups of the panel rows: A's rows (ray segments, ~8 entries at
// 128^3) unpadded, A^T's (voxels, ~17) in groups of 4.  Measured at 128^3,
// 35.8M FP16 entries, blocks of 2048: A 71.1 / 71.5 / 74.5 us with groups
// of 1 / 2 / 4; A^T 57.6 us with 4, 65.6 us with 2.
#ifndef TV3_PANEL_GROUP_A
#define TV3_PANEL_GROUP_A 1
#endif
#ifndef TV3_PANEL_GROUP_T
#define TV3_PANEL_GROUP_T 4
#endif
constexpr int kPanelGroupA = TV3_PANEL_GROUP_A, kPanelGroupT = TV3_PANEL_GROUP_T;
// A's packed entries: rows padded to pairs.  The block's products then take
// half the LDS (16^3 boxes stage 16 KB of x besides), so more workgroups are
// resident: at 128^3 A takes 57.0 / 56.1 us in row blocks of 4096 / 2048
// entries against 59.0 / 64.0 us with unpadded rows, for about 1.2M padding
// entries (4 MB).  Groups of 4 (8 workgroups per CU also in blocks of 4096,
// 3.6M padding entries): 51.0 against 50.3 us with work items sized to one
// round, equal in blocks of 2048.
#ifndef TV3_PANEL_GROUP_A12
#define TV3_PANEL_GROUP_A12 2
#endif
constexpr int kPanelGroupA12 = TV3_PANEL_GROUP_A12;

// Pad, block and upload a panelled matrix.  Panel p holds rows [pan_row[p],
// pan_row[p+1]) and sources src_off[p] .. src_off[p+1] (indices into *src,
// or the index range itself when src is NULL); row r's entries are
// rptr[r] .. rptr[r+1] with local source index lidx[k] and weight bytes
// wb[k * vb ..].  Rows are padded to multiples of G entries.
static void pack_panels(tv3_ctx *ctx, const std::vector<int32_t> &pan_row,
                        const std::vector<int32_t> &src_off, const std::vector<int32_t> *src,
                        const std::vector<int32_t> &rptr, const std::vector<uint16_t> &lidx,
                        const uint8_t *wb, size_t vb, tv3_dtype vt, int blk_nnz, int G,
                        int chunk_blocks, Panels &P, bool pack12 = false) {
    hipStream_t s = ctx->stream;
    TV3_REQUIRE(G == 1 || G == 2 || G == 4, TV3_E_ARG, "panel groups of 1, 2 or 4 entries");
    TV3_REQUIRE(!pack12 || ((vt == TV3_F16 || vt == TV3_BF16) &&
                            (blk_nnz == 2048 || blk_nnz == 4096)),
                TV3_E_ARG, "packed entries take FP16 or BF16 weights and blocks of 2048 or 4096");
    const int64_t align = pack12 ? 8 : 4;
    const int64_t npan = (int64_t)pan_row.size() - 1, rows = pan_row.back();
    auto plen = [&](int64_t r) { return ((int64_t)rptr[r + 1] - rptr[r] + G - 1) / G * G; };
    // row blocks: consecutive rows of one panel, at most blk_nnz padded
    // entries and 256 rows (a longer row alone); a panel without rows (a tile
    // no ray crosses, in the segment form of A) gets no row block at all
    std::vector<int32_t> blk{0}, pan_blk{0};
    for (int64_t p = 0; p < npan; ++p) {
        const int64_t r0 = pan_row[p], r1 = pan_row[p + 1];
        if (r1 > r0) {
            int64_t start = r0, used = 0;
            for (int64_t r = r0; r < r1; ++r) {
                if (r > start && (used + plen(r) > blk_nnz || r - start >= 256)) {
                    blk.push_back((int32_t)r);
                    start = r;
                    used = 0;
                }
                used += plen(r);
            }
            blk.push_back((int32_t)r1);
        }
        pan_blk.push_back((int32_t)blk.size() - 1);
    }
    // stored positions: a block's padded rows back to back from a 4-aligned
    // start (8-aligned for packed entries)
    std::vector<int32_t> pptr(rows + 1);
    int64_t acc = 0;
    for (size_t b = 0; b + 1 < blk.size(); ++b) {
        for (int64_t r = blk[b]; r < blk[b + 1]; ++r) {
            pptr[r] = (int32_t)acc;
            acc += plen(r);
        }
        acc = (acc + align - 1) & ~(align - 1);
    }
    TV3_REQUIRE(acc < (int64_t(1) << 31), TV3_E_UNSUPPORTED, "panel entries exceed 2^31");
    pptr[rows] = (int32_t)acc;
    int max_src = 0;
    for (int64_t p = 0; p < npan; ++p) max_src = std::max(max_src, src_off[p + 1] - src_off[p]);
    std::vector<uint32_t> pk, pw;
    std::vector<uint16_t> ps;
    std::vector<float> pv;
    const size_t np = acc > align ? (size_t)acc : (size_t)align;   // whole quads (octets)
    if (pack12) { pk.assign(np / 8 * 3, 0u); pw.assign(np / 8 * 4, 0u); }
    else if (vb == 2) pk.assign(np, 0u);
    else { ps.assign(np, 0); pv.assign(np, 0.f); }
    for (int64_t r = 0; r < rows; ++r)
        for (int32_t k = rptr[r]; k < rptr[r + 1]; ++k) {
            const int64_t q = (int64_t)pptr[r] + (k - rptr[r]);
            if (pack12) {
                // entry e of octet o: index at bits 12 e of the octet's 96,
                // weight in half e of its 16 bytes
                uint16_t bits;
                std::memcpy(&bits, wb + (size_t)k * 2, 2);
                TV3_REQUIRE(lidx[k] < 4096, TV3_E_INTERNAL, "a packed local index needs 12 bits");
                const int64_t o = q >> 3;
                const int e = (int)(q & 7), bit = 12 * e, w = bit >> 5, off = bit & 31;
                pk[3 * o + w] |= (uint32_t)lidx[k] << off;
                if (off > 20) pk[3 * o + w + 1] |= (uint32_t)lidx[k] >> (32 - off);
                pw[4 * o + (e >> 1)] |= (uint32_t)bits << (16 * (e & 1));
            } else if (vb == 2) {
                uint16_t bits;
                std::memcpy(&bits, wb + (size_t)k * 2, 2);
                pk[q] = (uint32_t)lidx[k] | ((uint32_t)bits << 16);
            } else {
                ps[q] = lidx[k];
                std::memcpy(&pv[q], wb + (size_t)k * 4, 4);
            }
        }
    if (chunk_blocks <= 0) {
        TV3_REQUIRE(pack12 && (G == 2 || G == 4), TV3_E_INTERNAL,
                    "chunks sized to one round: packed A only");
        const int64_t slots = packed_A_slots(ctx, blk_nnz, G, max_src);
        chunk_blocks = 64;
        for (int c = 3; c < 64; ++c) {
            int64_t items = 0;
            for (int64_t p = 0; p < npan; ++p) items += (pan_blk[p + 1] - pan_blk[p] + c - 1) / c;
            if (items <= slots) { chunk_blocks = c; break; }
        }
    }
    // work items: each panel's row blocks in chunks of at most chunk_blocks,
    // one workgroup each (staging the panel's sources itself), so that the
    // largest panels -- 29 blocks against 9 on average for the two-tile A
    // panels at 128^3 -- do not run long chains after the others have ended;
    // largest first
    std::vector<int32_t> chunk, order;
    std::vector<int64_t> cwork;
    for (int64_t p = 0; p < npan; ++p) {
        const int32_t b0 = pan_blk[p], nb = pan_blk[p + 1] - b0;
        const int32_t nc = (int32_t)(((int64_t)nb + chunk_blocks - 1) / chunk_blocks);
        const int ns = src_off[p + 1] - src_off[p];
        for (int32_t i = 0; i < nc; ++i) {
            const int32_t c0 = b0 + (int32_t)((int64_t)nb * i / nc);
            const int32_t c1 = b0 + (int32_t)((int64_t)nb * (i + 1) / nc);
            chunk.insert(chunk.end(), {(int32_t)p, c0, c1});
            cwork.push_back((int64_t)pptr[blk[c1]] - pptr[blk[c0]] + 4 * (int64_t)ns);
        }
    }
    order.resize(cwork.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return cwork[a] > cwork[b]; });
    P.rows = rows;
    P.npan = npan;
    P.nsrc = src ? (int64_t)src->size() : 0;
    P.max_src = max_src;
    P.blk_nnz = blk_nnz;
    P.group = G;
    P.fmt = pack12 ? 12 : 0;
    P.stored = (int64_t)np;
    P.contiguous = src == nullptr;
    P.vtype = vt;
    auto up = [&](DevBuf<int32_t> &d, const std::vector<int32_t> &hv) {
        d.alloc(hv.empty() ? 1 : hv.size());
        if (!hv.empty())
            TV3_HIP(hipMemcpyAsync(d.p, hv.data(), sizeof(int32_t) * hv.size(),
                                   hipMemcpyHostToDevice, s));
    };
    up(P.order, order);
    up(P.src_off, src_off);
    up(P.chunk, chunk);
    P.nitem = (int64_t)order.size();
    if (src) up(P.src, *src);
    else P.src.release();
    up(P.pptr, pptr);
    const std::vector<BlkDesc> d = make_blk_desc(blk, pptr.data(), nullptr);
    P.desc.alloc(d.size());
    TV3_HIP(hipMemcpyAsync(P.desc.p, d.data(), sizeof(BlkDesc) * d.size(), hipMemcpyHostToDevice,
                           s));
    P.pw.release();
    if (pack12) {
        P.pk.alloc(pk.size());
        P.pw.alloc(pw.size());
        P.ps.release();
        P.pv.release();
        TV3_HIP(hipMemcpyAsync(P.pk.p, pk.data(), sizeof(uint32_t) * pk.size(),
                               hipMemcpyHostToDevice, s));
        TV3_HIP(hipMemcpyAsync(P.pw.p, pw.data(), sizeof(uint32_t) * pw.size(),
                               hipMemcpyHostToDevice, s));
    } else if (vb == 2) {
        P.pk.alloc(pk.size());
        P.ps.release();
        P.pv.release();
        TV3_HIP(hipMemcpyAsync(P.pk.p, pk.data(), sizeof(uint32_t) * pk.size(),
                               hipMemcpyHostToDevice, s));
    } else {
        P.pk.release();
        P.ps.alloc(ps.size());
        P.pv.alloc(pv.size());
        TV3_HIP(hipMemcpyAsync(P.ps.p, ps.data(), sizeof(uint16_t) * ps.size(),
                               hipMemcpyHostToDevice, s));
        TV3_HIP(hipMemcpyAsync(P.pv.p, pv.data(), sizeof(float) * pv.size(),
                               hipMemcpyHostToDevice, s));
    }
    TV3_HIP(hipStreamSynchronize(s));
}

// A^T in panel form: rows = voxels, sources = the tile's rays.
// pack12: packed entries (the plan's panels hold at most 4096 rays)
static void build_panels_T(tv3_ctx *ctx, const HostCSR &h, const PanelPlan &pl, int blk_nnz,
                           Panels &P, bool pack12 = false) {
    std::vector<int32_t> local(h.cols > 0 ? h.cols : 1, 0);
    std::vector<uint16_t> lidx(h.nnz > 0 ? h.nnz : 1);
    const int64_t npan = (int64_t)pl.pan_row.size() - 1;
    for (int64_t p = 0; p < npan; ++p) {
        for (int32_t j = pl.pan_seg[p]; j < pl.pan_seg[p + 1]; ++j)
            local[pl.seg_ray[j]] = j - pl.pan_seg[p];
        for (int64_t k = h.ptr[pl.pan_row[p]]; k < h.ptr[pl.pan_row[p + 1]]; ++k)
            lidx[k] = (uint16_t)local[h.ind[k]];
    }
    pack_panels(ctx, pl.pan_row, pl.pan_seg, &pl.seg_ray, h.ptr, lidx, h.val.data(), h.vb,
                h.vtype, blk_nnz, kPanelGroupT, kPanelChunkT, P, pack12);
}

// A in panel form: rows = ray segments (segment s = pan_seg[p] + position of
// its ray in panel p's list), sources = the panel's voxel rows; and the
// per-ray segment lists of the second pass.
static void build_panels_A(tv3_ctx *ctx, const HostCSR &h, const PanelPlan &pl, int blk_nnz,
                           Panels &P, SegSum &R, int chunk_blocks = kPanelChunkA,
                           const std::vector<int32_t> *tmap = nullptr, bool pack12 = false) {
    const int64_t npan = (int64_t)pl.pan_row.size() - 1, nseg = (int64_t)pl.seg_ray.size();
    const int64_t m = h.cols;
    std::vector<int32_t> local(m > 0 ? m : 1, 0), sptr(nseg + 1, 0);
    // entries per segment
    for (int64_t p = 0; p < npan; ++p) {
        for (int32_t j = pl.pan_seg[p]; j < pl.pan_seg[p + 1]; ++j) local[pl.seg_ray[j]] = j;
        for (int64_t k = h.ptr[pl.pan_row[p]]; k < h.ptr[pl.pan_row[p + 1]]; ++k)
            ++sptr[local[h.ind[k]] + 1];
    }
    for (int64_t s = 0; s < nseg; ++s) sptr[s + 1] += sptr[s];
    std::vector<int32_t> pos(sptr.begin(), sptr.end() - 1);
    std::vector<uint16_t> lidx(h.nnz > 0 ? h.nnz : 1);
    std::vector<uint8_t> wb((h.nnz > 0 ? h.nnz : 1) * h.vb);
    for (int64_t p = 0; p < npan; ++p) {
        for (int32_t j = pl.pan_seg[p]; j < pl.pan_seg[p + 1]; ++j) local[pl.seg_ray[j]] = j;
        const int32_t v0 = pl.pan_row[p];
        for (int32_t v = v0; v < pl.pan_row[p + 1]; ++v)
            for (int32_t k = h.ptr[v]; k < h.ptr[v + 1]; ++k) {
                const int32_t q = pos[local[h.ind[k]]]++;
                lidx[q] = (uint16_t)(v - v0);
                std::memcpy(&wb[(size_t)q * h.vb], &h.val[(size_t)k * h.vb], h.vb);
            }
    }
    // panels of the segment matrix: segment ranges pan_seg, voxel ranges pan_row
    pack_panels(ctx, pl.pan_seg, pl.pan_row, nullptr, sptr, lidx, wb.data(), h.vb, h.vtype,
                blk_nnz, pack12 ? kPanelGroupA12 : kPanelGroupA, chunk_blocks, P, pack12);
    if (tmap) {
        // h's rows in super-tile order: the panels stage x through the tile map
        P.src.alloc(tmap->size());
        TV3_HIP(hipMemcpyAsync(P.src.p, tmap->data(), sizeof(int32_t) * tmap->size(),
                               hipMemcpyHostToDevice, ctx->stream));
    }
    // segments of each ray, in increasing segment number
    std::vector<int32_t> rptr(m + 1, 0), rseg(nseg > 0 ? nseg : 1);
    for (int64_t s = 0; s < nseg; ++s) ++rptr[pl.seg_ray[s] + 1];
    for (int64_t r = 0; r < m; ++r) rptr[r + 1] += rptr[r];
    std::vector<int32_t> rpos(rptr.begin(), rptr.end() - 1);
    for (int64_t s = 0; s < nseg; ++s) rseg[rpos[pl.seg_ray[s]]++] = (int32_t)s;
    hipStream_t st = ctx->stream;
    R.rows = m;
    R.nseg = nseg;
    const double per_ray = m > 0 ? (double)nseg / (double)m : 0.0;
    R.lanes = per_ray > 24 ? 16 : per_ray > 12 ? 8 : 4;
    R.ptr.alloc(m + 1);
    R.seg.alloc(rseg.size() > 0 ? rseg.size() : 1);   // the sum kernel reads seg[0] of empty rays
    R.part.alloc(nseg > 0 ? nseg : 1);
    TV3_HIP(hipMemcpyAsync(R.ptr.p, rptr.data(), sizeof(int32_t) * (m + 1), hipMemcpyHostToDevice,
                           st));
    if (rseg.empty())
        TV3_HIP(hipMemsetAsync(R.seg.p, 0, sizeof(int32_t), st));   // a valid index into part
    else
        TV3_HIP(hipMemcpyAsync(R.seg.p, rseg.data(), sizeof(int32_t) * rseg.size(),
                               hipMemcpyHostToDevice, st));
    TV3_HIP(hipStreamSynchronize(st));
}

// The rows of A^T (internal voxels) in super-tile order: blocks of 2 x 2 x 2
// tiles of the B = 8 numbering, the tiles of a block in (y, x, z) bit order,
// so that 2, 4 and 8 consecutive tiles form boxes of 16 x 8 x 8, 16 x 16 x 8
// and 16^3 voxels.  A ray crossing the volume leaves 14.9 segments in boxes of
// 16 x 8 x 8 voxels, 12.0 in 16 x 16 x 8 and 9.1 in 16^3 at 128^3 (spherical
// geometry, 256 x 1024 rays): fewer partial sums for the second pass.
// `vols` lists the volumes in the rows (one for a single operator, one per
// sparse problem of a batch): their first row (a multiple of 4096) and tile
// counts (even along every axis); rows outside them keep their place.
// tmap[t'] = internal tile of super-tile-order tile t'.
static bool supertile_rows(const std::vector<TileVolume> &vols, const HostCSR &h, HostCSR &hs,
                           std::vector<int32_t> &tmap) {
    if (h.rows % 512 != 0 || vols.empty()) return false;
    const int64_t ntile = h.rows / 512;
    tmap.resize((size_t)ntile);
    std::iota(tmap.begin(), tmap.end(), 0);
    for (const TileVolume &v : vols) {
        if (v.row0 % 4096 != 0 || v.nby % 2 || v.nbx % 2 || v.nbz % 2) return false;
        const int64_t t0 = v.row0 / 512;
        if (t0 + (int64_t)v.nby * v.nbx * v.nbz > ntile) return false;
        for (int tz = 0; tz < v.nbz; ++tz)
            for (int tx = 0; tx < v.nbx; ++tx)
                for (int ty = 0; ty < v.nby; ++ty) {
                    const int64_t t = ((int64_t)tz * v.nbx + tx) * v.nby + ty;
                    const int64_t S =
                        ((int64_t)(tz >> 1) * (v.nbx / 2) + (tx >> 1)) * (v.nby / 2) + (ty >> 1);
                    const int w = ((tz & 1) << 2) | ((tx & 1) << 1) | (ty & 1);
                    tmap[(size_t)(t0 + S * 8 + w)] = (int32_t)(t0 + t);
                }
    }
    hs.rows = h.rows; hs.cols = h.cols; hs.nnz = h.nnz; hs.vb = h.vb; hs.vtype = h.vtype;
    hs.ptr.assign((size_t)h.rows + 1, 0);
    hs.ind.resize(h.ind.size());
    hs.val.resize(h.val.size());
    int64_t at = 0;
    for (int64_t ts = 0; ts < ntile; ++ts) {
        const int64_t r0 = (int64_t)tmap[(size_t)ts] * 512;
        const int64_t k0 = h.ptr[r0], k1 = h.ptr[r0 + 512];
        for (int i = 0; i < 512; ++i) hs.ptr[ts * 512 + i] = (int32_t)(at + h.ptr[r0 + i] - k0);
        std::copy(h.ind.begin() + k0, h.ind.begin() + k1, hs.ind.begin() + at);
        std::copy(h.val.begin() + k0 * (int64_t)h.vb, h.val.begin() + k1 * (int64_t)h.vb,
                  hs.val.begin() + at * (int64_t)h.vb);
        at += k1 - k0;
    }
    hs.ptr[(size_t)h.rows] = (int32_t)at;
    return true;
}

// the fused layout's volume as one TileVolume (B = 8 only)
static bool layout_volume(const FusedLayout &F, std::vector<TileVolume> &vols) {
    vols.clear();
    if (F.B != 8) return false;
    vols.push_back({0, F.nby, F.nbx, F.dims[2] / 8});
    return true;
}

// Reads a buffer larger than the MALL (the 256 MiB memory-side cache) so
// that the next kernel starts with cold caches, as a product does inside the
// solver, where the other kernels of an iteration stream ~0.5 GB in between.
__global__ void k_scrub(const float4 *__restrict__ p, int64_t n, float *sink) {
    float a = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float4 v = p[i];
        a += v.x + v.y + v.z + v.w;
    }
    if (a == -1.f) *sink = a;   // the buffer is zero: never taken, keeps the loads
}

// Median duration (ms) of `reps` launches of f on the context stream, each
// timed alone after k_scrub has read 512 MiB, which evicts the operator from
// the 256 MiB memory-side cache as the solver's other kernels do between two
// products (back-to-back launches would time a cache-resident operator and
// can rank the shapes differently); one untimed launch first.  A 512 MiB
// write instead leaves the caches full of dirty lines whose write-back
// lands in the timed launch: at 128^3 it ranked A^T's packed row blocks of
// 2048 entries first, whose product takes 5% longer in the solver than
// with the blocks of 4096 the read ranks first.  15 launches: candidates
// 2% apart (A's packed and 4-byte entries) are ranked as the solver ranks
// them, which the median of 5 did not always do.
struct ColdTimer {
    static constexpr int64_t kScrub = (int64_t)(512 << 20) / 16;
    tv3_ctx *ctx;
    DevBuf<float4> scrub;
    DevBuf<float> sink;
    explicit ColdTimer(tv3_ctx *c) : ctx(c), scrub(kScrub), sink(1) {
        TV3_HIP(hipMemsetAsync(scrub.p, 0, scrub.bytes(), c->stream));
    }
    template <typename F>
    float operator()(F &&f, int reps = 15) {
        hipStream_t s = ctx->stream;
        std::vector<hipEvent_t> ev(2 * reps);
        for (auto &e : ev) TV3_HIP(hipEventCreate(&e));
        f();
        for (int r = 0; r < reps; ++r) {
            hipLaunchKernelGGL(k_scrub, dim3(4096), dim3(256), 0, s, scrub.p, kScrub, sink.p);
            TV3_HIP(hipEventRecord(ev[2 * r], s));
            f();
            TV3_HIP(hipEventRecord(ev[2 * r + 1], s));
        }
        TV3_HIP(hipEventSynchronize(ev[2 * reps - 1]));
        std::vector<float> t(reps);
        for (int r = 0; r < reps; ++r)
            TV3_HIP(hipEventElapsedTime(&t[r], ev[2 * r], ev[2 * r + 1]));
        for (auto &e : ev) (void)hipEventDestroy(e);
        std::sort(t.begin(), t.end());
        return t[reps / 2];
    }
};

// Autotuner: time the product of S over the candidate launch shapes (row
// blocks of 1024/2048/4096 entries, row groups of 8..64 lanes) and keep the
// fastest.  Runs once per operator and volume shape.  Returns its time (ms).
static float tune_shape(tv3_ctx *ctx, ColdTimer &time_launches, SparseRows &S, int64_t xlen) {
    hipStream_t s = ctx->stream;
    DevBuf<float> x(xlen > 0 ? xlen : 1), y(S.rows > 0 ? S.rows : 1);
    TV3_HIP(hipMemsetAsync(x.p, 0, x.bytes(), s));
    auto run = [&]() { launch_product<float>(S, x.p, EpiStore<float>{y.p}, s); };
    float best = 1e30f;
    int choice = 0;
    for (int nb : {1024, 2048, 4096}) {
        build_row_blocks(ctx, S, nb);
        const float t = time_launches(run);
        if (t < best) { best = t; choice = -nb; }
    }
    for (int L : {8, 16, 32, 64}) {
        S.blk_nnz = 0;
        S.lanes = L;
        const float t = time_launches(run);
        if (t < best) { best = t; choice = L; }
    }
    if (choice < 0) build_row_blocks(ctx, S, -choice);
    else { S.blk_nnz = 0; S.lanes = choice; }
    return best;
}

// Panel forms of A^T and of A, each over panels of one tile (512 internal
// voxels, at most 16384 entries) or of two consecutive tiles (16 x 8 x 8
// voxels, at most 32768 entries: A's rays leave fewer, longer segments and
// so fewer partial sums for the second pass; A^T stages the union of the two
// tiles' rays once), timed over row blocks of 1024 and 2048 entries (A^T
// also 4096, and with FP16 or BF16 weights packed entries in blocks of 2048 or
// 4096); A also over panels of 4 and 8 tiles where the tile counts are
// even; each kept when faster than the tuned plain kernel (t_A, t_T in ms).
static void tune_panels(tv3_ctx *ctx, ColdTimer &time_launches, FusedLayout &F, float t_A,
                        float t_T) {
    hipStream_t s = ctx->stream;
    const int64_t n = F.AT.rows, m = F.A.rows;
    const size_t vb = dtype_size(F.AT.vtype);
    if (vb != 2 && F.AT.vtype != TV3_F32) return;
    const HostCSR h = download_csr(ctx, F.AT);
    PanelPlan pl;
    if (!plan_panels(h, 512, 16384, pl)) return;
    F.npan = (int64_t)pl.pan_row.size() - 1;
    F.nseg = (int64_t)pl.seg_ray.size();
    F.max_seg = (int)pl.max_seg;
    DevBuf<float> xr(m > 0 ? m : 1), xv(n > 0 ? n : 1), y(n > m ? n : (m > 0 ? m : 1));
    TV3_HIP(hipMemsetAsync(xr.p, 0, xr.bytes(), s));
    TV3_HIP(hipMemsetAsync(xv.p, 0, xv.bytes(), s));
    // 16-bit weights: also with packed entries (panels of at most 4096 rays)
    const bool f16 = F.AT.vtype == TV3_F16 || F.AT.vtype == TV3_BF16;   // 16-bit weights
    for (int unit : {512, 1024})
        for (int packed = 0; packed <= (f16 ? 1 : 0); ++packed) {
            PanelPlan pl2;
            const bool own = unit != 512 || packed;
            if (own && !plan_panels(h, unit, 16384 * unit / 512, pl2,
                                    packed ? 4096 : kPanelSegCap))
                continue;
            const PanelPlan &pt = own ? pl2 : pl;
            for (int nb : {1024, 2048, 4096}) {
                if (packed && nb == 1024) continue;      // packed blocks: 2048 or 4096
                Panels P;
                build_panels_T(ctx, h, pt, nb, P, packed != 0);
                const float t =
                    time_launches([&]() { launch_panel(P, xr.p, EpiStoreIf{y.p, nullptr}, s); });
                if (t < t_T) {
                    t_T = t;
                    F.PT = std::move(P);
                    F.panels_T = true;
                    F.unit_T = unit;
                    F.npan = (int64_t)pt.pan_row.size() - 1;
                    F.nseg = (int64_t)pt.seg_ray.size();
                    F.max_seg = (int)pt.max_seg;
                }
            }
        }
    for (int unit : {512, 1024}) {
        PanelPlan pl2;
        if (unit != 512 && !plan_panels(h, unit, 16384 * unit / 512, pl2)) continue;
        const PanelPlan &pa = unit == 512 ? pl : pl2;
        for (int nb : {1024, 2048}) {
            Panels P;
            SegSum R;
            build_panels_A(ctx, h, pa, nb, P, R);
            const float t = time_launches([&]() {
                launch_panel(P, xv.p, EpiStoreIf{R.part.p, nullptr}, s);
                launch_seg_sum(R, EpiStoreIf{y.p, nullptr}, s);
            });
            if (t < t_A) {
                t_A = t;
                F.PA = std::move(P);
                F.RS = std::move(R);
                F.panels_A = true;
                F.unit_A = unit;
            }
        }
    }
    // A over panels of 4 and 8 tiles (16 x 16 x 8 and 16^3 voxels), row
    // blocks of 2048 or 4096 entries.  Measured at 128^3 (A in us, work items
    // of 6 blocks): two tiles 67.9 (blocks of 2048); four tiles 61.9 (2048),
    // 63.7 (4096); eight tiles 64.3 (2048), 60.4 (4096).  Items of 3, 12 or
    // 24 blocks were slower for every panel size.
    HostCSR hs;
    std::vector<int32_t> tmap;
    std::vector<TileVolume> vols;
    if (!layout_volume(F, vols) || !supertile_rows(vols, h, hs, tmap)) return;
    // 16-bit weights: also with packed entries (12-bit indices, 28 bytes per 8
    // entries instead of 32)
    for (int unit : {2048, 4096}) {
        PanelPlan pl2;
        if (!plan_panels(hs, unit, 16384 * unit / 512, pl2)) continue;
        for (int packed = 0; packed <= (f16 ? 1 : 0); ++packed)
            for (int nb : {2048, 4096}) {
                Panels P;
                SegSum R;
                build_panels_A(ctx, hs, pl2, nb, P, R, packed ? kPanelChunkA12 : kPanelChunkA,
                               &tmap, packed != 0);
                const float t = time_launches([&]() {
                    launch_panel(P, xv.p, EpiStoreIf{R.part.p, nullptr}, s);
                    launch_seg_sum(R, EpiStoreIf{y.p, nullptr}, s);
                });
                if (t < t_A) {
                    t_A = t;
                    F.PA = std::move(P);
                    F.RS = std::move(R);
                    F.panels_A = true;
                    F.unit_A = unit;
                }
            }
    }
}

bool build_transposed_panels(tv3_ctx *ctx, const SparseRows &AT, const std::vector<int> &blks,
                             std::vector<Panels> &out) {
    if (dtype_size(AT.vtype) != 2 && AT.vtype != TV3_F32) return false;
    const HostCSR h = download_csr(ctx, AT);
    PanelPlan pl, pl12;
    if (!plan_panels(h, 512, 16384, pl)) return false;
    // 16-bit weights in row blocks of 2048 or 4096: packed entries, over a
    // plan whose panels hold at most 4096 rays
    const bool half_w = AT.vtype == TV3_F16 || AT.vtype == TV3_BF16;
    const bool have12 = half_w && plan_panels(h, 512, 16384, pl12, 4096);
    out.clear();
    out.resize(blks.size());
    for (size_t i = 0; i < blks.size(); ++i) {
        const bool pk = have12 && blks[i] >= 2048;
        build_panels_T(ctx, h, pk ? pl12 : pl, blks[i], out[i], pk);
    }
    return true;
}

bool build_forward_panels(tv3_ctx *ctx, const SparseRows &AT, const std::vector<int> &blks,
                          const std::vector<int64_t> &row_starts, std::vector<Panels> &P,
                          std::vector<SegSum> &R, std::vector<int32_t> &seg_starts,
                          const std::vector<TileVolume> *boxes) {
    if (dtype_size(AT.vtype) != 2 && AT.vtype != TV3_F32) return false;
    const HostCSR h0 = download_csr(ctx, AT);
    HostCSR hs;
    std::vector<int32_t> tmap;
    if (boxes && !supertile_rows(*boxes, h0, hs, tmap)) return false;
    const HostCSR &h = boxes ? hs : h0;
    const int unit = boxes ? 4096 : 512;
    PanelPlan pl;
    if (!plan_panels(h, unit, 16384 * unit / 512, pl)) return false;
    // panels never straddle a unit (512 rows, or 4096 for 16^3 boxes, whose
    // volumes start at multiples of 4096 rows), so each range's segments are
    // those of the panels from its first row on
    seg_starts.assign(row_starts.size() + 1, (int32_t)pl.seg_ray.size());
    size_t p = 0;
    for (size_t k = 0; k < row_starts.size(); ++k) {
        while (p + 1 < pl.pan_row.size() && pl.pan_row[p] < row_starts[k]) ++p;
        seg_starts[k] = pl.pan_seg[p];
    }
    P.clear(); R.clear();
    P.resize(blks.size()); R.resize(blks.size());
    // 16^3 boxes with 16-bit weights in row blocks of 2048 or 4096: packed
    // entries, rows in pairs, work items sized to one round
    const bool half_w = AT.vtype == TV3_F16 || AT.vtype == TV3_BF16;
    for (size_t i = 0; i < blks.size(); ++i) {
        const bool pk = boxes && half_w && blks[i] >= 2048;
        build_panels_A(ctx, h, pl, blks[i], P[i], R[i], pk ? kPanelChunkA12 : kPanelChunkA,
                       boxes ? &tmap : nullptr, pk);
    }
    return true;
}

void renumber_operator(tv3_op *op, const int32_t dims[3], FusedLayout *L) {
    TV3_REQUIRE(op->kind == OpKind::CSR, TV3_E_UNSUPPORTED,
                "the fused solver needs a sparse operator or a matrix-free Siddon operator with "
                "every dimension a multiple of 4");
    renumber_rows(op->ctx, op->A, op->ray_order, dims, L);
}

void renumber_rows(tv3_ctx *ctx, const SparseRows &A, const std::vector<int32_t> &ray_order,
                   const int32_t dims[3], FusedLayout *L) {
    hipStream_t s = ctx->stream;
    for (int a = 0; a < 3; ++a) L->dims[a] = dims[a];
    const int p = dims[0], q = dims[1], r = dims[2];
    const int64_t n = (int64_t)p * q * r, m = A.rows;
    L->B = (p % 8 == 0 && q % 8 == 0 && r % 8 == 0) ? 8
           : (p % 4 == 0 && q % 4 == 0 && r % 4 == 0) ? 4 : 1;
    L->nby = L->B == 8 ? p / 8 : p / 4;
    L->nbx = L->B == 8 ? q / 8 : q / 4;
    L->v_nat2int.alloc(n);
    L->v_int2nat.alloc(n);
    hipLaunchKernelGGL(k_vox_maps, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, q, r, L->B,
                       L->nby, L->nbx, L->v_nat2int.p, L->v_int2nat.p);
    std::vector<int32_t> order = ray_order;
    if ((int64_t)order.size() != m) {
        order.resize(m);
        std::iota(order.begin(), order.end(), 0);
    }
    L->r_int2nat.alloc(m);
    TV3_HIP(hipMemcpyAsync(L->r_int2nat.p, order.data(), sizeof(int32_t) * m,
                           hipMemcpyHostToDevice, s));
    SparseRows &N = L->A;
    N.rows = m; N.cols = n; N.nnz = A.nnz; N.vtype = A.vtype;
    N.ptr.alloc(m + 1);
    DevBuf<int32_t> len(m > 0 ? m : 1);
    const unsigned gm = (unsigned)((m + 255) / 256);
    hipLaunchKernelGGL(k_perm_rowlen, dim3(gm), dim3(256), 0, s, m, L->r_int2nat.p, A.ptr.p, len.p);
    exclusive_scan_i32(ctx, len.p, N.ptr.p, m);
    N.ind.alloc(A.nnz > 0 ? A.nnz : 1);
    const size_t vb = dtype_size(A.vtype);
    N.val.alloc((A.nnz > 0 ? A.nnz : 1) * vb);
    if (vb == 2)
        hipLaunchKernelGGL(k_perm_rows<2>, dim3(gm), dim3(256), 0, s, m, L->r_int2nat.p, A.ptr.p,
                           A.ind.p, A.val.p, N.ptr.p, L->v_nat2int.p, N.ind.p, N.val.p);
    else if (vb == 4)
        hipLaunchKernelGGL(k_perm_rows<4>, dim3(gm), dim3(256), 0, s, m, L->r_int2nat.p, A.ptr.p,
                           A.ind.p, A.val.p, N.ptr.p, L->v_nat2int.p, N.ind.p, N.val.p);
    else
        throw tv3_error(TV3_E_FORMAT, "the fused solver needs F32, F16 or BF16 weights");
    TV3_HIP(hipGetLastError());
    build_transpose(ctx, N, L->AT);
}

FusedLayout &fused_layout(tv3_op *op, const int32_t dims[3], bool tune) {
    if (op->fused && op->fused->dims[0] == dims[0] && op->fused->dims[1] == dims[1] &&
        op->fused->dims[2] == dims[2])
        return *op->fused;
    TV3_REQUIRE(op->kind != OpKind::DYNAMIC || !op->fused, TV3_E_SHAPE,
                "a matrix-free operator's layout is the volume of its geometry: dims must be "
                "(rv_y, rv_x, rv_z)");
    tv3_ctx *ctx = op->ctx;
    hipStream_t s = ctx->stream;
    std::unique_ptr<FusedLayout> L(new FusedLayout());
    renumber_operator(op, dims, L.get());
    SparseRows &N = L->A;
    const int64_t n = N.cols, m = N.rows;
    if (tune) {
        ColdTimer timer(ctx);
        const float t_a = tune_shape(ctx, timer, N, n);
        const float t_at = tune_shape(ctx, timer, L->AT, m);
        tune_panels(ctx, timer, *L, t_a, t_at);
    }
    TV3_HIP(hipStreamSynchronize(s));
    op->fused = std::move(L);
    return *op->fused;
}

void layout_matvec(tv3_op *op, const int32_t dims[3], bool trans, const float *x, float *y,
                   hipStream_t s) {
    FusedLayout &F = fused_layout(op, dims);
    const int64_t n = F.AT.rows, m = F.A.rows;
    DevBuf<float> xi(trans ? (m > 0 ? m : 1) : (n > 0 ? n : 1));
    DevBuf<float> yi(trans ? (n > 0 ? n : 1) : (m > 0 ? m : 1));
    if (trans) {
        permute_gather(x, F.r_int2nat.p, xi.p, m, s);
        launch_transposed<float>(F, xi.p, EpiStore<float>{yi.p}, s);
        permute_gather(yi.p, F.v_nat2int.p, y, n, s);
    } else {
        permute_gather(x, F.v_int2nat.p, xi.p, n, s);
        launch_forward<float>(F, xi.p, EpiStore<float>{yi.p}, s);
        permute_scatter(yi.p, F.r_int2nat.p, y, m, s);
    }
    TV3_HIP(hipGetLastError());
    TV3_HIP(hipStreamSynchronize(s));   // xi, yi are freed on return
}

static void apply_code(tv3_ctx *ctx, SparseRows &S, int code) {
    if (code > 0) { S.lanes = code; S.blk_nnz = 0; }
    else if (code < 0) build_row_blocks(ctx, S, -code);
}

void layout_force(tv3_op *op, const int32_t dims[3], int code_N, int code_T, int pan_nnz,
                  int panels) {
    // a layout not built yet is built without timing: the shapes given here
    // replace whatever the timing would pick
    FusedLayout &F = fused_layout(op, dims, false);
    TV3_REQUIRE(!F.dyn, TV3_E_UNSUPPORTED, "a matrix-free operator's tile products have no "
                "launch shapes to force");
    F.session.reset();               // its graphs hold the old launch shapes
    tv3_ctx *ctx = op->ctx;
    const bool pa = (panels & 1) != 0, pt = (panels & 2) != 0, packed_T = (panels & 512) != 0;
    const int unit_A = (panels & 32) ? 4096 : (panels & 16) ? 2048 : (panels & 4) ? 1024 : 512;
    const int unit_T = (panels & 8) ? 1024 : 512;
    const bool half_w = F.AT.vtype == TV3_F16 || F.AT.vtype == TV3_BF16;
    TV3_REQUIRE(!packed_T || (pt && half_w && code_T != -1024), TV3_E_UNSUPPORTED,
                "packed panel entries take FP16 or BF16 weights and blocks of 2048 or 4096");
    if (pa || pt) {
        TV3_REQUIRE(pan_nnz > 0, TV3_E_ARG, "panels need pan_nnz > 0");
        for (int c : {pa ? code_N : -1024, pt ? code_T : -1024})
            TV3_REQUIRE(c == -1024 || c == -2048 || c == -4096, TV3_E_ARG,
                        "panel block sizes are -1024, -2048 or -4096");
        const HostCSR h = download_csr(ctx, F.AT);
        PanelPlan pl;
        TV3_REQUIRE(plan_panels(h, 512, pan_nnz, pl), TV3_E_UNSUPPORTED,
                    "a panel of this operator would need more rays than LDS holds");
        F.npan = (int64_t)pl.pan_row.size() - 1;
        F.nseg = (int64_t)pl.seg_ray.size();
        F.max_seg = (int)pl.max_seg;
        if (pt && unit_T == 512 && !packed_T) build_panels_T(ctx, h, pl, -code_T, F.PT);
        if (pt && (unit_T != 512 || packed_T)) {
            PanelPlan pl2;
            TV3_REQUIRE(plan_panels(h, unit_T, pan_nnz * (unit_T / 512), pl2,
                                    packed_T ? 4096 : kPanelSegCap),
                        TV3_E_UNSUPPORTED,
                        "a panel of this operator would need more rays than LDS holds");
            build_panels_T(ctx, h, pl2, -code_T, F.PT, packed_T);
            F.npan = (int64_t)pl2.pan_row.size() - 1;
            F.nseg = (int64_t)pl2.seg_ray.size();
            F.max_seg = (int)pl2.max_seg;
        }
        F.unit_T = unit_T;
        if (pa && unit_A == 512) build_panels_A(ctx, h, pl, -code_N, F.PA, F.RS);
        if (pa && unit_A > 1024) {
            HostCSR hs;
            std::vector<int32_t> tmap;
            std::vector<TileVolume> vols;
            TV3_REQUIRE(layout_volume(F, vols) && supertile_rows(vols, h, hs, tmap),
                        TV3_E_UNSUPPORTED,
                        "panels of 4 or 8 tiles need 8^3 tiles and an even tile count along "
                        "every axis");
            PanelPlan pl2;
            TV3_REQUIRE(plan_panels(hs, unit_A, pan_nnz * (unit_A / 512), pl2), TV3_E_UNSUPPORTED,
                        "a panel of this operator would need more rays than LDS holds");
            const bool packed = (panels & 256) != 0;
            TV3_REQUIRE(!packed || (half_w && code_N != -1024), TV3_E_UNSUPPORTED,
                        "packed panel entries take FP16 or BF16 weights and blocks of 2048 or "
                        "4096");
            build_panels_A(ctx, hs, pl2, -code_N, F.PA, F.RS,
                           packed ? kPanelChunkA12 : kPanelChunkA, &tmap, packed);
        }
        if (pa && unit_A == 1024) {
            PanelPlan pl2;
            TV3_REQUIRE(plan_panels(h, unit_A, 2 * pan_nnz, pl2), TV3_E_UNSUPPORTED,
                        "a panel of this operator would need more rays than LDS holds");
            build_panels_A(ctx, h, pl2, -code_N, F.PA, F.RS);
        }
        F.unit_A = unit_A;
    }
    F.panels_A = pa;
    F.panels_T = pt;
    if (!pa) apply_code(ctx, F.A, code_N);
    if (!pt) apply_code(ctx, F.AT, code_T);
}

__global__ void k_fill(float *p, float v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Bandwidth probes: a streaming copy (read n, write n float4) and a streaming
// read, four independent 16-byte accesses per thread per pass.
__global__ __launch_bounds__(256) void k_probe_copy(const float4 *__restrict__ src,
                                                    float4 *__restrict__ dst, int64_t n) {
    const int64_t G = (int64_t)gridDim.x * 256;
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + 3 * G < n; i += 4 * G) {
        const float4 a = src[i], b = src[i + G], c = src[i + 2 * G], d = src[i + 3 * G];
        dst[i] = a; dst[i + G] = b; dst[i + 2 * G] = c; dst[i + 3 * G] = d;
    }
    for (; i < n; i += G) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_probe_read(const float4 *__restrict__ src, int64_t n,
                                                    float *sink) {
    const int64_t G = (int64_t)gridDim.x * 256;
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float s = 0.f;
    for (; i + 3 * G < n; i += 4 * G) {
        const float4 a = src[i], b = src[i + G], c = src[i + 2 * G], d = src[i + 3 * G];
        s += (a.x + b.x) + (c.x + d.x);
    }
    for (; i < n; i += G) s += src[i].x;
    if (s == -1.f) *sink = s;   // the buffer is zero: never taken, keeps the loads
}

void bandwidth_probe(tv3_ctx *ctx, int64_t bytes, int reps, double *copy_gbps,
                     double *read_gbps) {
    hipStream_t s = ctx->stream;
    const int64_t n = bytes / 16;
    TV3_REQUIRE(n >= 1024, TV3_E_ARG, "probe size below 16 KiB");
    DevBuf<float4> a(n), b(n);
    DevBuf<float> sink(1);
    TV3_HIP(hipMemsetAsync(a.p, 0, a.bytes(), s));
    TV3_HIP(hipMemsetAsync(b.p, 0, b.bytes(), s));
    const unsigned grid = (unsigned)std::max(1, ctx->num_cus) * 8;
    hipEvent_t e0, e1;
    TV3_HIP(hipEventCreate(&e0));
    TV3_HIP(hipEventCreate(&e1));
    auto median_ms = [&](auto launch) {
        std::vector<float> t;
        launch();
        for (int r = 0; r < std::max(1, reps); ++r) {
            TV3_HIP(hipEventRecord(e0, s));
            launch();
            TV3_HIP(hipEventRecord(e1, s));
            TV3_HIP(hipEventSynchronize(e1));
            float ms = 0;
            TV3_HIP(hipEventElapsedTime(&ms, e0, e1));
            t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        return (double)t[t.size() / 2];
    };
    const double tc = median_ms([&] {
        hipLaunchKernelGGL(k_probe_copy, dim3(grid), dim3(256), 0, s, a.p, b.p, n);
    });
    const double tr = median_ms([&] {
        hipLaunchKernelGGL(k_probe_read, dim3(grid), dim3(256), 0, s, a.p, n, sink.p);
    });
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (copy_gbps) *copy_gbps = 2.0 * 16.0 * (double)n / (tc * 1e-3) / 1e9;
    if (read_gbps) *read_gbps = 16.0 * (double)n / (tr * 1e-3) / 1e9;
}

double layout_time(tv3_op *op, const int32_t dims[3], bool trans, int reps) {
    FusedLayout &F = fused_layout(op, dims);
    tv3_ctx *ctx = op->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = F.AT.rows, m = F.A.rows;
    const int64_t xl = trans ? m : n, yl = trans ? n : m;
    DevBuf<float> x(xl > 0 ? xl : 1), y(yl > 0 ? yl : 1);
    if (xl > 0)
        hipLaunchKernelGGL(k_fill, dim3((unsigned)((xl + 255) / 256)), dim3(256), 0, s, x.p, 1.f,
                           xl);
    const int64_t nscrub = (int64_t)(512 << 20) / 16;
    DevBuf<float4> scrub(nscrub);
    DevBuf<float> sink(1);
    TV3_HIP(hipMemsetAsync(scrub.p, 0, scrub.bytes(), s));
    auto run = [&]() {
        if (trans) launch_transposed<float>(F, x.p, EpiStore<float>{y.p}, s);
        else launch_forward<float>(F, x.p, EpiStore<float>{y.p}, s);
    };
    hipEvent_t e0, e1;
    TV3_HIP(hipEventCreate(&e0));
    TV3_HIP(hipEventCreate(&e1));
    run();
    double total = 0;
    for (int r = 0; r < (reps > 0 ? reps : 1); ++r) {
        hipLaunchKernelGGL(k_scrub, dim3(4096), dim3(256), 0, s, scrub.p, nscrub, sink.p);
        TV3_HIP(hipEventRecord(e0, s));
        run();
        TV3_HIP(hipEventRecord(e1, s));
        TV3_HIP(hipEventSynchronize(e1));
        float ms = 0;
        TV3_HIP(hipEventElapsedTime(&ms, e0, e1));
        total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 1e3 * total / (reps > 0 ? reps : 1);
}

}  // namespace tv3
