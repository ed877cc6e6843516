// layout.hip -- the fused solver's internal numbering of voxels and rays.
//
// Scattered gathers of x (A x) and of the residual (A^T r) are bound by the
// L1/L2 request rate, so the solver works in a numbering that puts
// neighbours in the same cache lines: voxels in 8x8x8 tiles of 4x4x4 bricks
// (brick_index) and rays ordered by emitter and receiver position.  A^T is
// further split into panels whose rays are staged in LDS (build_panels).
// Inputs and outputs are permuted at the boundary of a solve only.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "spmv_stream.hpp"

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

std::vector<int32_t> geometry_ray_order(const DynGeom &g) {
    std::vector<int32_t> order;
    const int64_t P = g.num_paths;
    if (P == 0 || g.epos_h.empty() || g.rpos_h.empty() || (int64_t)g.path_h.size() != 2 * P)
        return order;
    std::vector<int32_t> er = spatial_rank(g.epos_h), rr = spatial_rank(g.rpos_h);
    order.resize(P);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        const int32_t ea = g.path_h[2 * a], eb = g.path_h[2 * b];
        const int32_t ma = ea / g.E, mb = eb / g.E;
        if (ma != mb) return ma < mb;
        if (er[ea] != er[eb]) return er[ea] < er[eb];
        return rr[g.path_h[2 * a + 1]] < rr[g.path_h[2 * b + 1]];
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

// Largest ray list of a panel: its values take at most 48 KiB of LDS, which
// leaves room for the 16 KiB product buffer of 4096-entry blocks within the
// 64 KiB a workgroup may hold.
constexpr int kPanelSegCap = 12288;

bool build_panels(tv3_ctx *ctx, const SparseRows &S, int unit, int pan_nnz, int blk_nnz,
                  PanelT &P) {
    const int64_t n = S.rows, m = S.cols, nnz = S.nnz;
    hipStream_t s = ctx->stream;
    const size_t vb = dtype_size(S.vtype);
    TV3_REQUIRE(vb == 2 || S.vtype == TV3_F32, TV3_E_FORMAT, "panels need F32, F16 or BF16 weights");
    std::vector<int32_t> ptr(n + 1), ind(nnz > 0 ? nnz : 1);
    std::vector<uint8_t> val((nnz > 0 ? nnz : 1) * vb);
    TV3_HIP(hipMemcpyAsync(ptr.data(), S.ptr.p, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost,
                           s));
    if (nnz > 0) {
        TV3_HIP(hipMemcpyAsync(ind.data(), S.ind.p, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost,
                               s));
        TV3_HIP(hipMemcpyAsync(val.data(), S.val.p, vb * nnz, hipMemcpyDeviceToHost, s));
    }
    TV3_HIP(hipStreamSynchronize(s));
    const int64_t sub = unit >= 64 ? 64 : unit;   // split granularity: one brick
    // pass 1: panel row ranges and their sorted ray lists
    std::vector<int32_t> stamp(m > 0 ? m : 1, -1), local(m > 0 ? m : 1, 0);
    std::vector<int32_t> pan_row{0}, pan_seg{0}, seg_ray, rays, fresh;
    int32_t cur = 0;
    int64_t cur_nnz = 0, max_seg = 0;
    auto close = [&](int64_t row_end) {
        std::sort(rays.begin(), rays.end());
        seg_ray.insert(seg_ray.end(), rays.begin(), rays.end());
        max_seg = std::max<int64_t>(max_seg, (int64_t)rays.size());
        pan_seg.push_back((int32_t)seg_ray.size());
        pan_row.push_back((int32_t)row_end);
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
                const bool empty = (int64_t)pan_row.back() == c0;
                if (empty || (cur_nnz + cn <= pan_nnz &&
                              (int64_t)(rays.size() + fresh.size()) <= kPanelSegCap))
                    break;
                for (int32_t r : fresh) stamp[r] = -1;   // the chunk opens a new panel
                close(c0);
            }
            if ((int64_t)(rays.size() + fresh.size()) > kPanelSegCap) return false;
            rays.insert(rays.end(), fresh.begin(), fresh.end());
            cur_nnz += cn;
        }
        close(u1);
    }
    const int64_t npan = (int64_t)pan_row.size() - 1;
    // pass 2: local ray indices of the entries, row blocks inside panels
    std::vector<uint16_t> sidx(nnz > 0 ? nnz : 1);
    std::vector<int32_t> blk{0}, pan_blk{0};
    std::vector<int64_t> pan_work(npan);
    for (int64_t pn = 0; pn < npan; ++pn) {
        for (int32_t j = pan_seg[pn]; j < pan_seg[pn + 1]; ++j) local[seg_ray[j]] = j - pan_seg[pn];
        const int64_t r0 = pan_row[pn], r1 = pan_row[pn + 1];
        for (int64_t k = ptr[r0]; k < ptr[r1]; ++k) sidx[k] = (uint16_t)local[ind[k]];
        int64_t start = r0;
        for (int64_t r = r0; r < r1; ++r)
            if (r > start && ((int64_t)ptr[r + 1] - ptr[start] > blk_nnz || r - start >= 256)) {
                blk.push_back((int32_t)r);
                start = r;
            }
        blk.push_back((int32_t)r1);
        pan_blk.push_back((int32_t)blk.size() - 1);
        // cost: entries plus the staging gathers
        pan_work[pn] = ((int64_t)ptr[r1] - ptr[r0]) + 4 * (pan_seg[pn + 1] - pan_seg[pn]);
    }
    // blk holds each panel's boundaries once: first rows of all blocks, then n
    std::vector<int32_t> bp(blk.size()), bq(blk.size());
    for (size_t b = 0; b < blk.size(); ++b) bp[b] = ptr[blk[b]];
    bq[0] = 0;
    for (size_t b = 0; b + 1 < blk.size(); ++b) bq[b + 1] = bq[b] + ((bp[b + 1] - bp[b] + 3) & ~3);
    const int64_t npad = bq.back();
    std::vector<uint32_t> pk;
    std::vector<uint16_t> ps;
    std::vector<float> pv;
    if (vb == 2) pk.assign(npad > 0 ? npad : 1, 0u);
    else { ps.assign(npad > 0 ? npad : 1, 0); pv.assign(npad > 0 ? npad : 1, 0.f); }
    for (size_t b = 0; b + 1 < blk.size(); ++b)
        for (int32_t k = bp[b]; k < bp[b + 1]; ++k) {
            const int64_t q = bq[b] + (k - bp[b]);
            if (vb == 2) {
                uint16_t wbits;
                std::memcpy(&wbits, &val[(size_t)k * 2], 2);
                pk[q] = (uint32_t)sidx[k] | ((uint32_t)wbits << 16);
            } else {
                ps[q] = sidx[k];
                std::memcpy(&pv[q], &val[(size_t)k * 4], 4);
            }
        }
    std::vector<int32_t> order(npan);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return pan_work[a] > pan_work[b]; });
    P.npan = npan;
    P.nseg = (int64_t)seg_ray.size();
    P.max_seg = (int)max_seg;
    P.blk_nnz = blk_nnz;
    auto up = [&](DevBuf<int32_t> &d, const std::vector<int32_t> &h) {
        d.alloc(h.empty() ? 1 : h.size());
        if (!h.empty())
            TV3_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(int32_t) * h.size(),
                                   hipMemcpyHostToDevice, s));
    };
    up(P.order, order);
    up(P.pan_seg, pan_seg);
    up(P.pan_blk, pan_blk);
    up(P.seg_ray, seg_ray);
    up(P.blk, blk);
    up(P.bptr, bp);
    up(P.bpad, bq);
    P.vtype = S.vtype;
    if (vb == 2) {
        P.pk.alloc(pk.size());
        TV3_HIP(hipMemcpyAsync(P.pk.p, pk.data(), sizeof(uint32_t) * pk.size(),
                               hipMemcpyHostToDevice, s));
    } else {
        P.sidx.alloc(ps.size());
        P.pval.alloc(pv.size());
        TV3_HIP(hipMemcpyAsync(P.sidx.p, ps.data(), sizeof(uint16_t) * ps.size(),
                               hipMemcpyHostToDevice, s));
        TV3_HIP(hipMemcpyAsync(P.pval.p, pv.data(), sizeof(float) * pv.size(),
                               hipMemcpyHostToDevice, s));
    }
    TV3_HIP(hipStreamSynchronize(s));
    return true;
}

// Average duration (ms) of `reps` launches of f on the context stream after
// one untimed launch.
template <typename F>
static float time_launches(tv3_ctx *ctx, F &&f, int reps = 4) {
    hipStream_t s = ctx->stream;
    hipEvent_t e0, e1;
    TV3_HIP(hipEventCreate(&e0));
    TV3_HIP(hipEventCreate(&e1));
    f();
    TV3_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) f();
    TV3_HIP(hipEventRecord(e1, s));
    TV3_HIP(hipEventSynchronize(e1));
    float ms = 0;
    TV3_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

// Autotuner: time the product of S over the candidate launch shapes (row
// blocks of 1024/2048/4096 entries, row groups of 8..64 lanes) and keep the
// fastest.  Runs once per operator and volume shape.  Returns its time (ms).
static float tune_shape(tv3_ctx *ctx, SparseRows &S, int64_t xlen) {
    hipStream_t s = ctx->stream;
    DevBuf<float> x(xlen > 0 ? xlen : 1), y(S.rows > 0 ? S.rows : 1);
    TV3_HIP(hipMemsetAsync(x.p, 0, x.bytes(), s));
    auto run = [&]() { launch_product<float>(S, x.p, EpiStore<float>{y.p}, s); };
    float best = 1e30f;
    int choice = 0;
    for (int nb : {1024, 2048, 4096}) {
        build_row_blocks(ctx, S, nb);
        const float t = time_launches(ctx, run);
        if (t < best) { best = t; choice = -nb; }
    }
    for (int L : {8, 16, 32, 64}) {
        S.blk_nnz = 0;
        S.lanes = L;
        const float t = time_launches(ctx, run);
        if (t < best) { best = t; choice = L; }
    }
    if (choice < 0) build_row_blocks(ctx, S, -choice);
    else { S.blk_nnz = 0; S.lanes = choice; }
    return best;
}

// Panels of the CSC (tiles of 512 internal voxels, panels of at most 16384
// entries) timed over row blocks of 1024/2048/4096 entries; kept when faster
// than `best` (the tuned plain kernel's time).
static bool tune_panels(tv3_ctx *ctx, FusedLayout &F, int64_t xlen, float best) {
    hipStream_t s = ctx->stream;
    DevBuf<float> x(xlen > 0 ? xlen : 1), y(F.AT.rows > 0 ? F.AT.rows : 1);
    TV3_HIP(hipMemsetAsync(x.p, 0, x.bytes(), s));
    PanelT P;
    int choice = 0;
    for (int nb : {1024, 2048, 4096}) {
        if (!build_panels(ctx, F.AT, 512, 16384, nb, P)) return false;
        const float t = time_launches(
            ctx, [&]() { launch_panel<float>(P, F.AT, x.p, EpiStore<float>{y.p}, s); });
        if (t < best) { best = t; choice = nb; }
    }
    if (!choice) return false;
    if (choice != 4096) build_panels(ctx, F.AT, 512, 16384, choice, P);
    F.PT = std::move(P);
    return true;
}

FusedLayout &fused_layout(tv3_op *op, const int32_t dims[3]) {
    TV3_REQUIRE(op->kind == OpKind::CSR, TV3_E_UNSUPPORTED,
                "the fused solver needs a sparse operator (materialize a dynamic operator first)");
    if (op->fused && op->fused->dims[0] == dims[0] && op->fused->dims[1] == dims[1] &&
        op->fused->dims[2] == dims[2])
        return *op->fused;
    tv3_ctx *ctx = op->ctx;
    hipStream_t s = ctx->stream;
    std::unique_ptr<FusedLayout> L(new FusedLayout());
    for (int a = 0; a < 3; ++a) L->dims[a] = dims[a];
    const int p = dims[0], q = dims[1], r = dims[2];
    const int64_t n = (int64_t)p * q * r, m = op->m;
    L->B = (p % 8 == 0 && q % 8 == 0 && r % 8 == 0) ? 8
           : (p % 4 == 0 && q % 4 == 0 && r % 4 == 0) ? 4 : 1;
    L->nby = L->B == 8 ? p / 8 : p / 4;
    L->nbx = L->B == 8 ? q / 8 : q / 4;
    L->v_nat2int.alloc(n);
    L->v_int2nat.alloc(n);
    hipLaunchKernelGGL(k_vox_maps, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, q, r, L->B,
                       L->nby, L->nbx, L->v_nat2int.p, L->v_int2nat.p);
    std::vector<int32_t> order = op->ray_order;
    if ((int64_t)order.size() != m) {
        order.resize(m);
        std::iota(order.begin(), order.end(), 0);
    }
    L->r_int2nat.alloc(m);
    TV3_HIP(hipMemcpyAsync(L->r_int2nat.p, order.data(), sizeof(int32_t) * m,
                           hipMemcpyHostToDevice, s));
    const SparseRows &A = op->A;
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
    tune_shape(ctx, N, n);
    const float t_at = tune_shape(ctx, L->AT, m);
    L->use_panels = tune_panels(ctx, *L, m, t_at);
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
        launch_product<float>(F.A, xi.p, EpiStore<float>{yi.p}, s);
        permute_scatter(yi.p, F.r_int2nat.p, y, m, s);
    }
    TV3_HIP(hipGetLastError());
    TV3_HIP(hipStreamSynchronize(s));   // xi, yi are freed on return
}

static void apply_code(tv3_ctx *ctx, SparseRows &S, int code) {
    if (code > 0) { S.lanes = code; S.blk_nnz = 0; }
    else if (code < 0) build_row_blocks(ctx, S, -code);
}

void layout_force(tv3_op *op, const int32_t dims[3], int code_N, int code_T, int pan_nnz) {
    FusedLayout &F = fused_layout(op, dims);
    apply_code(op->ctx, F.A, code_N);
    if (pan_nnz > 0) {
        TV3_REQUIRE(code_T == -1024 || code_T == -2048 || code_T == -4096, TV3_E_ARG,
                    "panels need a block size code of -1024, -2048 or -4096");
        PanelT P;
        TV3_REQUIRE(build_panels(op->ctx, F.AT, 512, pan_nnz, -code_T, P), TV3_E_UNSUPPORTED,
                    "a panel of this operator would need more rays than LDS holds");
        F.PT = std::move(P);
        F.use_panels = true;
    } else {
        F.use_panels = false;
        apply_code(op->ctx, F.AT, code_T);
    }
}

__global__ void k_fill(float *p, float v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
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
    const float ms = time_launches(
        ctx,
        [&]() {
            if (trans) launch_transposed<float>(F, x.p, EpiStore<float>{y.p}, s);
            else launch_product<float>(F.A, x.p, EpiStore<float>{y.p}, s);
        },
        reps > 0 ? reps : 1);
    return 1e3 * (double)ms;
}

}  // namespace tv3
