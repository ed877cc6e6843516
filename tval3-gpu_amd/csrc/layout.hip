// layout.hip -- the fused solver's internal numbering of voxels and rays.
//
// Scattered gathers of x (A x) and of the residual (A^T r) are bound by the
// L2 request rate, so the solver works in a numbering that puts neighbours
// in the same cache lines: voxels in 4x4x4 bricks (64 voxels = 256 bytes,
// y fastest inside a brick, bricks ordered y, x, z) and rays ordered by
// emitter and receiver position.  Inputs and outputs are permuted at the
// boundary of a solve only.
#include <algorithm>
#include <cmath>
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
    int64_t i;
    if (B == 1) {
        i = v;
    } else {
        i = ((((int64_t)(z >> 2) * nbx + (x >> 2)) * nby + (y >> 2)) << 6) + ((z & 3) << 4) +
            ((x & 3) << 2) + (y & 3);
    }
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

// Autotuner: time the product of S over the candidate launch shapes (row
// blocks of 1024/2048/4096 entries, row groups of 8..64 lanes) and keep the
// fastest.  Runs once per operator and volume shape.
static void tune_shape(tv3_ctx *ctx, SparseRows &S, int64_t xlen) {
    hipStream_t s = ctx->stream;
    DevBuf<float> x(xlen > 0 ? xlen : 1), y(S.rows > 0 ? S.rows : 1);
    TV3_HIP(hipMemsetAsync(x.p, 0, x.bytes(), s));
    hipEvent_t e0, e1;
    TV3_HIP(hipEventCreate(&e0));
    TV3_HIP(hipEventCreate(&e1));
    auto timed = [&]() {
        launch_product<float>(S, x.p, EpiStore<float>{y.p}, s);
        TV3_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 4; ++r) launch_product<float>(S, x.p, EpiStore<float>{y.p}, s);
        TV3_HIP(hipEventRecord(e1, s));
        TV3_HIP(hipEventSynchronize(e1));
        float ms = 0;
        TV3_HIP(hipEventElapsedTime(&ms, e0, e1));
        return ms;
    };
    float best = 1e30f;
    int choice = 0;
    for (int nb : {1024, 2048, 4096}) {
        build_row_blocks(ctx, S, nb);
        const float t = timed();
        if (t < best) { best = t; choice = -nb; }
    }
    for (int L : {8, 16, 32, 64}) {
        S.blk_nnz = 0;
        S.lanes = L;
        const float t = timed();
        if (t < best) { best = t; choice = L; }
    }
    if (choice < 0) build_row_blocks(ctx, S, -choice);
    else { S.blk_nnz = 0; S.lanes = choice; }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
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
    L->B = (p % 4 == 0 && q % 4 == 0 && r % 4 == 0) ? 4 : 1;
    L->nby = p / 4;
    L->nbx = q / 4;
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
    tune_shape(ctx, L->AT, m);
    TV3_HIP(hipStreamSynchronize(s));
    op->fused = std::move(L);
    return *op->fused;
}

}  // namespace tv3
