// dynamic.hip -- matrix-free ray operators: ray-path weights are generated
// inside every product (geom_mm_multGPU / dyn_calc_kernel of the reference,
// 3d/mat_vec_mul_dynamic.h:682-967, 3d/tval3_gpu_3d_kernels.cu:1612-1764).
//
// Two weight modes:
//  * BRESENHAM -- the reference's integer 3-D walk from the emitter voxel to
//    the receiver voxel, constant weight scale_factor per visited voxel
//    (3d/tval3_gpu_3d_kernels.cu:1656-1759); coordinates are clamped into the
//    grid for both N and T (the reference clamps N through its texture,
//    :2167-2170, and does not check T).
//  * SIDDON -- exact intersection lengths of the emitter->receiver segment
//    with the voxel grid (Siddon, Med. Phys. 12(2), 1985), traversed in
//    increasing ray parameter by tracking the next plane of every axis
//    (siddon.hpp).
//
// FP32 products of a Siddon operator whose dimensions are multiples of 4
// run as tile products (dyntile.hip, panels.hpp: deterministic, no
// atomics).  The others -- Bresenham, FP64, other shapes -- run here: one
// thread walks one ray; N gathers x along the ray (x stays resident in L2 /
// MALL) and writes y[ray] once; T scatters x[ray]*w into y with FP32
// hardware atomics (no replicated sub-volumes).  The same walk, with FP
// contraction disabled, fills the CSR in dynamic_materialize so a
// precomputed matrix holds exactly the weights the dynamic product uses.
#include "siddon.hpp"
#include "tv3_internal.hpp"

#include <cmath>

namespace tv3 {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Bresenham walk; emit(linear voxel index) for each visited voxel.
template <typename F>
__device__ __forceinline__ void bresenham_walk(int x0, int y0, int z0, int x1, int y1, int z1,
                                               int rvx, int rvy, int rvz, F emit) {
    int c[3] = {x0, y0, z0};
    auto visit = [&]() {
        int xx = clampi(c[0], 0, rvx - 1), yy = clampi(c[1], 0, rvy - 1),
            zz = clampi(c[2], 0, rvz - 1);
        emit((int64_t)zz * rvx * rvy + (int64_t)xx * rvy + yy);
    };
    visit();
    int d1 = x1 - x0, d2 = y1 - y0, d3 = z1 - z0;
    int m1 = abs(d1), m2 = abs(d2), m3 = abs(d3);
    int a1, a2, a3;
    if (m1 >= m2 && m1 >= m3) {
        a1 = 0; a2 = 1; a3 = 2;
    } else if (m2 >= m1 && m2 >= m3) {
        int t = d1; d1 = d2; d2 = t;
        a1 = 1; a2 = 0; a3 = 2;
    } else {
        int t = d1; d1 = d3; d3 = d2; d2 = t;
        a1 = 2; a2 = 0; a3 = 1;
    }
    int inc1 = d1 < 0 ? -1 : 1, inc2 = d2 < 0 ? -1 : 1, inc3 = d3 < 0 ? -1 : 1;
    m1 = abs(d1); m2 = abs(d2); m3 = abs(d3);
    int e1 = 2 * m2 - m1, e2 = 2 * m3 - m1;
    for (int j = 1; j < m1 + 1; ++j) {
        if (e1 > 0) { c[a2] += inc2; e1 -= 2 * m1; }
        if (e2 > 0) { c[a3] += inc3; e2 -= 2 * m1; }
        e1 += 2 * m2; e2 += 2 * m3;
        c[a1] += inc1;
        visit();
    }
}

// Siddon traversal of p0->p1 through [org, org + N*h] (siddon.hpp);
// emit(linear voxel index, length) in traversal order.
template <typename R, typename F>
__device__ __forceinline__ void siddon_walk(const R p0[3], const R p1[3], const R org[3],
                                            const R h[3], const int N[3], int rvy_stride, F emit) {
    SidRay<R> s;
    sid_setup(s, p0, p1, org, h, N);
    if (!s.hit) return;
    sid_walk(s, s.amin, s.amax, [&](int cx, int cy, int cz, R w) {
        emit((int64_t)cz * N[0] * rvy_stride + (int64_t)cx * rvy_stride + cy, w);
    });
}

struct GeoView {
    int32_t rvx, rvy, rvz;
    float scale;
    const int32_t *path, *ec, *rc;
    const float *ep, *rp;
    float org[3], h[3];
    int64_t P;
};

static GeoView view(const DynGeom &g) {
    GeoView v;
    v.rvx = g.rv_x; v.rvy = g.rv_y; v.rvz = g.rv_z;
    v.scale = g.scale;
    v.path = g.path.p; v.ec = g.ecoord.p; v.rc = g.rcoord.p;
    v.ep = g.epos.p; v.rp = g.rpos.p;
    for (int a = 0; a < 3; ++a) { v.org[a] = g.origin[a]; v.h[a] = g.voxel[a]; }
    v.P = g.num_paths;
    return v;
}

template <typename R>
__device__ __forceinline__ void ray_ends(const GeoView &g, int64_t k, R p0[3], R p1[3]) {
    int e = g.path[2 * k], r = g.path[2 * k + 1];
#pragma unroll
    for (int a = 0; a < 3; ++a) { p0[a] = (R)g.ep[3 * e + a]; p1[a] = (R)g.rp[3 * r + a]; }
}

// The Siddon operator is the FP32 traversal: FP64 products accumulate its
// weights in FP64 (the same matrix in both precisions).
template <typename T, typename AccT>
__global__ __launch_bounds__(256) void k_dyn_N(GeoView g, int siddon, const T *__restrict__ x,
                                               T *__restrict__ y) {
    int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= g.P) return;
    AccT acc = 0;
    if (siddon) {
        float p0[3], p1[3];
        int N[3] = {g.rvx, g.rvy, g.rvz};
        ray_ends<float>(g, k, p0, p1);
        siddon_walk<float>(p0, p1, g.org, g.h, N, g.rvy,
                           [&](int64_t v, float w) { acc += (AccT)w * (AccT)x[v]; });
    } else {
        int e = g.path[2 * k], r = g.path[2 * k + 1];
        T w = (T)g.scale;
        bresenham_walk(g.ec[3 * e], g.ec[3 * e + 1], g.ec[3 * e + 2], g.rc[3 * r], g.rc[3 * r + 1],
                       g.rc[3 * r + 2], g.rvx, g.rvy, g.rvz,
                       [&](int64_t v) { acc += (AccT)x[v] * (AccT)w; });
    }
    y[k] = (T)acc;
}

template <typename T>
__global__ __launch_bounds__(256) void k_dyn_T(GeoView g, int siddon, const T *__restrict__ x,
                                               T *__restrict__ y) {
    int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= g.P) return;
    const T xr = x[k];
    if (xr == 0) return;
    if (siddon) {
        float p0[3], p1[3];
        int N[3] = {g.rvx, g.rvy, g.rvz};
        ray_ends<float>(g, k, p0, p1);
        siddon_walk<float>(p0, p1, g.org, g.h, N, g.rvy,
                           [&](int64_t v, float w) { atomicAdd(&y[v], (T)w * xr); });
    } else {
        int e = g.path[2 * k], r = g.path[2 * k + 1];
        const T v0 = xr * (T)g.scale;
        bresenham_walk(g.ec[3 * e], g.ec[3 * e + 1], g.ec[3 * e + 2], g.rc[3 * r], g.rc[3 * r + 1],
                       g.rc[3 * r + 2], g.rvx, g.rvy, g.rvz, [&](int64_t v) { atomicAdd(&y[v], v0); });
    }
}

template <typename T, typename AccT>
static void dyn_apply(tv3_op *op, bool trans, const T *x, T *y, hipStream_t s) {
    GeoView g = view(op->geo);
    int sid = op->geo.mode == TV3_WEIGHT_SIDDON;
    unsigned blocks = (unsigned)((g.P + 255) / 256);
    if (!trans) {
        if (blocks) hipLaunchKernelGGL((k_dyn_N<T, AccT>), dim3(blocks), dim3(256), 0, s, g, sid, x, y);
    } else {
        TV3_HIP(hipMemsetAsync(y, 0, sizeof(T) * op->n, s));
        if (blocks) hipLaunchKernelGGL((k_dyn_T<T>), dim3(blocks), dim3(256), 0, s, g, sid, x, y);
    }
    TV3_HIP(hipGetLastError());
}

void dynamic_apply(tv3_op *op, bool trans, const float *x, float *y, hipStream_t s, bool acc64) {
    // a Siddon operator with tile products: deterministic, no atomics
    if (op->kind == OpKind::DYNAMIC && dyn_tile_apply(op, trans, x, y, s)) return;
    acc64 ? dyn_apply<float, double>(op, trans, x, y, s) : dyn_apply<float, float>(op, trans, x, y, s);
}
void dynamic_apply(tv3_op *op, bool trans, const double *x, double *y, hipStream_t s, bool) {
    dyn_apply<double, double>(op, trans, x, y, s);
}

// ------------------------------------------------------------- materialize
// Pass 1: each ray walks once into a fixed-capacity slot of a scratch buffer
// (capacity = the longest possible walk) and records its count; the counts
// are scanned into the row pointer; pass 2 compacts the slots into the CSR,
// converting the weights to the storage type.
__global__ __launch_bounds__(256) void k_walk_to_slots(GeoView g, int siddon, int cap,
                                                       int32_t *cnt, int32_t *scol,
                                                       float *sval) {
    int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= g.P) return;
    int c = 0;
    int32_t *oc = scol + k * cap;
    float *ov = sval + k * cap;
    if (siddon) {
        float p0[3], p1[3];
        int N[3] = {g.rvx, g.rvy, g.rvz};
        ray_ends<float>(g, k, p0, p1);
        siddon_walk<float>(p0, p1, g.org, g.h, N, g.rvy, [&](int64_t v, float w) {
            if (c < cap) { oc[c] = (int32_t)v; ov[c] = w; }
            ++c;
        });
    } else {
        int e = g.path[2 * k], r = g.path[2 * k + 1];
        bresenham_walk(g.ec[3 * e], g.ec[3 * e + 1], g.ec[3 * e + 2], g.rc[3 * r], g.rc[3 * r + 1],
                       g.rc[3 * r + 2], g.rvx, g.rvy, g.rvz, [&](int64_t v) {
                           if (c < cap) { oc[c] = (int32_t)v; ov[c] = g.scale; }
                           ++c;
                       });
    }
    cnt[k] = c < cap ? c : cap;
}

template <typename V> __device__ __forceinline__ V cvt(float w);
template <> __device__ __forceinline__ float cvt<float>(float w) { return w; }
template <> __device__ __forceinline__ __half cvt<__half>(float w) { return __float2half(w); }
template <> __device__ __forceinline__ __hip_bfloat16 cvt<__hip_bfloat16>(float w) {
    return __float2bfloat16(w);
}

template <typename V>
__global__ __launch_bounds__(256) void k_compact(int64_t P, int cap, const int32_t *rowptr,
                                                 const int32_t *scol, const float *sval,
                                                 int32_t *col, V *val) {
    int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= P) return;
    int b = rowptr[k], e = rowptr[k + 1];
    for (int j = 0; j < e - b; ++j) {
        col[b + j] = scol[k * cap + j];
        val[b + j] = cvt<V>(sval[k * cap + j]);
    }
}

void dynamic_materialize(tv3_op *op, tv3_dtype vt, SparseRows &A) {
    TV3_REQUIRE(vt == TV3_F32 || vt == TV3_F16 || vt == TV3_BF16, TV3_E_FORMAT,
                "materialized weights must be F32, F16 or BF16");
    tv3_ctx *ctx = op->ctx;
    hipStream_t s = ctx->stream;
    GeoView g = view(op->geo);
    int cap = g.rvx + g.rvy + g.rvz + 4;
    if (op->geo.mode == TV3_WEIGHT_BRESENHAM) cap = op->geo.max_walk;
    A.rows = g.P;
    A.cols = op->n;
    A.vtype = vt;
    A.ptr.alloc(g.P + 1);
    // process rays in chunks so the slot buffer stays bounded (<= ~1 GiB)
    int64_t chunk = (int64_t)((1ll << 30) / (8ll * cap));
    if (chunk < 1) chunk = 1;
    if (chunk > g.P) chunk = g.P > 0 ? g.P : 1;
    DevBuf<int32_t> cnt(g.P > 0 ? g.P : 1), scol(chunk * cap);
    DevBuf<float> sval(chunk * cap);
    int sid = op->geo.mode == TV3_WEIGHT_SIDDON;
    // count pass over all chunks (walk into slots, keep only counts)
    for (int64_t c0 = 0; c0 < g.P; c0 += chunk) {
        GeoView gc = g;
        int64_t cn = g.P - c0 < chunk ? g.P - c0 : chunk;
        gc.P = cn;
        gc.path = g.path + 2 * c0;
        hipLaunchKernelGGL(k_walk_to_slots, dim3((unsigned)((cn + 255) / 256)), dim3(256), 0, s, gc,
                           sid, cap, cnt.p + c0, scol.p, sval.p);
    }
    TV3_HIP(hipGetLastError());
    A.nnz = exclusive_scan_i32(ctx, cnt.p, A.ptr.p, g.P);
    A.ind.alloc(A.nnz > 0 ? A.nnz : 1);
    A.val.alloc((A.nnz > 0 ? A.nnz : 1) * dtype_size(vt));
    for (int64_t c0 = 0; c0 < g.P; c0 += chunk) {
        GeoView gc = g;
        int64_t cn = g.P - c0 < chunk ? g.P - c0 : chunk;
        gc.P = cn;
        gc.path = g.path + 2 * c0;
        unsigned bl = (unsigned)((cn + 255) / 256);
        hipLaunchKernelGGL(k_walk_to_slots, dim3(bl), dim3(256), 0, s, gc, sid, cap, cnt.p + c0,
                           scol.p, sval.p);
        const int32_t *rp = A.ptr.p + c0;
        switch (vt) {
        case TV3_F32:
            hipLaunchKernelGGL(k_compact<float>, dim3(bl), dim3(256), 0, s, cn, cap, rp, scol.p,
                               sval.p, A.ind.p, reinterpret_cast<float *>(A.val.p));
            break;
        case TV3_F16:
            hipLaunchKernelGGL(k_compact<__half>, dim3(bl), dim3(256), 0, s, cn, cap, rp, scol.p,
                               sval.p, A.ind.p, reinterpret_cast<__half *>(A.val.p));
            break;
        default:
            hipLaunchKernelGGL(k_compact<__hip_bfloat16>, dim3(bl), dim3(256), 0, s, cn, cap, rp,
                               scol.p, sval.p, A.ind.p,
                               reinterpret_cast<__hip_bfloat16 *>(A.val.p));
            break;
        }
    }
    TV3_HIP(hipGetLastError());
    TV3_HIP(hipStreamSynchronize(s));
    A.lanes = default_lanes(A.rows, A.nnz);
    build_row_blocks(ctx, A, 2048);
}

}  // namespace tv3
