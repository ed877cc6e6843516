// spmv_stream.hpp -- row-blocked CSR product for gfx950 (templated so the
// fused solver can attach its per-row epilogues).
//
// A 256-thread workgroup owns a block of consecutive rows whose entries fit
// in NB slots (NB/256 per thread).  Phase 1 streams the block's (col, val)
// pairs with unit-stride loads and stores the products in LDS.  Phase 2 gives
// each row a power-of-two group of lanes (as many as 256 / rows allows, at
// most 64), which sums the row's products from LDS and reduces them with
// wavefront butterflies; the row total goes to the epilogue.  A block holding
// a single row longer than NB (never produced for ray matrices, kept for
// generality) is reduced by the whole workgroup.
//
// These kernels are bound by memory latency, not bandwidth: every load of a
// block is issued before the first one is waited on.  Loads are clamped to
// the block's last entry instead of being guarded by branches (a guarded load
// cannot be hoisted, and each guard costs a full memory round trip); phase
// 2's row pointers and the epilogue's operands are requested with phase 1's
// loads.
#pragma once

#include <type_traits>

#include "tv3_internal.hpp"

namespace tv3 {

template <typename V> __device__ __forceinline__ float ld_weight(const V *p, int64_t k);
template <> __device__ __forceinline__ float ld_weight<float>(const float *p, int64_t k) { return p[k]; }
template <> __device__ __forceinline__ float ld_weight<__half>(const __half *p, int64_t k) {
    return __half2float(p[k]);
}
template <> __device__ __forceinline__ float ld_weight<__hip_bfloat16>(const __hip_bfloat16 *p,
                                                                       int64_t k) {
    return __bfloat162float(p[k]);
}

__device__ __forceinline__ float to_f32(float v) { return v; }
__device__ __forceinline__ float to_f32(__half v) { return __half2float(v); }
__device__ __forceinline__ float to_f32(__hip_bfloat16 v) { return __bfloat162float(v); }

// Workgroups are dispatched round-robin over the 8 XCDs; remap the block
// index so that each XCD walks a contiguous range of logical blocks (its L2
// then sees spatially neighbouring rows).  A bijection on [0, G).
__device__ __forceinline__ int xcd_remap(int bid, int G) {
    const int q = G >> 3, r = G & 7, x = bid & 7;
    return x * q + (x < r ? x : r) + (bid >> 3);
}

// Epilogue that stores y[row] = sum.
// Phase 2 of the row-blocked kernels: lane `lane` of an L-lane group sums
// prod[a + lane + i L] for a + lane + i L < e.  Reads go out in batches of
// four (clamped into the NB-entry buffer, masked after the read) so that a
// lane waits for one LDS round trip per batch instead of one per entry.
template <int NB, typename Acc>
__device__ __forceinline__ Acc row_sum(const Acc *prod, int a, int e, int lane, int L) {
    Acc s0 = 0, s1 = 0;
    for (int k = a + lane; k < e; k += 4 * L) {
        Acc v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = prod[min(k + i * L, NB - 1)];
        s0 += v[0] + (k + 2 * L < e ? v[2] : (Acc)0);
        s1 += (k + L < e ? v[1] : (Acc)0) + (k + 3 * L < e ? v[3] : (Acc)0);
    }
    return s0 + s1;
}

// Epilogue interface: load(row) fetches the row's operands early (returns a
// Pre value held in registers until the row total is known), row(row, sum,
// active, pre) consumes them, finish() runs once per workgroup, skip() lets a
// predicated launch exit at once.
template <typename Acc>
struct EpiStore {
    float *y;
    struct Pre {};
    __device__ __forceinline__ Pre load(int) const { return {}; }
    __device__ __forceinline__ void row(int r, Acc s, bool active, const Pre &) const {
        if (active) y[r] = (float)s;
    }
    __device__ __forceinline__ void finish() const {}
    __device__ bool skip() const { return false; }
};

template <int NB, typename V, typename Acc, typename Epi>
__global__ __launch_bounds__(256) void k_csr_stream(const int32_t *__restrict__ blk,
                                                    const int32_t *__restrict__ bptr,
                                                    const int32_t *__restrict__ ptr,
                                                    const int32_t *__restrict__ ind,
                                                    const V *__restrict__ val,
                                                    const float *__restrict__ x, Epi epi,
                                                    int nblk) {
    __shared__ Acc prod[NB];
    if (epi.skip()) return;
    // each workgroup walks a contiguous chunk of row blocks; consecutive
    // chunks stay on one XCD (xcd_remap)
    const int G = gridDim.x;
    const int chunk = (nblk + G - 1) / G;
    const int w = xcd_remap(blockIdx.x, G);
    const int b0 = w * chunk, b1 = min(nblk, b0 + chunk);
    const int t = threadIdx.x;
    for (int b = b0; b < b1; ++b) {
        const int r0 = blk[b], r1 = blk[b + 1];
        const int p0 = bptr[b], p1 = bptr[b + 1];
        const int len = p1 - p0;
        if (len <= NB) {
            const int nr = r1 - r0;                  // <= 256 by construction
            int L = 64;
            while (L > 1 && L * nr > 256) L >>= 1;
            const int rr = t / L, lane = t & (L - 1);
            const bool has = rr < nr;
            const int row = r0 + (has ? rr : nr - 1);
            const int a = ptr[row] - p0, e = ptr[row + 1] - p0;
            const auto pre = epi.load(row);
            if (len > 0) {
                constexpr int J = NB / 256;
                const int last = p0 + len - 1;
                int c[J];
                float w[J];
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    const int k = min(p0 + j * 256 + t, last);
                    c[j] = ind[k];
                    w[j] = ld_weight<V>(val, k);
                }
#pragma unroll
                for (int j = 0; j < J; ++j) prod[j * 256 + t] = (Acc)w[j] * (Acc)x[c[j]];
            }
            __syncthreads();
            Acc s = 0;
            if (has) s = row_sum<NB>(prod, a, e, lane, L);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                Acc o = __shfl_xor(s, off, 64);
                if (off < L) s += o;
            }
            epi.row(row, s, has && lane == 0, pre);
        } else {
            Acc s = 0;
            const auto pre = epi.load(r0);
            for (int k = p0 + t; k < p1; k += 256) s += (Acc)ld_weight<V>(val, k) * (Acc)x[ind[k]];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            if ((t & 63) == 0) prod[t >> 6] = s;
            __syncthreads();
            if (t == 0) s = (prod[0] + prod[1]) + (prod[2] + prod[3]);
            epi.row(r0, s, t == 0, pre);
        }
        __syncthreads();             // prod is rewritten by the next row block
    }
    epi.finish();
}

// Largest grid of the product kernels (beyond it a workgroup walks a chunk
// of row blocks).  Measured on MI355X: capping at 2048 workgroups costs ~40 %
// on 128^3 ray matrices (static chunks lose the dispatcher's load balance).
constexpr int kMaxProductGrid = 1 << 30;

// Row-group kernel with an epilogue: L lanes own one row (L divides 64).
template <int L, typename V, typename Acc, typename Epi>
__global__ __launch_bounds__(256) void k_csr_rows(int64_t rows, const int32_t *__restrict__ ptr,
                                                  const int32_t *__restrict__ ind,
                                                  const V *__restrict__ val,
                                                  const float *__restrict__ x, Epi epi,
                                                  int64_t nlogical) {
    if (epi.skip()) return;
    const int64_t G = gridDim.x;
    const int64_t chunk = (nlogical + G - 1) / G;
    const int64_t w = xcd_remap(blockIdx.x, (int)G);
    const int64_t l0 = w * chunk, l1 = min(nlogical, l0 + chunk);
    const int lane = threadIdx.x & (L - 1);
    for (int64_t lb = l0; lb < l1; ++lb) {
        const int64_t row0 = (lb * 256 + threadIdx.x) / L;
        const bool has = row0 < rows;
        const int64_t row = has ? row0 : rows - 1;
        const int beg = ptr[row], end = has ? ptr[row + 1] : beg;
        const auto pre = epi.load((int)row);
        // batches of 4 entries per lane, loads clamped to the row's last entry
        Acc s = 0, s2 = 0;
        for (int k0 = beg + lane; k0 < end; k0 += 4 * L) {
            int c[4];
            float w[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = min(k0 + i * L, end - 1);
                c[i] = ind[k];
                w[i] = ld_weight<V>(val, k);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) w[i] = k0 + i * L < end ? w[i] : 0.f;
            s += (Acc)w[0] * (Acc)x[c[0]] + (Acc)w[2] * (Acc)x[c[2]];
            s2 += (Acc)w[1] * (Acc)x[c[1]] + (Acc)w[3] * (Acc)x[c[3]];
        }
        s += s2;
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
        epi.row((int)row, s, has && lane == 0, pre);
    }
    epi.finish();
}

// Panel product of A^T (rows = voxels) with epilogue: one workgroup per panel
// (largest panels first, so the tail of the launch is short).  The panel's
// rays are gathered from x into LDS once; phase 1 of each row block then
// reads x through the 16-bit local index from LDS instead of issuing one
// scattered global load per entry (4 entries per 16-byte load), phase 2 is
// the stream kernel's.

// weight of a packed entry (bits 16..31)
template <typename V> __device__ __forceinline__ float pk_weight(uint32_t u);
template <> __device__ __forceinline__ float pk_weight<__half>(uint32_t u) {
    return __half2float(__ushort_as_half((unsigned short)(u >> 16)));
}
template <> __device__ __forceinline__ float pk_weight<__hip_bfloat16>(uint32_t u) {
    return __uint_as_float(u & 0xffff0000u);
}

// entries q .. q+3 of the padded arrays (q a multiple of 4)
template <typename V>
__device__ __forceinline__ void panel_ld4(const uint32_t *pk, const uint16_t *, const float *,
                                          int64_t q, int c[4], float w[4]) {
    const uint4 u = *reinterpret_cast<const uint4 *>(pk + q);
    const uint32_t v[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) { c[i] = (int)(v[i] & 0xffffu); w[i] = pk_weight<V>(v[i]); }
}
template <>
__device__ __forceinline__ void panel_ld4<float>(const uint32_t *, const uint16_t *ps,
                                                 const float *pv, int64_t q, int c[4], float w[4]) {
    const uint2 u = *reinterpret_cast<const uint2 *>(ps + q);
    const float4 f = *reinterpret_cast<const float4 *>(pv + q);
    c[0] = (int)(u.x & 0xffffu); c[1] = (int)(u.x >> 16);
    c[2] = (int)(u.y & 0xffffu); c[3] = (int)(u.y >> 16);
    w[0] = f.x; w[1] = f.y; w[2] = f.z; w[3] = f.w;
}
template <typename V>
__device__ __forceinline__ void panel_ld1(const uint32_t *pk, const uint16_t *, const float *,
                                          int64_t q, int &c, float &w) {
    const uint32_t u = pk[q];
    c = (int)(u & 0xffffu);
    w = pk_weight<V>(u);
}
template <>
__device__ __forceinline__ void panel_ld1<float>(const uint32_t *, const uint16_t *ps,
                                                 const float *pv, int64_t q, int &c, float &w) {
    c = ps[q];
    w = pv[q];
}

template <int NB, typename V, typename Acc, typename Epi>
__global__ __launch_bounds__(256) void k_panel_t(const int32_t *__restrict__ order,
                                                 const int32_t *__restrict__ pan_seg,
                                                 const int32_t *__restrict__ pan_blk,
                                                 const int32_t *__restrict__ seg_ray,
                                                 const int32_t *__restrict__ blk,
                                                 const int32_t *__restrict__ bptr,
                                                 const int32_t *__restrict__ bpad,
                                                 const int32_t *__restrict__ ptr,
                                                 const uint32_t *__restrict__ pk,
                                                 const uint16_t *__restrict__ ps,
                                                 const float *__restrict__ pv,
                                                 const float *__restrict__ x, Epi epi) {
    static_assert(std::is_same<Acc, float>::value, "panel products accumulate in FP32");
    extern __shared__ float xs[];            // the panel's ray values (max_seg)
    __shared__ __align__(16) Acc prod[NB];
    if (epi.skip()) return;
    const int t = threadIdx.x;
    const int pan = order[blockIdx.x];
    const int s0 = pan_seg[pan], ns = pan_seg[pan + 1] - s0;
    // stage the panel's rays: batches of 8 per thread, all loads in flight
    for (int k0 = 0; k0 < ns; k0 += 8 * 256) {
        int rv[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) rv[i] = seg_ray[s0 + min(k0 + i * 256 + t, ns - 1)];
        float xv[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) xv[i] = x[rv[i]];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (k0 + i * 256 + t < ns) xs[k0 + i * 256 + t] = xv[i];
    }
    __syncthreads();
    const int b1 = pan_blk[pan + 1];
    for (int b = pan_blk[pan]; b < b1; ++b) {
        const int r0 = blk[b], r1 = blk[b + 1];
        const int p0 = bptr[b], len = bptr[b + 1] - p0;
        const int64_t q0 = bpad[b];
        if (len <= NB) {
            const int nr = r1 - r0;                  // <= 256 by construction
            int L = 64;
            while (L > 1 && L * nr > 256) L >>= 1;
            const int rr = t / L, lane = t & (L - 1);
            const bool has = rr < nr;
            const int row = r0 + (has ? rr : nr - 1);
            const int a = ptr[row] - p0, e = ptr[row + 1] - p0;
            const auto pre = epi.load(row);
            if (len > 0) {
                constexpr int J = NB / 1024;         // 16-byte loads per thread
                const int qlast = (len - 1) & ~3;
                int c[J][4];
                float w[J][4];
#pragma unroll
                for (int j = 0; j < J; ++j)
                    panel_ld4<V>(pk, ps, pv, q0 + min(j * 1024 + 4 * t, qlast), c[j], w[j]);
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    float4 o;
                    o.x = w[j][0] * xs[c[j][0]];
                    o.y = w[j][1] * xs[c[j][1]];
                    o.z = w[j][2] * xs[c[j][2]];
                    o.w = w[j][3] * xs[c[j][3]];
                    *reinterpret_cast<float4 *>(&prod[j * 1024 + 4 * t]) = o;
                }
            }
            __syncthreads();
            Acc s = 0;
            if (has) s = row_sum<NB>(prod, a, e, lane, L);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                Acc o = __shfl_xor(s, off, 64);
                if (off < L) s += o;
            }
            epi.row(row, s, has && lane == 0, pre);
        } else {
            // one row longer than NB
            const auto pre = epi.load(r0);
            Acc s = 0;
            for (int k = t; k < len; k += 256) {
                int c;
                float w;
                panel_ld1<V>(pk, ps, pv, q0 + k, c, w);
                s += (Acc)w * (Acc)xs[c];
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            if ((t & 63) == 0) prod[t >> 6] = s;
            __syncthreads();
            if (t == 0) s = (prod[0] + prod[1]) + (prod[2] + prod[3]);
            epi.row(r0, s, t == 0, pre);
        }
        __syncthreads();             // prod is rewritten by the next row block
    }
    epi.finish();
}

template <typename V, typename Acc, typename Epi>
void launch_panel_v(const PanelT &P, const SparseRows &S, const float *x, const Epi &epi,
                    hipStream_t s) {
    const unsigned g = (unsigned)P.npan;
    const size_t lds = sizeof(float) * (size_t)(P.max_seg > 0 ? P.max_seg : 1);
    if (g == 0) return;
#define TV3_PCASE(NBV)                                                                        \
    case NBV:                                                                                 \
        hipLaunchKernelGGL((k_panel_t<NBV, V, Acc, Epi>), dim3(g), dim3(256), lds, s,           \
                           P.order.p, P.pan_seg.p, P.pan_blk.p, P.seg_ray.p, P.blk.p, P.bptr.p, \
                           P.bpad.p, S.ptr.p, P.pk.p, P.sidx.p, P.pval.p, x, epi);            \
        break;
    switch (P.blk_nnz) {
        TV3_PCASE(1024) TV3_PCASE(2048) TV3_PCASE(4096)
    default: throw tv3_error(TV3_E_ARG, "panel block size must be 1024, 2048 or 4096");
    }
#undef TV3_PCASE
}

// Panel product of S (= A^T with panels P) with epilogue epi.
template <typename Acc, typename Epi>
void launch_panel(const PanelT &P, const SparseRows &S, const float *x, const Epi &epi,
                  hipStream_t s) {
    switch (S.vtype) {
    case TV3_F32: launch_panel_v<float, Acc, Epi>(P, S, x, epi, s); break;
    case TV3_F16: launch_panel_v<__half, Acc, Epi>(P, S, x, epi, s); break;
    case TV3_BF16: launch_panel_v<__hip_bfloat16, Acc, Epi>(P, S, x, epi, s); break;
    default: throw tv3_error(TV3_E_FORMAT, "FP32 products need F32, F16 or BF16 values");
    }
}

// Launch S's product (stream kernel when S.blk_nnz != 0, else the row-group
// kernel with S.lanes) with epilogue epi.
template <typename V, typename Acc, typename Epi>
void launch_rows(const SparseRows &S, const float *x, const Epi &epi, hipStream_t s) {
    const V *val = reinterpret_cast<const V *>(S.val.p);
    const int L = S.lanes ? S.lanes : 32;
    const int64_t nl = (S.rows * L + 255) / 256;
    const unsigned g = (unsigned)(nl < kMaxProductGrid ? nl : kMaxProductGrid);
    if (g == 0) return;
#define TV3_RCASE(LL)                                                                     \
    case LL:                                                                              \
        hipLaunchKernelGGL((k_csr_rows<LL, V, Acc, Epi>), dim3(g), dim3(256), 0, s, S.rows, \
                           S.ptr.p, S.ind.p, val, x, epi, nl);                            \
        break;
    switch (L) {
        TV3_RCASE(2) TV3_RCASE(4) TV3_RCASE(8) TV3_RCASE(16) TV3_RCASE(32) TV3_RCASE(64)
    default: throw tv3_error(TV3_E_ARG, "lanes per row must be a power of two in [2, 64]");
    }
#undef TV3_RCASE
}

// Launch the stream kernel of S with epilogue epi (S.blk must be built).
template <typename V, typename Acc, typename Epi>
void launch_stream(const SparseRows &S, const float *x, const Epi &epi, hipStream_t s) {
    const V *val = reinterpret_cast<const V *>(S.val.p);
    const int nb = (int)S.nblk;
    const unsigned g = (unsigned)(nb < kMaxProductGrid ? nb : kMaxProductGrid);
    if (g == 0) return;
    switch (S.blk_nnz) {
    case 1024:
        hipLaunchKernelGGL((k_csr_stream<1024, V, Acc, Epi>), dim3(g), dim3(256), 0, s, S.blk.p,
                           S.bptr.p, S.ptr.p, S.ind.p, val, x, epi, nb);
        break;
    case 2048:
        hipLaunchKernelGGL((k_csr_stream<2048, V, Acc, Epi>), dim3(g), dim3(256), 0, s, S.blk.p,
                           S.bptr.p, S.ptr.p, S.ind.p, val, x, epi, nb);
        break;
    case 4096:
        hipLaunchKernelGGL((k_csr_stream<4096, V, Acc, Epi>), dim3(g), dim3(256), 0, s, S.blk.p,
                           S.bptr.p, S.ptr.p, S.ind.p, val, x, epi, nb);
        break;
    default: throw tv3_error(TV3_E_ARG, "stream block size must be 1024, 2048 or 4096");
    }
}

template <typename Acc, typename Epi>
void launch_stream_any(const SparseRows &S, const float *x, const Epi &epi, hipStream_t s) {
    switch (S.vtype) {
    case TV3_F32: launch_stream<float, Acc, Epi>(S, x, epi, s); break;
    case TV3_F16: launch_stream<__half, Acc, Epi>(S, x, epi, s); break;
    case TV3_BF16: launch_stream<__hip_bfloat16, Acc, Epi>(S, x, epi, s); break;
    default: throw tv3_error(TV3_E_FORMAT, "FP32 products need F32, F16 or BF16 values");
    }
}

// Product of S with epilogue, dispatched on the tuned kernel and value type.
template <typename Acc, typename Epi>
void launch_product(const SparseRows &S, const float *x, const Epi &epi, hipStream_t s) {
    if (S.blk_nnz) {
        launch_stream_any<Acc, Epi>(S, x, epi, s);
        return;
    }
    switch (S.vtype) {
    case TV3_F32: launch_rows<float, Acc, Epi>(S, x, epi, s); break;
    case TV3_F16: launch_rows<__half, Acc, Epi>(S, x, epi, s); break;
    case TV3_BF16: launch_rows<__hip_bfloat16, Acc, Epi>(S, x, epi, s); break;
    default: throw tv3_error(TV3_E_FORMAT, "FP32 products need F32, F16 or BF16 values");
    }
}

// A^T product of a fused layout: panels when tuned in, else the plain kernel.
template <typename Acc, typename Epi>
void launch_transposed(const FusedLayout &F, const float *x, const Epi &epi, hipStream_t s) {
    if (F.use_panels) launch_panel<Acc, Epi>(F.PT, F.AT, x, epi, s);
    else launch_product<Acc, Epi>(F.AT, x, epi, s);
}

}  // namespace tv3
