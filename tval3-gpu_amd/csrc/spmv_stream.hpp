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
#include <utility>

#include "lanes.hpp"
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

// An epilogue may also define skip_rows(first_row): true skips a whole row
// block (the stream kernel) or row group (the row-group kernel) -- the
// batched solver's way of leaving out the rows of finished problems.  Row
// blocks and groups must then not cross the boundaries it tests.
template <typename E, typename = void>
struct HasSkipRows : std::false_type {};
template <typename E>
struct HasSkipRows<E, std::void_t<decltype(std::declval<const E &>().skip_rows(0))>>
    : std::true_type {};

// Phase 2 and the epilogue of one row block whose products are in prod.
template <int NB, typename Acc, typename Epi>
__device__ __forceinline__ void block_rows(const Acc *prod, const BlkDesc &d, int a, int e,
                                           const Epi &epi, int row,
                                           const typename Epi::Pre &pre) {
    const int t = threadIdx.x;
    const int L = 1 << d.lg, rr = t >> d.lg, lane = t & (L - 1);
    const bool has = rr < d.nr;
    Acc s = 0;
    if (has) s = row_sum<NB>(prod, a, e, lane, L);
    s = lane_sum(s, L);
    epi.row(row, s, has && lane == 0, pre);
}

// the row a thread serves in phase 2 (clamped to the block's last row)
__device__ __forceinline__ int block_row(const BlkDesc &d) {
    const int rr = threadIdx.x >> d.lg;
    return d.r0 + (rr < d.nr ? rr : d.nr - 1);
}

template <int NB, typename V, typename Acc, typename Epi>
__global__ __launch_bounds__(256) void k_csr_stream(const BlkDesc *__restrict__ desc,
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
    const int wg = xcd_remap(blockIdx.x, G);
    const int b0 = wg * chunk, b1 = min(nblk, b0 + chunk);
    const int t = threadIdx.x;
    if (b0 >= b1) { epi.finish(); return; }
    BlkDesc d = desc[b0];
    for (int b = b0; b < b1; ++b) {
        const BlkDesc dn = desc[b + 1];      // in flight while this block runs
        if constexpr (HasSkipRows<Epi>::value) {
            if (epi.skip_rows(d.r0)) {       // uniform over the workgroup
                d = dn;
                continue;
            }
        }
        if (d.len <= NB) {
            const int row = block_row(d);
            const int a = ptr[row] - d.p0, e = ptr[row + 1] - d.p0;
            const auto pre = epi.load(row);
            if (d.len > 0) {
                constexpr int J = NB / 256;
                const int last = d.p0 + d.len - 1;
                int c[J];
                float w[J];
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    const int k = min(d.p0 + j * 256 + t, last);
                    c[j] = ind[k];
                    w[j] = ld_weight<V>(val, k);
                }
#pragma unroll
                for (int j = 0; j < J; ++j) prod[j * 256 + t] = (Acc)w[j] * (Acc)x[c[j]];
            }
            __syncthreads();
            block_rows<NB>(prod, d, a, e, epi, row, pre);
        } else {
            // one row longer than NB
            const auto pre = epi.load(d.r0);
            Acc s = 0;
            for (int k = d.p0 + t; k < d.p0 + d.len; k += 256)
                s += (Acc)ld_weight<V>(val, k) * (Acc)x[ind[k]];
            s = lane_sum<64>(s);
            if ((t & 63) == 0) prod[t >> 6] = s;
            __syncthreads();
            if (t == 0) s = (prod[0] + prod[1]) + (prod[2] + prod[3]);
            epi.row(d.r0, s, t == 0, pre);
        }
        __syncthreads();             // prod is rewritten by the next row block
        d = dn;
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
        if constexpr (HasSkipRows<Epi>::value) {
            if (epi.skip_rows((int)(lb * (256 / L)))) continue;   // uniform
        }
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
        s = lane_sum<L>(s + s2);
        epi.row((int)row, s, has && lane == 0, pre);
    }
    epi.finish();
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
        hipLaunchKernelGGL((k_csr_stream<1024, V, Acc, Epi>), dim3(g), dim3(256), 0, s, S.desc.p,
                           S.ptr.p, S.ind.p, val, x, epi, nb);
        break;
    case 2048:
        hipLaunchKernelGGL((k_csr_stream<2048, V, Acc, Epi>), dim3(g), dim3(256), 0, s, S.desc.p,
                           S.ptr.p, S.ind.p, val, x, epi, nb);
        break;
    case 4096:
        hipLaunchKernelGGL((k_csr_stream<4096, V, Acc, Epi>), dim3(g), dim3(256), 0, s, S.desc.p,
                           S.ptr.p, S.ind.p, val, x, epi, nb);
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

}  // namespace tv3
