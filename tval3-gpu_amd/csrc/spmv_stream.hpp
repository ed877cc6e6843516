// spmv_stream.hpp -- row-blocked CSR product for gfx950 (templated so the
// fused solver can attach its per-row epilogues).
//
// A 256-thread workgroup owns a block of consecutive rows whose entries fit
// in NB slots (NB/256 per thread).  Phase 1 streams the block's (col, val)
// pairs with unit-stride loads -- every thread has NB/256 independent gathers
// of x in flight -- and stores the products in LDS.  Phase 2 gives each row a
// power-of-two group of lanes (as many as 256 / rows allows, at most 64),
// which sums the row's products from LDS and reduces them with wavefront
// butterflies; the row total goes to the epilogue.  A block holding a single
// row longer than NB (never produced for ray matrices, kept for generality)
// is reduced by the whole workgroup.
#pragma once

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

// Workgroups are dispatched round-robin over the 8 XCDs; remap the block
// index so that each XCD walks a contiguous range of logical blocks (its L2
// then sees spatially neighbouring rows).  A bijection on [0, G).
__device__ __forceinline__ int xcd_remap(int bid, int G) {
    const int q = G >> 3, r = G & 7, x = bid & 7;
    return x * q + (x < r ? x : r) + (bid >> 3);
}

// Epilogue that stores y[row] = sum.
template <typename Acc>
struct EpiStore {
    float *y;
    __device__ __forceinline__ void row(int r, Acc s, bool active) const {
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
#pragma unroll
            for (int j = 0; j < NB / 256; ++j) {
                const int k = j * 256 + t;
                if (k < len) prod[k] = (Acc)ld_weight<V>(val, p0 + k) * (Acc)x[ind[p0 + k]];
            }
            __syncthreads();
            const int nr = r1 - r0;                  // <= 256 by construction
            int L = 64;
            while (L > 1 && L * nr > 256) L >>= 1;
            const int rr = t / L, lane = t & (L - 1);
            const bool has = rr < nr;
            const int row = r0 + (has ? rr : 0);
            Acc s = 0;
            if (has) {
                const int a = ptr[row] - p0, e = ptr[row + 1] - p0;
                for (int k = a + lane; k < e; k += L) s += prod[k];
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                Acc o = __shfl_xor(s, off, 64);
                if (off < L) s += o;
            }
            epi.row(row, s, has && lane == 0);
        } else {
            Acc s = 0;
            for (int k = p0 + t; k < p1; k += 256) s += (Acc)ld_weight<V>(val, k) * (Acc)x[ind[k]];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            if ((t & 63) == 0) prod[t >> 6] = s;
            __syncthreads();
            if (t == 0) s = (prod[0] + prod[1]) + (prod[2] + prod[3]);
            epi.row(r0, s, t == 0);
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
        const int64_t row = (lb * 256 + threadIdx.x) / L;
        const bool has = row < rows;
        Acc s = 0;
        if (has) {
            const int beg = ptr[row], end = ptr[row + 1];
            Acc s2 = 0;
            int k = beg + lane;
            for (; k + L < end; k += 2 * L) {
                const int c0 = ind[k], c1 = ind[k + L];
                s += (Acc)ld_weight<V>(val, k) * (Acc)x[c0];
                s2 += (Acc)ld_weight<V>(val, k + L) * (Acc)x[c1];
            }
            if (k < end) s += (Acc)ld_weight<V>(val, k) * (Acc)x[ind[k]];
            s += s2;
        }
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
        epi.row((int)(has ? row : 0), s, has && lane == 0);
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

}  // namespace tv3
