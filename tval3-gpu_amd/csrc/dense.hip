// dense.hip -- column-major dense products (gemv N and T), the semantics of
// dense_mm / mat_vec_mul (3d/mat_vec_mul_dense.h:25-73: gemv with alpha = 1,
// beta = 0, leading dimension lda).
//
// N (y = A x): a thread owns a row and walks a slab of columns, so a
// wavefront reads 64 consecutive entries of one column (coalesced); the
// column range is split over gridDim.y slabs and the slab partials are summed
// in a fixed order by a second kernel (deterministic, no atomics).
// T (y = A^T x): a wavefront owns a column; lanes read it contiguously.
#include "tv3_internal.hpp"

namespace tv3 {

template <typename V> __device__ __forceinline__ float ldv(const V *p, int64_t k);
template <> __device__ __forceinline__ float ldv<float>(const float *p, int64_t k) { return p[k]; }
template <> __device__ __forceinline__ float ldv<__half>(const __half *p, int64_t k) {
    return __half2float(p[k]);
}
template <> __device__ __forceinline__ float ldv<__hip_bfloat16>(const __hip_bfloat16 *p, int64_t k) {
    return __bfloat162float(p[k]);
}
__device__ __forceinline__ double ldd(const double *p, int64_t k) { return p[k]; }
__device__ __forceinline__ double ldd(const float *p, int64_t k) { return (double)p[k]; }

template <typename V, typename AccT>
__device__ __forceinline__ AccT loadA(const V *A, int64_t k) {
    if constexpr (std::is_same<V, double>::value) return (AccT)A[k];
    else if constexpr (std::is_same<V, float>::value) return (AccT)A[k];
    else return (AccT)ldv<V>(A, k);
}

template <typename V, typename T, typename AccT>
__global__ __launch_bounds__(256) void k_gemv_n_part(int64_t m, int64_t n, int64_t lda,
                                                     const V *__restrict__ A,
                                                     const T *__restrict__ x,
                                                     AccT *__restrict__ part, int64_t cols_per) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t j0 = (int64_t)blockIdx.y * cols_per, j1 = j0 + cols_per < n ? j0 + cols_per : n;
    if (i >= m) return;
    AccT s = 0;
    for (int64_t j = j0; j < j1; ++j) s += loadA<V, AccT>(A, i + j * lda) * (AccT)x[j];
    part[(int64_t)blockIdx.y * m + i] = s;
}

template <typename T, typename AccT>
__global__ void k_gemv_n_sum(int64_t m, int ks, const AccT *part, T *y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    AccT s = 0;
    for (int k = 0; k < ks; ++k) s += part[(int64_t)k * m + i];
    y[i] = (T)s;
}

template <typename V, typename T, typename AccT>
__global__ __launch_bounds__(256) void k_gemv_t(int64_t m, int64_t n, int64_t lda,
                                                const V *__restrict__ A,
                                                const T *__restrict__ x, T *__restrict__ y) {
    int64_t j = ((int64_t)blockIdx.x * 256 + threadIdx.x) / 64;
    int lane = threadIdx.x & 63;
    if (j >= n) return;
    const V *col = A + j * lda;
    AccT s = 0;
    for (int64_t i = lane; i < m; i += 64) s += loadA<V, AccT>(col, i) * (AccT)x[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[j] = (T)s;
}

// ws: the split-K partials of A x go to this workspace of ws_bytes (the
// allocation-free path) instead of the context's, which grows on demand
template <typename V, typename T, typename AccT>
static void gemv(tv3_op *op, bool trans, const T *x, T *y, hipStream_t s, void *ws = nullptr,
                 size_t ws_bytes = 0) {
    const V *A = reinterpret_cast<const V *>(op->dense.p);
    int64_t m = op->m, n = op->n, lda = op->lda;
    if (!trans) {
        int ks = (int)((n + 63) / 64);
        if (ks > 128) ks = 128;
        if (ks < 1) ks = 1;
        int64_t cols_per = (n + ks - 1) / ks;
        ks = (int)((n + cols_per - 1) / cols_per);
        size_t need = sizeof(AccT) * (size_t)ks * (size_t)m;
        AccT *part;
        if (ws) {
            TV3_REQUIRE(need <= ws_bytes, TV3_E_UNSUPPORTED,
                        "dense A x needs " + std::to_string(need) + " bytes of partial sums; the "
                        "allocation-free path holds " + std::to_string(ws_bytes));
            part = reinterpret_cast<AccT *>(ws);
        } else {
            if (op->ctx->scratch.bytes() < need) op->ctx->scratch.alloc(need);
            part = reinterpret_cast<AccT *>(op->ctx->scratch.p);
        }
        hipLaunchKernelGGL((k_gemv_n_part<V, T, AccT>), dim3((unsigned)((m + 255) / 256), ks), dim3(256),
                           0, s, m, n, lda, A, x, part, cols_per);
        hipLaunchKernelGGL((k_gemv_n_sum<T, AccT>), dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m,
                           ks, part, y);
    } else {
        hipLaunchKernelGGL((k_gemv_t<V, T, AccT>), dim3((unsigned)((n * 64 + 255) / 256)), dim3(256), 0,
                           s, m, n, lda, A, x, y);
    }
    TV3_HIP(hipGetLastError());
}

void dense_gemv(tv3_op *op, bool trans, const float *x, float *y, hipStream_t s, bool acc64) {
    switch (op->dtype) {
    case TV3_F32:
        acc64 ? gemv<float, float, double>(op, trans, x, y, s) : gemv<float, float, float>(op, trans, x, y, s);
        break;
    case TV3_F16:
        acc64 ? gemv<__half, float, double>(op, trans, x, y, s)
              : gemv<__half, float, float>(op, trans, x, y, s);
        break;
    case TV3_BF16:
        acc64 ? gemv<__hip_bfloat16, float, double>(op, trans, x, y, s)
              : gemv<__hip_bfloat16, float, float>(op, trans, x, y, s);
        break;
    default: throw tv3_error(TV3_E_FORMAT, "FP32 dense products need F32, F16 or BF16 values");
    }
}

void dense_gemv_ws(tv3_op *op, const float *x, float *y, hipStream_t s, void *ws,
                   size_t ws_bytes) {
    switch (op->dtype) {
    case TV3_F32: gemv<float, float, float>(op, false, x, y, s, ws, ws_bytes); break;
    case TV3_F16: gemv<__half, float, float>(op, false, x, y, s, ws, ws_bytes); break;
    case TV3_BF16: gemv<__hip_bfloat16, float, float>(op, false, x, y, s, ws, ws_bytes); break;
    default: throw tv3_error(TV3_E_FORMAT, "FP32 dense products need F32, F16 or BF16 values");
    }
}

void dense_gemv(tv3_op *op, bool trans, const double *x, double *y, hipStream_t s, bool) {
    switch (op->dtype) {
    case TV3_F64: gemv<double, double, double>(op, trans, x, y, s); break;
    case TV3_F32: gemv<float, double, double>(op, trans, x, y, s); break;
    default: throw tv3_error(TV3_E_FORMAT, "FP64 dense products need F64 or F32 values");
    }
}

}  // namespace tv3
