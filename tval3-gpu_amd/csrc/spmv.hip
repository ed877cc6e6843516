// spmv.hip -- CSR products y = A x and y = A^T x (via the CSC, stored as the
// CSR of A^T) for gfx950, plus the device scan and the stable transpose.
//
// Semantics follow the single-GPU sparse operator of the reference
// (3d/mat_vec_mul_sparse.h:1703-1780: csrmv with alpha = 1, beta = 0 on the
// CSR for N and on the CSC arrays for T); the kernels are new.
//
// Kernel shape: a group of L lanes (L a power of two, 2..64) owns one row.
// Lanes stride over the row's entries, so one group reads L consecutive
// (col, val) pairs per step and a 64-wide wavefront reads 64/L rows' worth of
// consecutive memory: the col/val streams are fully coalesced whatever L is.
// x is gathered (it stays in L2 / MALL); the group sum is a butterfly over L
// lanes.  L is chosen from the mean row length or by the autotuner.
#include "spmv_stream.hpp"

namespace tv3 {

template <typename V> __device__ __forceinline__ float ldw(const V *p, int64_t k);
template <> __device__ __forceinline__ float ldw<float>(const float *p, int64_t k) { return p[k]; }
template <> __device__ __forceinline__ float ldw<__half>(const __half *p, int64_t k) {
    return __half2float(p[k]);
}
template <> __device__ __forceinline__ float ldw<__hip_bfloat16>(const __hip_bfloat16 *p, int64_t k) {
    return __bfloat162float(p[k]);
}
__device__ __forceinline__ double ldw64(const double *p, int64_t k) { return p[k]; }
__device__ __forceinline__ double ldw64(const float *p, int64_t k) { return (double)p[k]; }

// Acc = float: FP32 accumulation (the fast path).  Acc = double: each row
// sum is formed in FP64 and rounded once, which restates the oracle's
// rounding (used by the reference-shaped solver).
template <int L, typename V, typename Acc>
__global__ __launch_bounds__(256) void k_csr_f32(int64_t rows, const int32_t *__restrict__ ptr,
                                                 const int32_t *__restrict__ ind,
                                                 const V *__restrict__ val,
                                                 const float *__restrict__ x,
                                                 float *__restrict__ y) {
    const int64_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t row = (bid * 256 + threadIdx.x) / L;
    const int lane = threadIdx.x & (L - 1);
    if (row >= rows) return;          // whole groups leave together (L divides 256)
    const int beg = ptr[row], end = ptr[row + 1];
    Acc s = 0;
    int k = beg + lane;
    // two independent chains keep two gathers in flight per lane
    Acc s2 = 0;
    for (; k + L < end; k += 2 * L) {
        const int c0 = ind[k], c1 = ind[k + L];
        const Acc w0 = (Acc)ldw<V>(val, k), w1 = (Acc)ldw<V>(val, k + L);
        s = fma(w0, (Acc)x[c0], s);
        s2 = fma(w1, (Acc)x[c1], s2);
    }
    if (k < end) s = fma((Acc)ldw<V>(val, k), (Acc)x[ind[k]], s);
    s = lane_sum<L>(s + s2);
    if (lane == 0) y[row] = (float)s;
}

template <int L, typename V>
__global__ __launch_bounds__(256) void k_csr_f64(int64_t rows, const int32_t *__restrict__ ptr,
                                                 const int32_t *__restrict__ ind,
                                                 const V *__restrict__ val,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ y) {
    const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) / L;
    const int lane = threadIdx.x & (L - 1);
    if (row >= rows) return;
    const int beg = ptr[row], end = ptr[row + 1];
    double s = 0.0;
    for (int k = beg + lane; k < end; k += L) s = fma(ldw64(val, k), x[ind[k]], s);
    s = lane_sum<L>(s);
    if (lane == 0) y[row] = s;
}

int default_lanes(int64_t rows, int64_t nnz) {
    double avg = rows > 0 ? (double)nnz / (double)rows : 0.0;
    int L = 2;
    while (L < 64 && L < avg / 2.0) L <<= 1;
    return L;
}

template <typename V, typename Acc>
static void launch_f32(const SparseRows &S, const float *x, float *y, hipStream_t s, int L) {
    const V *val = reinterpret_cast<const V *>(S.val.p);
    const int64_t threads = S.rows * L;
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    if (blocks == 0) return;
#define TV3_CASE(LL)                                                                      \
    case LL:                                                                              \
        hipLaunchKernelGGL((k_csr_f32<LL, V, Acc>), dim3(blocks), dim3(256), 0, s, S.rows,\
                           S.ptr.p, S.ind.p, val, x, y);                                  \
        break;
    switch (L) {
        TV3_CASE(2) TV3_CASE(4) TV3_CASE(8) TV3_CASE(16) TV3_CASE(32) TV3_CASE(64)
    default: throw tv3_error(TV3_E_ARG, "lanes per row must be a power of two in [2, 64]");
    }
#undef TV3_CASE
    TV3_HIP(hipGetLastError());
}

template <typename V>
static void launch_f64(const SparseRows &S, const double *x, double *y, hipStream_t s, int L) {
    const V *val = reinterpret_cast<const V *>(S.val.p);
    const int64_t threads = S.rows * L;
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    if (blocks == 0) return;
#define TV3_CASE(LL)                                                                      \
    case LL:                                                                              \
        hipLaunchKernelGGL((k_csr_f64<LL, V>), dim3(blocks), dim3(256), 0, s, S.rows,     \
                           S.ptr.p, S.ind.p, val, x, y);                                  \
        break;
    switch (L) {
        TV3_CASE(2) TV3_CASE(4) TV3_CASE(8) TV3_CASE(16) TV3_CASE(32) TV3_CASE(64)
    default: throw tv3_error(TV3_E_ARG, "lanes per row must be a power of two in [2, 64]");
    }
#undef TV3_CASE
    TV3_HIP(hipGetLastError());
}

void build_row_blocks(tv3_ctx *ctx, SparseRows &S, int blk_nnz,
                      const std::vector<int64_t> *starts) {
    std::vector<int32_t> ptr(S.rows + 1);
    TV3_HIP(hipMemcpyAsync(ptr.data(), S.ptr.p, sizeof(int32_t) * (S.rows + 1),
                           hipMemcpyDeviceToHost, ctx->stream));
    TV3_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> blk;
    blk.reserve(S.nnz / (blk_nnz / 2) + S.rows / 256 + 2);
    int64_t start = 0;
    size_t si = 0;   // next forced block start
    blk.push_back(0);
    for (int64_t r = 0; r < S.rows; ++r) {
        while (starts && si < starts->size() && (*starts)[si] < r) ++si;
        const bool forced = starts && si < starts->size() && (*starts)[si] == r;
        // close the current block before row r if adding it would overflow
        if (r > start &&
            (forced || (int64_t)ptr[r + 1] - ptr[start] > blk_nnz || r - start >= 256)) {
            blk.push_back((int32_t)r);
            start = r;
        }
    }
    blk.push_back((int32_t)S.rows);
    S.nblk = (int64_t)blk.size() - 1;
    const std::vector<BlkDesc> d = make_blk_desc(blk, ptr.data(), nullptr);
    S.desc.alloc(d.size());
    TV3_HIP(hipMemcpyAsync(S.desc.p, d.data(), sizeof(BlkDesc) * d.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    TV3_HIP(hipStreamSynchronize(ctx->stream));
    S.blk_nnz = blk_nnz;
}

std::vector<BlkDesc> make_blk_desc(const std::vector<int32_t> &blk, const int32_t *ptr,
                                   const std::vector<int32_t> *pad_start) {
    const size_t nb = blk.empty() ? 0 : blk.size() - 1;
    std::vector<BlkDesc> d(nb + 1);
    for (size_t b = 0; b < nb; ++b) {
        BlkDesc &e = d[b];
        e.r0 = blk[b];
        e.nr = blk[b + 1] - blk[b];
        // kernels serve row r0 + min(lane group, nr - 1): a block needs a row
        TV3_REQUIRE(e.nr > 0, TV3_E_FORMAT, "row block without rows");
        e.p0 = ptr[blk[b]];
        e.len = ptr[blk[b + 1]] - e.p0;
        e.q0 = pad_start ? (*pad_start)[b] : e.p0;
        int lg = 6;
        while (lg > 0 && (1 << lg) * e.nr > 256) --lg;
        e.lg = lg;
        e.pad0 = e.pad1 = 0;
    }
    BlkDesc &z = d[nb];
    z.r0 = nb ? blk[nb] : 0;
    z.nr = z.len = z.lg = z.pad0 = z.pad1 = 0;
    z.p0 = z.q0 = nb ? ptr[blk[nb]] : 0;
    if (pad_start && nb) z.q0 = (*pad_start)[nb];
    return d;
}

void spmv_csr(const SparseRows &S, const float *x, float *y, hipStream_t s, int lanes,
              bool acc64) {
    if (!lanes && S.blk_nnz) {
        if (acc64) launch_stream_any<double>(S, x, EpiStore<double>{y}, s);
        else launch_stream_any<float>(S, x, EpiStore<float>{y}, s);
        TV3_HIP(hipGetLastError());
        return;
    }
    int L = lanes ? lanes : (S.lanes ? S.lanes : default_lanes(S.rows, S.nnz));
    switch (S.vtype) {
    case TV3_F32:
        acc64 ? launch_f32<float, double>(S, x, y, s, L) : launch_f32<float, float>(S, x, y, s, L);
        break;
    case TV3_F16:
        acc64 ? launch_f32<__half, double>(S, x, y, s, L) : launch_f32<__half, float>(S, x, y, s, L);
        break;
    case TV3_BF16:
        acc64 ? launch_f32<__hip_bfloat16, double>(S, x, y, s, L)
              : launch_f32<__hip_bfloat16, float>(S, x, y, s, L);
        break;
    default: throw tv3_error(TV3_E_FORMAT, "FP32 products need F32, F16 or BF16 values");
    }
}

void spmv_csr(const SparseRows &S, const double *x, double *y, hipStream_t s, int lanes,
              bool) {
    int L = lanes ? lanes : (S.lanes ? S.lanes : default_lanes(S.rows, S.nnz));
    switch (S.vtype) {
    case TV3_F64: launch_f64<double>(S, x, y, s, L); break;
    case TV3_F32: launch_f64<float>(S, x, y, s, L); break;
    default: throw tv3_error(TV3_E_FORMAT, "FP64 products need F64 or F32 values");
    }
}

// ------------------------------------------------------------------ scan
// Three passes: per-block sums of 1024 elements, a single-block carried scan
// of the block sums, then the per-block exclusive scan with its offset.
__global__ __launch_bounds__(256) void k_block_sums(const int32_t *cnt, int64_t n,
                                                    int64_t *bsum) {
    __shared__ int64_t sh[256];
    int64_t base = (int64_t)blockIdx.x * 1024;
    int64_t v = 0;
    for (int j = 0; j < 4; ++j) {
        int64_t i = base + threadIdx.x * 4 + j;
        if (i < n) v += cnt[i];
    }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_scan_sums(int64_t *bsum, int64_t nb, int64_t *total) {
    // exclusive scan of nb values, 256 at a time with a running carry
    __shared__ int64_t sh[256];
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += 256) {
        int64_t i = base + threadIdx.x;
        int64_t v = i < nb ? bsum[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            int64_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + sh[threadIdx.x] - v;
        int64_t blocktot = sh[255];
        __syncthreads();
        carry += blocktot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(256) void k_block_scan(const int32_t *cnt, int64_t n,
                                                    const int64_t *boff, int32_t *out) {
    __shared__ int64_t sh[256];
    int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x * 4;
    int64_t v[4], tsum = 0;
    for (int j = 0; j < 4; ++j) {
        v[j] = (base + j < n) ? cnt[base + j] : 0;
        tsum += v[j];
    }
    sh[threadIdx.x] = tsum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int64_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int64_t run = boff[blockIdx.x] + sh[threadIdx.x] - tsum;
    for (int j = 0; j < 4; ++j) {
        if (base + j < n) out[base + j] = (int32_t)run;
        run += v[j];
    }
}

__global__ void k_set_last(int32_t *out, int64_t n, const int64_t *total) {
    out[n] = (int32_t)*total;
}

int64_t exclusive_scan_i32(tv3_ctx *ctx, const int32_t *cnt, int32_t *out, int64_t n) {
    hipStream_t s = ctx->stream;
    int64_t nb = (n + 1023) / 1024;
    DevBuf<int64_t> bsum(nb + 1);
    int64_t *total = bsum.p + nb;
    if (nb > 0) {
        hipLaunchKernelGGL(k_block_sums, dim3((unsigned)nb), dim3(256), 0, s, cnt, n, bsum.p);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, bsum.p, nb, total);
        hipLaunchKernelGGL(k_block_scan, dim3((unsigned)nb), dim3(256), 0, s, cnt, n, bsum.p,
                           out);
    } else {
        TV3_HIP(hipMemsetAsync(total, 0, sizeof(int64_t), s));
    }
    hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, s, out, n, total);
    TV3_HIP(hipGetLastError());
    int64_t h = 0;
    TV3_HIP(hipMemcpyAsync(&h, total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    TV3_HIP(hipStreamSynchronize(s));
    TV3_REQUIRE(h <= INT32_MAX, TV3_E_SHAPE, "more than 2^31-1 entries: int32 indices overflow");
    return h;
}

// ------------------------------------------------------------------ transpose
__global__ void k_col_count(const int32_t *ind, int64_t nnz, int32_t *cnt) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) atomicAdd(&cnt[ind[k]], 1);
}

template <int VB>
__global__ void k_scatter(int64_t rows, const int32_t *ptr, const int32_t *ind,
                          const uint8_t *val, int32_t *cursor, int32_t *orow, uint8_t *oval) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
        int pos = atomicAdd(&cursor[ind[k]], 1);
        orow[pos] = (int32_t)i;
#pragma unroll
        for (int b = 0; b < VB; ++b) oval[(int64_t)pos * VB + b] = val[(int64_t)k * VB + b];
    }
}

// Sort each column's entries by row: makes the transpose independent of the
// atomic arrival order.  Entries are packed as (row << 32 | value bits)
// keys.  Columns of up to 64 entries: one thread, insertion sort in private
// memory; longer columns (rays start inside the volume, so a transducer's
// voxel can hold thousands of entries): one workgroup each, bitonic sort in
// LDS.
template <int VB>
__device__ __forceinline__ uint64_t pack_key(const int32_t *row, const uint8_t *val, int64_t k) {
    uint32_t v = 0;
#pragma unroll
    for (int t = 0; t < VB; ++t) v |= (uint32_t)val[k * VB + t] << (8 * t);
    return ((uint64_t)(uint32_t)row[k] << 32) | v;
}
template <int VB>
__device__ __forceinline__ void unpack_key(uint64_t key, int32_t *row, uint8_t *val, int64_t k) {
    row[k] = (int32_t)(key >> 32);
    const uint32_t v = (uint32_t)key;
#pragma unroll
    for (int t = 0; t < VB; ++t) val[k * VB + t] = (uint8_t)(v >> (8 * t));
}

template <int VB>
__global__ void k_sort_cols(int64_t cols, const int32_t *cptr, int32_t *row, uint8_t *val,
                            int32_t *long_list, int32_t *long_count) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    const int b = cptr[j], e = cptr[j + 1], len = e - b;
    if (len < 2) return;
    if (len > 64) {
        long_list[atomicAdd(long_count, 1)] = (int32_t)j;
        return;
    }
    uint64_t key[64];
    for (int a = 0; a < len; ++a) key[a] = pack_key<VB>(row, val, b + a);
    for (int a = 1; a < len; ++a) {
        const uint64_t k = key[a];
        int c = a - 1;
        while (c >= 0 && key[c] > k) { key[c + 1] = key[c]; --c; }
        key[c + 1] = k;
    }
    for (int a = 0; a < len; ++a) unpack_key<VB>(key[a], row, val, b + a);
}

// 8-byte values (FP64 operators): in-place insertion sort by row.
__global__ void k_sort_cols_f64(int64_t cols, const int32_t *cptr, int32_t *row, double *val) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    const int b = cptr[j], e = cptr[j + 1];
    for (int a = b + 1; a < e; ++a) {
        const int32_t r = row[a];
        const double v = val[a];
        int c = a - 1;
        while (c >= b && row[c] > r) { row[c + 1] = row[c]; val[c + 1] = val[c]; --c; }
        row[c + 1] = r;
        val[c + 1] = v;
    }
}

constexpr int kLongSortCap = 8192;   // LDS keys per workgroup (64 KiB)

template <int VB>
__global__ __launch_bounds__(256) void k_sort_long(const int32_t *list, const int32_t *cptr,
                                                   int32_t *row, uint8_t *val) {
    __shared__ uint64_t key[kLongSortCap];
    const int j = list[blockIdx.x];
    const int b = cptr[j], len = cptr[j + 1] - b;
    int P = 1;
    while (P < len) P <<= 1;
    for (int a = threadIdx.x; a < P; a += 256)
        key[a] = a < len ? pack_key<VB>(row, val, b + a) : ~0ull;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int i = threadIdx.x; i < P; i += 256) {
                const int l = i ^ jj;
                if (l > i) {
                    const uint64_t x = key[i], y = key[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) { key[i] = y; key[l] = x; }
                }
            }
            __syncthreads();
        }
    }
    for (int a = threadIdx.x; a < len; a += 256) unpack_key<VB>(key[a], row, val, b + a);
}

void build_transpose(tv3_ctx *ctx, const SparseRows &A, SparseRows &AT) {
    hipStream_t s = ctx->stream;
    AT.rows = A.cols;
    AT.cols = A.rows;
    AT.nnz = A.nnz;
    AT.vtype = A.vtype;
    AT.ptr.alloc(AT.rows + 1);
    AT.ind.alloc(A.nnz > 0 ? A.nnz : 1);
    const size_t vb = dtype_size(A.vtype);
    AT.val.alloc((A.nnz > 0 ? A.nnz : 1) * vb);
    DevBuf<int32_t> cnt(AT.rows + 1);
    TV3_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
    if (A.nnz > 0)
        hipLaunchKernelGGL(k_col_count, dim3((unsigned)((A.nnz + 255) / 256)), dim3(256), 0, s,
                           A.ind.p, A.nnz, cnt.p);
    exclusive_scan_i32(ctx, cnt.p, AT.ptr.p, AT.rows);
    TV3_HIP(hipMemcpyAsync(cnt.p, AT.ptr.p, sizeof(int32_t) * AT.rows, hipMemcpyDeviceToDevice, s));
    unsigned gb = (unsigned)((A.rows + 255) / 256), gc = (unsigned)((AT.rows + 255) / 256);
    if (A.nnz > 0 && vb == 8) {
        hipLaunchKernelGGL(k_scatter<8>, dim3(gb), dim3(256), 0, s, A.rows, A.ptr.p, A.ind.p,
                           A.val.p, cnt.p, AT.ind.p, AT.val.p);
        hipLaunchKernelGGL(k_sort_cols_f64, dim3(gc), dim3(256), 0, s, AT.rows, AT.ptr.p, AT.ind.p,
                           reinterpret_cast<double *>(AT.val.p));
    } else if (A.nnz > 0) {
        DevBuf<int32_t> list(AT.rows > 0 ? AT.rows : 1), lcount(1);
        TV3_HIP(hipMemsetAsync(lcount.p, 0, sizeof(int32_t), s));
        if (vb == 2) {
            hipLaunchKernelGGL(k_scatter<2>, dim3(gb), dim3(256), 0, s, A.rows, A.ptr.p, A.ind.p,
                               A.val.p, cnt.p, AT.ind.p, AT.val.p);
            hipLaunchKernelGGL(k_sort_cols<2>, dim3(gc), dim3(256), 0, s, AT.rows, AT.ptr.p,
                               AT.ind.p, AT.val.p, list.p, lcount.p);
        } else {
            hipLaunchKernelGGL(k_scatter<4>, dim3(gb), dim3(256), 0, s, A.rows, A.ptr.p, A.ind.p,
                               A.val.p, cnt.p, AT.ind.p, AT.val.p);
            hipLaunchKernelGGL(k_sort_cols<4>, dim3(gc), dim3(256), 0, s, AT.rows, AT.ptr.p,
                               AT.ind.p, AT.val.p, list.p, lcount.p);
        }
        int32_t nlong = 0;
        TV3_HIP(hipMemcpyAsync(&nlong, lcount.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        TV3_HIP(hipStreamSynchronize(s));
        if (nlong > 0) {
            // every long column must fit the LDS sort
            std::vector<int32_t> lst(nlong), cp(AT.rows + 1);
            TV3_HIP(hipMemcpy(lst.data(), list.p, sizeof(int32_t) * nlong, hipMemcpyDeviceToHost));
            TV3_HIP(hipMemcpy(cp.data(), AT.ptr.p, sizeof(int32_t) * (AT.rows + 1),
                              hipMemcpyDeviceToHost));
            for (int32_t j : lst)
                TV3_REQUIRE(cp[j + 1] - cp[j] <= kLongSortCap, TV3_E_SHAPE,
                            "a column holds more than 8192 entries (device transpose limit)");
            if (vb == 2)
                hipLaunchKernelGGL(k_sort_long<2>, dim3(nlong), dim3(256), 0, s, list.p, AT.ptr.p,
                                   AT.ind.p, AT.val.p);
            else
                hipLaunchKernelGGL(k_sort_long<4>, dim3(nlong), dim3(256), 0, s, list.p, AT.ptr.p,
                                   AT.ind.p, AT.val.p);
        }
    }
    TV3_HIP(hipGetLastError());
    TV3_HIP(hipStreamSynchronize(s));
    AT.lanes = default_lanes(AT.rows, AT.nnz);
    build_row_blocks(ctx, AT, 2048);
}

}  // namespace tv3
