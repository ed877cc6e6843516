// dense_mfma.hip -- products of one dense BF16 matrix with up to 64
// right-hand sides on the matrix cores (v_mfma_f32_32x32x16_bf16).
//
// Y (R x G) = M (R x K) X (K x G), M row-major BF16 (row i at M + i ldm),
// X FP32 column-major (right-hand side j at X + j ldx), Y FP32
// column-major (Y + j ldy), optionally minus a vector of the same layout.
// This is the dense operator's product (3d/mat_vec_mul_dense.h:25-73, gemv N
// and T) for several vectors at once: A x uses a row-major copy of A, A^T r
// uses A's own column-major storage (which is A^T row-major), so both read
// 16 contiguous bytes per lane -- the dense counterpart of keeping a CSR and
// a CSC.
//
// Precision: A is BF16 by definition of the operator; each FP32 value of X
// is split into three BF16 terms, hi = bf16(x), mid = bf16(x - hi),
// lo = bf16(x - hi - mid), each multiplied into one FP32 accumulator (three
// MFMAs), so X enters with about 24 significant bits -- the products are as
// accurate as FP32 ones (two terms, 16 bits, leave a relative error near
// 2e-6 that TVAL3's step-size recursion amplifies).
//
// Work: a 512-thread workgroup owns 256 rows x 64 columns (wavefront w: rows
// 32w..32w+31, two 32 x 32 MFMA tiles) over one of S slices of K.  Per
// chunk of 64 K-values the workgroup converts X's 64 x 64 block to its three
// BF16 terms in LDS (read by all eight wavefronts); each lane loads its M operands
// (8 BF16 of one row, 16 bytes) straight from memory.  The S partial tiles
// go to a row-major scratch; a second kernel sums them in a fixed order,
// subtracts, and writes Y column-major through an LDS transpose.
//
// MFMA operand layout (v_mfma_f32_32x32x16_bf16, checked on the device):
// lane l supplies A[l % 32][8 (l / 32) + e] and B[8 (l / 32) + e][l % 32],
// e = 0..7; accumulator r of lane l is D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
#include "tv3_internal.hpp"

namespace tv3 {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kRows = 256;   // rows per workgroup (8 wavefronts)
constexpr int kThreads = 512;
constexpr int kCols = 64;    // right-hand sides per launch
constexpr int kChunk = 64;   // K-values per LDS chunk
constexpr int kPadK = kChunk + 8;

__device__ __forceinline__ uint16_t bf16_rne(float x) {
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__global__ __launch_bounds__(512) void k_gemm_rhs(const uint16_t *__restrict__ M, int64_t ldm,
                                                  int64_t R, int64_t K, int S,
                                                  const float *__restrict__ X, int64_t ldx,
                                                  int G, float *__restrict__ part,
                                                  const int32_t *gate) {
    if (gate && !*gate) return;
    __shared__ __attribute__((aligned(16))) uint16_t xh[kCols][kPadK];
    __shared__ __attribute__((aligned(16))) uint16_t xm[kCols][kPadK];
    __shared__ __attribute__((aligned(16))) uint16_t xl[kCols][kPadK];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t tiles = R / kRows;
    const int64_t tile = blockIdx.x % tiles;
    const int s = (int)(blockIdx.x / tiles);
    const int64_t kper = K / S;
    const int64_t k_beg = (int64_t)s * kper, k_end = k_beg + kper;
    const int64_t row = tile * kRows + wave * 32 + (lane & 31);
    const uint16_t *mrow = M + row * ldm + 8 * (lane >> 5);
    // X staging: thread t converts 8 values of column t / 8 (columns >= G
    // read column G - 1 and are zeroed: no branch around the loads)
    const int xc = t >> 3, xk = (t & 7) * 8;
    const float *xcol = X + (int64_t)(xc < G ? xc : G - 1) * ldx + xk;
    const float xon = xc < G ? 1.f : 0.f;
    f32x16 acc0 = {}, acc1 = {};
    for (int64_t kc = k_beg; kc < k_end; kc += kChunk) {
        bf16x8 a[4];
#pragma unroll
        for (int st = 0; st < 4; ++st)
            a[st] = *reinterpret_cast<const bf16x8 *>(mrow + kc + 16 * st);
        float4 x[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) x[q] = *reinterpret_cast<const float4 *>(xcol + kc + 4 * q);
        __syncthreads();   // the previous chunk's LDS operands have been read
        // the three BF16 terms of 8 values, packed two per word: 16-byte LDS stores
        uint32_t ph[4], pm[4], pl[4];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const float v[4] = {x[q].x * xon, x[q].y * xon, x[q].z * xon, x[q].w * xon};
#pragma unroll
            for (int e = 0; e < 4; e += 2) {
                uint32_t h2 = 0, m2 = 0, l2 = 0;
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const float w = v[e + u];
                    const uint32_t h = bf16_rne(w);
                    const float r1 = w - __uint_as_float(h << 16);
                    const uint32_t md = bf16_rne(r1);
                    const float r2 = r1 - __uint_as_float(md << 16);
                    const uint32_t lo = bf16_rne(r2);
                    h2 |= h << (16 * u); m2 |= md << (16 * u); l2 |= lo << (16 * u);
                }
                ph[2 * q + e / 2] = h2; pm[2 * q + e / 2] = m2; pl[2 * q + e / 2] = l2;
            }
        }
        *reinterpret_cast<uint4 *>(&xh[xc][xk]) = make_uint4(ph[0], ph[1], ph[2], ph[3]);
        *reinterpret_cast<uint4 *>(&xm[xc][xk]) = make_uint4(pm[0], pm[1], pm[2], pm[3]);
        *reinterpret_cast<uint4 *>(&xl[xc][xk]) = make_uint4(pl[0], pl[1], pl[2], pl[3]);
        __syncthreads();
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            const int kk = 16 * st + 8 * (lane >> 5);
            const int c0 = lane & 31, c1 = 32 + (lane & 31);
            // smallest terms first: the accumulator adds the large ones last
            const bf16x8 b0l = *reinterpret_cast<const bf16x8 *>(&xl[c0][kk]);
            const bf16x8 b1l = *reinterpret_cast<const bf16x8 *>(&xl[c1][kk]);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[st], b0l, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[st], b1l, acc1, 0, 0, 0);
            const bf16x8 b0m = *reinterpret_cast<const bf16x8 *>(&xm[c0][kk]);
            const bf16x8 b1m = *reinterpret_cast<const bf16x8 *>(&xm[c1][kk]);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[st], b0m, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[st], b1m, acc1, 0, 0, 0);
            const bf16x8 b0h = *reinterpret_cast<const bf16x8 *>(&xh[c0][kk]);
            const bf16x8 b1h = *reinterpret_cast<const bf16x8 *>(&xh[c1][kk]);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[st], b0h, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[st], b1h, acc1, 0, 0, 0);
        }
    }
    // partial tile, row-major [s][row][64]
    float *out = part + ((int64_t)s * R + tile * kRows + wave * 32) * kCols;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        out[(int64_t)i * kCols + (lane & 31)] = acc0[r];
        out[(int64_t)i * kCols + 32 + (lane & 31)] = acc1[r];
    }
}

// Y[j ldy + row] = sum_s part[s][row][j] (- sub[j ldy + row]), j < G, the
// S partials added in order s = 0, 1, ...; one workgroup per 32 rows (a
// thread sums 8 rows of one column: eight independent load chains), LDS
// transpose for column-major stores of 32 consecutive rows
__global__ __launch_bounds__(256) void k_gemm_reduce(const float *__restrict__ part, int64_t R,
                                                     int S, int G, float *__restrict__ Y,
                                                     int64_t ldy, const float *__restrict__ sub,
                                                     const int32_t *gate) {
    if (gate && !*gate) return;
    __shared__ float tile[kCols][33];
    const int t = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * 32;
    const int c = t & 63, rg = (t >> 6) * 8;
    float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 0; s < S; ++s) {
        const float *p = part + ((int64_t)s * R + r0 + rg) * kCols + c;
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] += p[(int64_t)i * kCols];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) tile[c][rg + i] = v[i];
    __syncthreads();
    const int rr = t & 31, cg = (t >> 5) * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int cc = cg + i;
        if (cc >= G) break;
        const int64_t o = (int64_t)cc * ldy + r0 + rr;
        const float w = tile[cc][rr];
        Y[o] = sub ? w - sub[o] : w;
    }
}

// row-major BF16 copy of a column-major m x n BF16 matrix (32 x 32 tiles)
__global__ __launch_bounds__(256) void k_transpose_bf16(const uint16_t *__restrict__ A,
                                                        int64_t lda, int64_t m, int64_t n,
                                                        uint16_t *__restrict__ T) {
    __shared__ uint16_t tl[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    const int64_t i0 = (int64_t)blockIdx.x * 32, j0 = (int64_t)blockIdx.y * 32;
    for (int k = ty; k < 32; k += 8) {
        const int64_t i = i0 + tx, j = j0 + k;
        tl[k][tx] = (i < m && j < n) ? A[j * lda + i] : (uint16_t)0;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int64_t i = i0 + k, j = j0 + tx;
        if (i < m && j < n) T[i * n + j] = tl[tx][k];
    }
}

}  // namespace

int gemm_rhs_rows() { return kRows; }

bool gemm_rhs_shape_ok(int64_t R, int64_t K, int splits) {
    return R > 0 && K > 0 && R % kRows == 0 && splits > 0 && K % ((int64_t)splits * kChunk) == 0;
}

void gemm_rhs(const uint16_t *M, int64_t ldm, int64_t R, int64_t K, const float *X, int64_t ldx,
              int G, float *Y, int64_t ldy, const float *sub, int splits, float *part,
              hipStream_t s, const int32_t *gate) {
    TV3_REQUIRE(G > 0 && G <= kCols, TV3_E_ARG, "1 to 64 right-hand sides per product");
    TV3_REQUIRE(gemm_rhs_shape_ok(R, K, splits), TV3_E_SHAPE,
                "dense multi-RHS products need rows a multiple of 256 and K a multiple of "
                "64 x splits");
    TV3_REQUIRE(ldm % 8 == 0 && ldx % 4 == 0, TV3_E_SHAPE,
                "dense multi-RHS products need 16-byte aligned rows");
    const unsigned grid = (unsigned)((R / kRows) * splits);
    hipLaunchKernelGGL(k_gemm_rhs, dim3(grid), dim3(kThreads), 0, s, M, ldm, R, K, splits, X,
                       ldx, G, part, gate);
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)(R / 32)), dim3(256), 0, s, part, R, splits,
                       G, Y, ldy, sub, gate);
}

size_t gemm_rhs_scratch(int64_t R, int splits) { return (size_t)splits * R * kCols; }

void dense_rowmajor(tv3_op *op, hipStream_t s) {
    if (op->dense_rm.p) return;
    TV3_REQUIRE(op->dtype == TV3_BF16, TV3_E_FORMAT,
                "dense multi-RHS products take BF16 matrices");
    op->dense_rm.alloc((size_t)op->m * op->n * 2);
    const dim3 grid((unsigned)((op->m + 31) / 32), (unsigned)((op->n + 31) / 32));
    hipLaunchKernelGGL(k_transpose_bf16, grid, dim3(256), 0, s,
                       (const uint16_t *)op->dense.p, op->lda, op->m, op->n,
                       (uint16_t *)op->dense_rm.p);
    TV3_HIP(hipGetLastError());
}

}  // namespace tv3
