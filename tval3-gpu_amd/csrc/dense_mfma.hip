// dense_mfma.hip -- products of one dense BF16 matrix with up to 64
// right-hand sides on the matrix cores (v_mfma_f32_32x32x16_bf16).
//
// Y (R x G) = M (R x K) X (K x G), M BF16 in MFMA operand order (below),
// X FP32 column-major (right-hand side j at X + j ldx), Y FP32
// column-major (Y + j ldy), optionally minus a vector of the same layout.
// This is the dense operator's product (3d/mat_vec_mul_dense.h:25-73, gemv N
// and T) for several vectors at once: A x and A^T r each read a copy of A
// (M = A, M = A^T) built once on the device -- the dense counterpart of
// keeping a CSR and a CSC.
//
// Operand order: M is stored in blocks of 32 rows x 16 columns (1 KB), the
// blocks of a 32-row band consecutive along K; inside a block lane l's 8
// operands M[32 rb + l % 32][16 kb + 8 (l / 32) + e], e = 0..7, lie at
// (l * 8 + e).  One wavefront load of its A operands is then 1 KB of
// contiguous memory (eight whole cache lines) instead of 16 bytes from each
// of 32 rows K apart.
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
// (16 contiguous bytes of the operand-ordered copy) straight from memory,
// one chunk ahead.  The S partial tiles
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

__global__ __launch_bounds__(512) void k_gemm_rhs(const uint16_t *__restrict__ M,
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
    // this wavefront's 32-row band of M in operand order, lane's 16 bytes
    const uint16_t *mrow = M + (tile * (kRows / 32) + wave) * (K / 16) * 512 + lane * 8;
    // X staging: thread t converts 8 values of column t / 8 (columns >= G
    // read column G - 1 and are zeroed: no branch around the loads)
    const int xc = t >> 3, xk = (t & 7) * 8;
    const float *xcol = X + (int64_t)(xc < G ? xc : G - 1) * ldx + xk;
    const float xon = xc < G ? 1.f : 0.f;
    f32x16 acc0 = {}, acc1 = {};
    // Chunk c's operands are requested one chunk ahead: while chunk c is
    // converted and multiplied, chunk c + 1's M and X loads are in flight.
    // Two register sets, unrolled by two, so no copy ties the wait for the
    // next chunk's loads to the current chunk's work.
    const int64_t nchunk = (k_end - k_beg) / kChunk;
    auto load_a = [&](bf16x8 *a, int64_t c) {
        const int64_t kc = k_beg + min(c, nchunk - 1) * kChunk;   // clamped: in range
#pragma unroll
        for (int st = 0; st < 4; ++st)
            a[st] = *reinterpret_cast<const bf16x8 *>(mrow + (kc / 16 + st) * 512);
    };
    auto load_x = [&](float4 *x, int64_t c) {
        const int64_t kc = k_beg + min(c, nchunk - 1) * kChunk;
#pragma unroll
        for (int q = 0; q < 2; ++q) x[q] = *reinterpret_cast<const float4 *>(xcol + kc + 4 * q);
    };
    auto chunk = [&](const float4 *x, const bf16x8 *a) {
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
    };
    bf16x8 a0[4], a1[4];
    float4 x0[2], x1[2];
    load_x(x0, 0);
    load_a(a0, 0);
    // (no exit between the two halves, so the compiler does not sink the
    // next chunk's loads below the current chunk's work; an odd last chunk
    // runs after the loop)
    int64_t c = 0;
    for (; c + 1 < nchunk; c += 2) {
        load_x(x1, c + 1);
        load_a(a1, c + 1);
        chunk(x0, a0);
        load_x(x0, c + 2);
        load_a(a0, c + 2);
        chunk(x1, a1);
    }
    if (c < nchunk) chunk(x0, a0);
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

// Operand-ordered copy of M = A (trans = 0, R = m, K = n) or M = A^T
// (trans = 1, R = n, K = m) from a column-major m x n BF16 matrix A (leading
// dimension lda): thread (block, lane) writes lane's 8 operands of one
// 32 x 16 block (16 bytes)
__global__ __launch_bounds__(256) void k_operand_order_bf16(const uint16_t *__restrict__ A,
                                                            int64_t lda, int trans, int64_t R,
                                                            int64_t K, uint16_t *__restrict__ Mt) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t nkb = K / 16;
    if (g >= (R / 32) * nkb * 64) return;
    const int lane = (int)(g & 63);
    const int64_t blk = g >> 6, rb = blk / nkb, kb = blk % nkb;
    const int64_t row = 32 * rb + (lane & 31), col = 16 * kb + 8 * (lane >> 5);
    uint16_t v[8];
    if (trans) {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = A[row * lda + col + e];   // column row of A
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = A[(col + e) * lda + row];
    }
    uint4 w;
    w.x = v[0] | ((uint32_t)v[1] << 16); w.y = v[2] | ((uint32_t)v[3] << 16);
    w.z = v[4] | ((uint32_t)v[5] << 16); w.w = v[6] | ((uint32_t)v[7] << 16);
    *reinterpret_cast<uint4 *>(Mt + g * 8) = w;
}

}  // namespace

int gemm_rhs_rows() { return kRows; }

bool gemm_rhs_shape_ok(int64_t R, int64_t K, int splits) {
    return R > 0 && K > 0 && R % kRows == 0 && splits > 0 && K % ((int64_t)splits * kChunk) == 0;
}

void gemm_rhs(const uint16_t *M, int64_t R, int64_t K, const float *X, int64_t ldx, int G,
              float *Y, int64_t ldy, const float *sub, int splits, float *part, hipStream_t s,
              const int32_t *gate) {
    TV3_REQUIRE(G > 0 && G <= kCols, TV3_E_ARG, "1 to 64 right-hand sides per product");
    TV3_REQUIRE(gemm_rhs_shape_ok(R, K, splits), TV3_E_SHAPE,
                "dense multi-RHS products need rows a multiple of 256 and K a multiple of "
                "64 x splits");
    TV3_REQUIRE(ldx % 4 == 0, TV3_E_SHAPE,
                "dense multi-RHS products need 16-byte aligned right-hand sides");
    const unsigned grid = (unsigned)((R / kRows) * splits);
    hipLaunchKernelGGL(k_gemm_rhs, dim3(grid), dim3(kThreads), 0, s, M, R, K, splits, X, ldx, G,
                       part, gate);
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)(R / 32)), dim3(256), 0, s, part, R, splits,
                       G, Y, ldy, sub, gate);
}

size_t gemm_rhs_scratch(int64_t R, int splits) { return (size_t)splits * R * kCols; }

void dense_operand_order(tv3_op *op, hipStream_t s) {
    if (op->dense_oN.p) return;
    TV3_REQUIRE(op->dtype == TV3_BF16, TV3_E_FORMAT,
                "dense multi-RHS products take BF16 matrices");
    TV3_REQUIRE(op->m % 32 == 0 && op->n % 32 == 0, TV3_E_SHAPE,
                "dense multi-RHS products need rows and columns multiples of 32");
    const size_t bytes = (size_t)op->m * op->n * 2;
    DevBuf<uint8_t> oN(bytes), oT(bytes);
    const int64_t threads = op->m * op->n / 8;
    const unsigned grid = (unsigned)((threads + 255) / 256);
    hipLaunchKernelGGL(k_operand_order_bf16, dim3(grid), dim3(256), 0, s,
                       (const uint16_t *)op->dense.p, op->lda, 0, op->m, op->n,
                       (uint16_t *)oN.p);
    hipLaunchKernelGGL(k_operand_order_bf16, dim3(grid), dim3(256), 0, s,
                       (const uint16_t *)op->dense.p, op->lda, 1, op->n, op->m,
                       (uint16_t *)oT.p);
    TV3_HIP(hipGetLastError());
    op->dense_oN = std::move(oN);
    op->dense_oT = std::move(oT);
}

}  // namespace tv3
