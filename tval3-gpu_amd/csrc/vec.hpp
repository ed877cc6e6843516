// vec.hpp -- declarations of the plain-path vector kernels and the shared
// block-reduction helper.
#pragma once

#include <array>

#include "lanes.hpp"
#include "tv3_internal.hpp"

namespace tv3 {

// Sum K per-thread doubles over a workgroup of NW wavefronts (4: 256
// threads, 8: 512) -- wavefront sums (wave_sum), then the wavefront totals
// in a fixed order; thread 0 stores partials[k*nblk + slot].  Every thread of
// the workgroup calls it.
template <int K, int NW = 4>
__device__ __forceinline__ void block_sum_store_at(double (&acc)[K], double *partials, int nblk,
                                                   int slot) {
    __shared__ double sh[K][NW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double v = wave_sum(acc[k]);
        if (lane == 0) sh[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double v = (sh[k][0] + sh[k][1]) + (sh[k][2] + sh[k][3]);
            if (NW == 8) v += (sh[k][4] + sh[k][5]) + (sh[k][6] + sh[k][7]);
            partials[(int64_t)k * nblk + slot] = v;
        }
    }
}

// The same at the workgroup's own slot (blockIdx.x).
template <int K, int NW = 4>
__device__ __forceinline__ void block_sum_store(double (&acc)[K], double *partials, int nblk) {
    block_sum_store_at<K, NW>(acc, partials, nblk, (int)blockIdx.x);
}

template <typename T> void axpy3(int64_t n, T *z, double a, const T *x, const T *y, hipStream_t s);
template <typename T>
void grad(int64_t n, T *d, double a, const T *g, const T *g2, const T *dts, hipStream_t s);
template <typename T> void nonneg(int64_t n, T *u, hipStream_t s);
template <typename T> void scale(int64_t n, T *y, double a, const T *x, hipStream_t s);
template <typename T> void divide(int64_t n, T *y, double a, const T *x, hipStream_t s);
template <typename T> double dot(tv3_ctx *ctx, int64_t n, const T *a, const T *b, hipStream_t s);
template <typename T>
std::array<double, 2> lam24(tv3_ctx *ctx, int64_t n, const T *Ux, const T *Uy, const T *Uz,
                            const T *Wx, const T *Wy, const T *Wz, const T *sx, const T *sy,
                            const T *sz, T *Vx, T *Vy, T *Vz, hipStream_t s);
template <typename T>
std::array<double, 2> lam35(tv3_ctx *ctx, int64_t m, const T *Au, const T *b, const T *delta,
                            hipStream_t s);
template <typename T>
std::array<double, 2> tau_bb(tv3_ctx *ctx, int64_t n, const T *g, const T *gp, const T *g2,
                             const T *g2p, const T *uup, double mdb, T *dg, T *dg2, hipStream_t s);
template <typename T>
double mlp_sigma(tv3_ctx *ctx, int64_t n, const T *Ux, const T *Uy, const T *Uz, const T *Wx,
                 const T *Wy, const T *Wz, double beta, T *sx, T *sy, T *sz, hipStream_t s);
template <typename T>
double mlp_delta(tv3_ctx *ctx, int64_t m, const T *Au, const T *b, double mu, T *delta,
                 hipStream_t s);
template <typename T>
std::array<double, 2> diff_norm(tv3_ctx *ctx, int64_t n, const T *a, const T *b, hipStream_t s);

}  // namespace tv3
