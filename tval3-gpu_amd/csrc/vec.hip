// vec.hip -- kernels of the reference-shaped ("plain") TVAL3 loop: one kernel
// per vector operation, reductions as FP64 block partials summed in a fixed
// order.  Float and double instantiations.
//
// Reference kernels restated (3d/tval3_gpu_3d_kernels.cu): vec_add :75-108,
// D :170-215, D^T :479-510, lam2/lam4 :724-771 (all three components, sigma
// in X,Y,Z order), lam3/lam5 :876-919, BB step :1017-1067, nonneg
// :1176-1188, shrinkage :1249-1300 (plus the isotropic variant).
#include "vec.hpp"

namespace tv3 {

// ---------------------------------------------------------------- elementwise
template <typename T>
__global__ void k_axpy3(int64_t n, T *z, T a, const T *x, const T *y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        z[i] = a * x[i] + y[i];
}

template <typename T>
__global__ void k_grad(int64_t n, T *d, T a, const T *g, const T *g2, const T *dts) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        T v = a * g[i] + g2[i];
        d[i] = v + dts[i];
    }
}

template <typename T>
__global__ void k_nonneg(int64_t n, T *u) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        T v = u[i];
        u[i] = v < 0 ? (T)0 : v;
    }
}

template <typename T>
__global__ void k_scale(int64_t n, T *y, T a, const T *x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = a * x[i];
}

template <typename T>
__global__ void k_div(int64_t n, T *y, T a, const T *x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = x[i] / a;
}

static inline unsigned ew_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b > 0 ? b : 1);
}

template <typename T>
void axpy3(int64_t n, T *z, double a, const T *x, const T *y, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_axpy3<T>, dim3(ew_blocks(n)), dim3(256), 0, s, n, z, (T)a, x, y);
}
template <typename T>
void grad(int64_t n, T *d, double a, const T *g, const T *g2, const T *dts, hipStream_t s) {
    hipLaunchKernelGGL(k_grad<T>, dim3(ew_blocks(n)), dim3(256), 0, s, n, d, (T)a, g, g2, dts);
}
template <typename T>
void nonneg(int64_t n, T *u, hipStream_t s) {
    hipLaunchKernelGGL(k_nonneg<T>, dim3(ew_blocks(n)), dim3(256), 0, s, n, u);
}
template <typename T>
void divide(int64_t n, T *y, double a, const T *x, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_div<T>, dim3(ew_blocks(n)), dim3(256), 0, s, n, y, (T)a, x);
}
template <typename T>
void scale(int64_t n, T *y, double a, const T *x, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scale<T>, dim3(ew_blocks(n)), dim3(256), 0, s, n, y, (T)a, x);
}

// ---------------------------------------------------------------- stencils
// thread (tx, ty) of block (bx, by, bz) owns voxel y = bx*64+tx, x = by*4+ty, z = bz
template <typename T>
__global__ __launch_bounds__(256) void k_D(int p, int q, int r, const T *__restrict__ U,
                                           T *__restrict__ Ux, T *__restrict__ Uy,
                                           T *__restrict__ Uz) {
    int y = blockIdx.x * 64 + threadIdx.x, x = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (y >= p || x >= q) return;
    int64_t pq = (int64_t)p * q;
    int64_t i = (int64_t)z * pq + (int64_t)x * p + y;
    T t = U[i];
    Uy[i] = (y < p - 1 ? U[i + 1] : U[i - y]) - t;
    Ux[i] = (x < q - 1 ? U[i + p] : U[i - (int64_t)x * p]) - t;
    Uz[i] = (z < r - 1 ? U[i + pq] : U[i - (int64_t)z * pq]) - t;
}

template <typename T>
__global__ __launch_bounds__(256) void k_Dt(int p, int q, int r, const T *__restrict__ X,
                                            const T *__restrict__ Y, const T *__restrict__ Z,
                                            T *__restrict__ res) {
    int y = blockIdx.x * 64 + threadIdx.x, x = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (y >= p || x >= q) return;
    int64_t pq = (int64_t)p * q;
    int64_t i = (int64_t)z * pq + (int64_t)x * p + y;
    T yp = (y == 0) ? Y[i + p - 1] : Y[i - 1];
    T xp = (x == 0) ? X[i + (int64_t)(q - 1) * p] : X[i - p];
    T zp = (z == 0) ? Z[i + (int64_t)(r - 1) * pq] : Z[i - pq];
    res[i] = yp - Y[i] + xp - X[i] + zp - Z[i];
}

static inline dim3 st_grid(const int32_t dims[3]) {
    return dim3((unsigned)((dims[0] + 63) / 64), (unsigned)((dims[1] + 3) / 4), (unsigned)dims[2]);
}

template <typename T>
void launch_D(const int32_t dims[3], const T *U, T *Ux, T *Uy, T *Uz, hipStream_t s) {
    hipLaunchKernelGGL(k_D<T>, st_grid(dims), dim3(64, 4), 0, s, dims[0], dims[1], dims[2], U, Ux,
                       Uy, Uz);
}
template <typename T>
void launch_Dt(const int32_t dims[3], const T *X, const T *Y, const T *Z, T *res, hipStream_t s) {
    hipLaunchKernelGGL(k_Dt<T>, st_grid(dims), dim3(64, 4), 0, s, dims[0], dims[1], dims[2], X, Y,
                       Z, res);
}

// ---------------------------------------------------------------- reductions
template <typename T> struct FDot {
    const T *a, *b;
    __device__ void operator()(int64_t i, double *acc) const { acc[0] += (double)a[i] * (double)b[i]; }
};

template <typename T> struct FLam24 {
    const T *Ux, *Uy, *Uz, *Wx, *Wy, *Wz, *sx, *sy, *sz;
    T *Vx, *Vy, *Vz;
    __device__ void operator()(int64_t i, double *acc) const {
        T vx = Ux[i] - Wx[i], vy = Uy[i] - Wy[i], vz = Uz[i] - Wz[i];
        Vx[i] = vx; Vy[i] = vy; Vz[i] = vz;
        acc[0] += (double)vx * vx + (double)vy * vy + (double)vz * vz;
        acc[1] += (double)vx * sx[i] + (double)vy * sy[i] + (double)vz * sz[i];
    }
};

template <typename T> struct FLam35 {
    const T *Au, *b, *delta;
    __device__ void operator()(int64_t i, double *acc) const {
        T r = Au[i] - b[i];
        acc[0] += (double)r * r;
        acc[1] += (double)delta[i] * r;
    }
};

template <typename T> struct FTau {
    const T *g, *gp, *g2, *g2p, *uup;
    T mdb;
    T *dg, *dg2;
    __device__ void operator()(int64_t i, double *acc) const {
        T a = g[i] - gp[i], b = g2[i] - g2p[i], u = uup[i];
        acc[0] += (double)u * u;
        acc[1] += (double)u * (double)(b + mdb * a);
        dg[i] = a; dg2[i] = b;
    }
};

template <typename T> struct FShrink {
    const T *Ux, *Uy, *Uz, *sx, *sy, *sz;
    T beta;
    int iso;
    T *Wx, *Wy, *Wz;
    __device__ void operator()(int64_t i, double *acc) const {
        T ib = (T)1 / beta;
        T ux = Ux[i] - sx[i] / beta, uy = Uy[i] - sy[i] / beta, uz = Uz[i] - sz[i] / beta;
        if (!iso) {
            T tx = fabs(ux) - ib, ty = fabs(uy) - ib, tz = fabs(uz) - ib;
            tx = tx >= 0 ? tx : (T)0; ty = ty >= 0 ? ty : (T)0; tz = tz >= 0 ? tz : (T)0;
            Wx[i] = ux >= 0 ? tx : -tx;
            Wy[i] = uy >= 0 ? ty : -ty;
            Wz[i] = uz >= 0 ? tz : -tz;
            acc[0] += (double)tx + (double)ty + (double)tz;
        } else {
            T nrm = (T)sqrt((double)ux * ux + (double)uy * uy + (double)uz * uz);
            T sh = nrm - ib;
            if (sh > 0) {
                T f = sh / nrm;
                Wx[i] = ux * f; Wy[i] = uy * f; Wz[i] = uz * f;
                acc[0] += (double)sh;
            } else {
                Wx[i] = 0; Wy[i] = 0; Wz[i] = 0;
            }
        }
    }
};

template <typename T> struct FMlpSigma {
    const T *Ux, *Uy, *Uz, *Wx, *Wy, *Wz;
    T nb;
    T *sx, *sy, *sz;
    __device__ void operator()(int64_t i, double *acc) const {
        T vx = Ux[i] - Wx[i], vy = Uy[i] - Wy[i], vz = Uz[i] - Wz[i];
        T a = nb * vx + sx[i], b = nb * vy + sy[i], c = nb * vz + sz[i];
        sx[i] = a; sy[i] = b; sz[i] = c;
        acc[0] += (double)a * vx + (double)b * vy + (double)c * vz;
    }
};

template <typename T> struct FMlpDelta {
    const T *Au, *b;
    T nm;
    T *delta;
    __device__ void operator()(int64_t i, double *acc) const {
        T aub = Au[i] - b[i];
        T d = nm * aub + delta[i];
        delta[i] = d;
        acc[0] += (double)d * aub;
    }
};

template <typename T> struct FDiffNorm {
    const T *a, *b;
    __device__ void operator()(int64_t i, double *acc) const {
        T t = a[i] - b[i];
        acc[0] += (double)t * t;
        acc[1] += (double)a[i] * a[i];
    }
};

template <int K, typename F>
__global__ __launch_bounds__(256) void k_reduce(int64_t n, F f, double *partials) {
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        f(i, acc);
    block_sum_store<K>(acc, partials, gridDim.x);
}

template <int K, typename F>
std::array<double, K> reduce_now(tv3_ctx *ctx, int64_t n, const F &f, hipStream_t s) {
    int64_t nb = (n + 255) / 256;
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    size_t need = (size_t)K * nb;
    if (ctx->red_partials.n < need) ctx->red_partials.alloc(need < 8192 ? 8192 : need);
    ctx->red_host.alloc(ctx->red_partials.n);
    hipLaunchKernelGGL((k_reduce<K, F>), dim3((unsigned)nb), dim3(256), 0, s, n, f,
                       ctx->red_partials.p);
    TV3_HIP(hipGetLastError());
    TV3_HIP(hipMemcpyAsync(ctx->red_host.p, ctx->red_partials.p, need * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    TV3_HIP(hipStreamSynchronize(s));
    std::array<double, K> out{};
    for (int k = 0; k < K; ++k) {
        double t = 0.0;
        for (int64_t b = 0; b < nb; ++b) t += ctx->red_host.p[k * nb + b];
        out[k] = t;
    }
    return out;
}

template <typename T>
double dot(tv3_ctx *ctx, int64_t n, const T *a, const T *b, hipStream_t s) {
    return reduce_now<1>(ctx, n, FDot<T>{a, b}, s)[0];
}
template <typename T>
std::array<double, 2> lam24(tv3_ctx *ctx, int64_t n, const T *Ux, const T *Uy, const T *Uz,
                            const T *Wx, const T *Wy, const T *Wz, const T *sx, const T *sy,
                            const T *sz, T *Vx, T *Vy, T *Vz, hipStream_t s) {
    return reduce_now<2>(ctx, n, FLam24<T>{Ux, Uy, Uz, Wx, Wy, Wz, sx, sy, sz, Vx, Vy, Vz}, s);
}
template <typename T>
std::array<double, 2> lam35(tv3_ctx *ctx, int64_t m, const T *Au, const T *b, const T *delta,
                            hipStream_t s) {
    return reduce_now<2>(ctx, m, FLam35<T>{Au, b, delta}, s);
}
template <typename T>
std::array<double, 2> tau_bb(tv3_ctx *ctx, int64_t n, const T *g, const T *gp, const T *g2,
                             const T *g2p, const T *uup, double mdb, T *dg, T *dg2,
                             hipStream_t s) {
    return reduce_now<2>(ctx, n, FTau<T>{g, gp, g2, g2p, uup, (T)mdb, dg, dg2}, s);
}
template <typename T>
double shrink_host(tv3_ctx *ctx, int64_t n, int iso, const T *Ux, const T *Uy, const T *Uz,
                   const T *sx, const T *sy, const T *sz, double beta, T *Wx, T *Wy, T *Wz,
                   hipStream_t s) {
    return reduce_now<1>(ctx, n, FShrink<T>{Ux, Uy, Uz, sx, sy, sz, (T)beta, iso, Wx, Wy, Wz},
                         s)[0];
}
template <typename T>
double mlp_sigma(tv3_ctx *ctx, int64_t n, const T *Ux, const T *Uy, const T *Uz, const T *Wx,
                 const T *Wy, const T *Wz, double beta, T *sx, T *sy, T *sz, hipStream_t s) {
    return reduce_now<1>(ctx, n, FMlpSigma<T>{Ux, Uy, Uz, Wx, Wy, Wz, (T)(-beta), sx, sy, sz}, s)[0];
}
template <typename T>
double mlp_delta(tv3_ctx *ctx, int64_t m, const T *Au, const T *b, double mu, T *delta,
                 hipStream_t s) {
    return reduce_now<1>(ctx, m, FMlpDelta<T>{Au, b, (T)(-mu), delta}, s)[0];
}
template <typename T>
std::array<double, 2> diff_norm(tv3_ctx *ctx, int64_t n, const T *a, const T *b, hipStream_t s) {
    return reduce_now<2>(ctx, n, FDiffNorm<T>{a, b}, s);
}

#define TV3_INST(T)                                                                              \
    template void axpy3<T>(int64_t, T *, double, const T *, const T *, hipStream_t);            \
    template void grad<T>(int64_t, T *, double, const T *, const T *, const T *, hipStream_t);  \
    template void nonneg<T>(int64_t, T *, hipStream_t);                                         \
    template void scale<T>(int64_t, T *, double, const T *, hipStream_t);                       \
    template void divide<T>(int64_t, T *, double, const T *, hipStream_t);                      \
    template void launch_D<T>(const int32_t *, const T *, T *, T *, T *, hipStream_t);          \
    template void launch_Dt<T>(const int32_t *, const T *, const T *, const T *, T *,           \
                               hipStream_t);                                                    \
    template double dot<T>(tv3_ctx *, int64_t, const T *, const T *, hipStream_t);              \
    template std::array<double, 2> lam24<T>(tv3_ctx *, int64_t, const T *, const T *,           \
                                            const T *, const T *, const T *, const T *,         \
                                            const T *, const T *, const T *, T *, T *, T *,     \
                                            hipStream_t);                                       \
    template std::array<double, 2> lam35<T>(tv3_ctx *, int64_t, const T *, const T *,           \
                                            const T *, hipStream_t);                            \
    template std::array<double, 2> tau_bb<T>(tv3_ctx *, int64_t, const T *, const T *,          \
                                             const T *, const T *, const T *, double, T *, T *, \
                                             hipStream_t);                                      \
    template double shrink_host<T>(tv3_ctx *, int64_t, int, const T *, const T *, const T *,    \
                                   const T *, const T *, const T *, double, T *, T *, T *,      \
                                   hipStream_t);                                                \
    template double mlp_sigma<T>(tv3_ctx *, int64_t, const T *, const T *, const T *,           \
                                 const T *, const T *, const T *, double, T *, T *, T *,        \
                                 hipStream_t);                                                  \
    template double mlp_delta<T>(tv3_ctx *, int64_t, const T *, const T *, double, T *,         \
                                 hipStream_t);                                                  \
    template std::array<double, 2> diff_norm<T>(tv3_ctx *, int64_t, const T *, const T *,       \
                                                hipStream_t);

TV3_INST(float)
TV3_INST(double)

}  // namespace tv3
