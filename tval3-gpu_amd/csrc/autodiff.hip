// autodiff.hip -- the shrinkage map as a function of its input, with its
// vector-Jacobian product, for the differentiable building blocks
// (torch.autograd.Function bindings).  Float and double.
//
// shrinkage_gpu (3d/tval3_gpu_3d.cpp:924-960, kernel
// 3d/tval3_gpu_3d_kernels.cu:1249-1300) computes W = shrink(DU - sigma/beta)
// at threshold 1/beta; here the argument V = DU - sigma/beta and the threshold
// t are inputs:
//   aniso  W_c = sign(V_c) max(|V_c| - t, 0),      s = sum_c max(|V_c| - t, 0)
//   iso    W = max(r - t, 0) V / r,  r = |V|_2,     s = max(r - t, 0)
// s is the per-voxel term of lam1 (lam1 = sum s).  Both maps are Lipschitz
// and differentiable except on the threshold surface (|V_c| = t, r = t);
// the VJP takes the inactive side there:
//   aniso  gV_c = [|V_c| > t] (gW_c + gs sign(V_c))
//   iso    gV = [r > t] ((r-t)/r gW + t (V.gW)/r^3 V + gs V/r)
// The forward arithmetic follows the plain solver's shrinkage kernel operation
// for operation, so W matches it bit for bit in either precision.
//
// Active sets: mask_out (optional) receives per voxel the active components
// (aniso: bit c for |V_c| > t; iso: bit 0 for r > t).  mask_in (optional)
// replaces the comparison by a recorded active set, which evaluates the
// smooth branch the record selected (active: W = V (r-t)/r, s = r-t, or
// W_c = V_c - sign(V_c) t; inactive: 0) -- the function whose derivative the
// VJP is, used to check gradients by finite differences across steps that
// would otherwise cross the threshold.
#include "tv3_internal.hpp"

namespace tv3 {

namespace {

template <typename T>
__global__ __launch_bounds__(256) void k_shrink_apply(int64_t n, int iso, T t,
                                                      const T *__restrict__ Vx,
                                                      const T *__restrict__ Vy,
                                                      const T *__restrict__ Vz,
                                                      T *__restrict__ Wx, T *__restrict__ Wy,
                                                      T *__restrict__ Wz, T *__restrict__ s,
                                                      const uint8_t *__restrict__ mask_in,
                                                      uint8_t *__restrict__ mask_out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        T ux = Vx[i], uy = Vy[i], uz = Vz[i];
        T wx, wy, wz, si;
        unsigned act;
        if (!iso) {
            T tx = (T)fabs((double)ux) - t, ty = (T)fabs((double)uy) - t,
              tz = (T)fabs((double)uz) - t;
            act = (tx > 0 ? 1u : 0u) | (ty > 0 ? 2u : 0u) | (tz > 0 ? 4u : 0u);
            if (mask_in) {
                act = mask_in[i];
                tx = (act & 1u) ? tx : (T)0; ty = (act & 2u) ? ty : (T)0;
                tz = (act & 4u) ? tz : (T)0;
            } else {
                tx = tx >= 0 ? tx : (T)0; ty = ty >= 0 ? ty : (T)0; tz = tz >= 0 ? tz : (T)0;
            }
            wx = ux >= 0 ? tx : -tx;
            wy = uy >= 0 ? ty : -ty;
            wz = uz >= 0 ? tz : -tz;
            si = (T)((double)tx + (double)ty + (double)tz);
        } else {
            T nrm = (T)sqrt((double)ux * ux + (double)uy * uy + (double)uz * uz);
            T sh = nrm - t;
            act = sh > 0 ? 1u : 0u;
            if (mask_in) act = mask_in[i] & 1u;
            if (act) {
                T f = sh / nrm;
                wx = ux * f; wy = uy * f; wz = uz * f;
                si = sh;
            } else {
                wx = wy = wz = si = 0;
            }
        }
        Wx[i] = wx; Wy[i] = wy; Wz[i] = wz;
        if (s) s[i] = si;
        if (mask_out) mask_out[i] = (uint8_t)act;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_shrink_adjoint(
    int64_t n, int iso, T t, const T *__restrict__ Vx, const T *__restrict__ Vy,
    const T *__restrict__ Vz, const T *__restrict__ gWx, const T *__restrict__ gWy,
    const T *__restrict__ gWz, const T *__restrict__ gs, T *__restrict__ gVx,
    T *__restrict__ gVy, T *__restrict__ gVz) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double vx = Vx[i], vy = Vy[i], vz = Vz[i];
        double gx = gWx[i], gy = gWy[i], gz = gWz[i];
        double g0 = gs ? (double)gs[i] : 0.0;
        double ox = 0, oy = 0, oz = 0;
        if (!iso) {
            double td = t;
            if (fabs(vx) > td) ox = gx + (vx >= 0 ? g0 : -g0);
            if (fabs(vy) > td) oy = gy + (vy >= 0 ? g0 : -g0);
            if (fabs(vz) > td) oz = gz + (vz >= 0 ? g0 : -g0);
        } else {
            double r = sqrt(vx * vx + vy * vy + vz * vz);
            if (r > (double)t) {
                double ir = 1.0 / r;
                double a = (r - (double)t) * ir;
                double c = ((double)t * (vx * gx + vy * gy + vz * gz) * ir * ir + g0) * ir;
                ox = a * gx + c * vx;
                oy = a * gy + c * vy;
                oz = a * gz + c * vz;
            }
        }
        gVx[i] = (T)ox; gVy[i] = (T)oy; gVz[i] = (T)oz;
    }
}

inline unsigned blocks_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

}  // namespace

template <typename T>
void shrink_apply(int64_t n, int iso, double t, const T *Vx, const T *Vy, const T *Vz, T *Wx,
                  T *Wy, T *Wz, T *s, const uint8_t *mask_in, uint8_t *mask_out,
                  hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_shrink_apply<T>, dim3(blocks_for(n)), dim3(256), 0, st, n, iso, (T)t, Vx,
                       Vy, Vz, Wx, Wy, Wz, s, mask_in, mask_out);
}

template <typename T>
void shrink_adjoint(int64_t n, int iso, double t, const T *Vx, const T *Vy, const T *Vz,
                    const T *gWx, const T *gWy, const T *gWz, const T *gs, T *gVx, T *gVy,
                    T *gVz, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_shrink_adjoint<T>, dim3(blocks_for(n)), dim3(256), 0, st, n, iso, (T)t,
                       Vx, Vy, Vz, gWx, gWy, gWz, gs, gVx, gVy, gVz);
}

#define TV3_AD_INST(T)                                                                         \
    template void shrink_apply<T>(int64_t, int, double, const T *, const T *, const T *, T *, \
                                  T *, T *, T *, const uint8_t *, uint8_t *, hipStream_t);    \
    template void shrink_adjoint<T>(int64_t, int, double, const T *, const T *, const T *,    \
                                    const T *, const T *, const T *, const T *, T *, T *, T *, \
                                    hipStream_t);
TV3_AD_INST(float)
TV3_AD_INST(double)

}  // namespace tv3
