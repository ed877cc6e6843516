// solver_plain.hip -- the reference-shaped TVAL3 loop on the device: one
// kernel per vector operation and host-side scalars, in the order of
// 3d/tval3_gpu_3d.cpp:1088-1345 (get_g :366-572, get_tau :577-666, descend
// :671-701, update_g :706-818, min_u :823-891, get_gradient :896-919,
// shrinkage :924-960, update_W :965-1029, update_mlp :1034-1083).
//
// Differences from the reference, all deliberate:
//  * one stream, so no cross-stream races (the reference's get_g/update_g/
//    update_W read V on a stream that did not wait for its producer);
//  * lam2/lam4 use all three components with sigma in X,Y,Z order;
//  * inner products accumulate in FP64 (block partials summed in a fixed
//    order on the host), scalars are FP64, and every entry of a product
//    A x is summed in FP64 and rounded once, so the FP32 iterates restate
//    the oracle's rounding instead of the accumulation order of a library;
//  * vector kernels are compiled without FP contraction (a*x+y rounds
//    twice, as written);
//  * every work vector is allocated once per solve (the reference allocates
//    and frees 3 n-vectors inside get_g/update_g/update_W/update_mlp);
//  * the A^T-skip heuristic (:398-468) leaves a skipped g as A^T(last Au)
//    - Atb; the reference subtracts Atb a second time from the stale g
//    (:463-467 then :565-567).
// This path is the parity anchor and the baseline the fused path is
// measured against.
#include <chrono>
#include <cmath>

#include "vec.hpp"

namespace tv3 {

namespace {

template <typename T>
struct PlainState {
    tv3_ctx *ctx;
    tv3_op *op;
    const tv3_opts *o;
    hipStream_t s;
    int32_t dims[3];
    int64_t n, m;
    double mu, beta, muDbeta, f, lam1, lam2, lam3, lam4, lam5;
    DevBuf<T> pool;
    T *delta, *sx, *sy, *sz, *Up, *dU, *uup, *Ux, *Uxp, *dUx, *Uy, *Uyp, *dUy, *Uz, *Uzp, *dUz,
        *Wx, *Wy, *Wz, *Atb, *Au, *Aup, *dAu, *d, *g, *gp, *dg, *g2, *g2p, *dg2, *Vx, *Vy, *Vz,
        *dx, *dy, *dz, *Ad, *DtsAtd, *rcdU, *bs, *lastAu, *dlast;
    int64_t nN = 0, nT = 0;
    int cur_init = 0;        // A^T-skip state (:155-157)
    double cur_skip = 0.0;
    bool last_mul = true;    // the last get_g multiplied by A^T

    void alloc() {
        T **nv[] = {&sx, &sy, &sz, &Up, &dU, &uup, &Ux, &Uxp, &dUx, &Uy, &Uyp, &dUy, &Uz, &Uzp,
                    &dUz, &Wx, &Wy, &Wz, &Atb, &d, &g, &gp, &dg, &g2, &g2p, &dg2, &Vx, &Vy, &Vz,
                    &dx, &dy, &dz, &DtsAtd, &rcdU};
        T **mv[] = {&delta, &Au, &Aup, &dAu, &Ad, &bs, &lastAu, &dlast};
        const size_t cn = sizeof(nv) / sizeof(nv[0]), cm = sizeof(mv) / sizeof(mv[0]);
        // pad each vector to 256 bytes
        auto pad = [](int64_t k) { return (k + 63) / 64 * 64; };
        pool.alloc(cn * pad(n) + cm * pad(m));
        TV3_HIP(hipMemsetAsync(pool.p, 0, pool.bytes(), s));
        T *p = pool.p;
        for (size_t k = 0; k < cn; ++k) { *nv[k] = p; p += pad(n); }
        for (size_t k = 0; k < cm; ++k) { *mv[k] = p; p += pad(m); }
    }
    void copy(T *dst, const T *src, int64_t k) {
        TV3_HIP(hipMemcpyAsync(dst, src, sizeof(T) * k, hipMemcpyDeviceToDevice, s));
    }
    void A(bool trans, const T *x, T *y) {
        op_apply(op, trans, x, y, s, /*acc64=*/true);
        (trans ? nT : nN)++;
    }
    double fval() const { return lam1 + beta / 2 * lam2 + mu / 2 * lam3 - lam4 - lam5; }

    void get_lams() {
        auto a = lam24<T>(ctx, n, Ux, Uy, Uz, Wx, Wy, Wz, sx, sy, sz, Vx, Vy, Vz, s);
        lam2 = a[0]; lam4 = a[1];
        auto c = lam35<T>(ctx, m, Au, bs, delta, s);
        lam3 = c[0]; lam5 = c[1];
    }
    // :366-572; force: the initial and the fallback call (:1148, :880)
    void get_g(const T *U, bool force) {
        A(false, U, Au);
        bool mul = true;
        if (o->skipMulRatio != 0.0) {
            // relative change of Au since the last A^T product (:398-412)
            axpy3<T>(m, dlast, -1.0, lastAu, Au, s);
            const double dn = dot<T>(ctx, m, dlast, dlast, s), an = dot<T>(ctx, m, Au, Au, s);
            const double rel = std::sqrt(dn) / std::sqrt(an);
            if (force) {
                cur_skip = 0.0;
            } else if ((uint32_t)cur_init < o->numInitIterations) {
                ++cur_init;
                cur_skip = 0.0;
            } else if (cur_skip < 1.0) {
                cur_skip += o->skipMulRatio;
            } else if (rel > o->maxRelativeChange) {
                // Au moved too far: multiply
            } else {
                cur_skip -= 1.0;
                mul = false;
            }
            if (mul) copy(lastAu, Au, m);
        }
        if (mul) A(true, Au, g);
        get_lams();
        launch_Dt<T>(dims, Vx, Vy, Vz, g2, s);
        if (mul) axpy3<T>(n, g, -1.0, Atb, g, s);
        last_mul = mul;
        f = fval();
    }
    double get_tau(bool fst) {
        if (fst) {
            launch_D<T>(dims, d, dx, dy, dz, s);
            double dDd = dot<T>(ctx, n, dx, dx, s) + dot<T>(ctx, n, dy, dy, s) +
                         dot<T>(ctx, n, dz, dz, s);
            double dd = dot<T>(ctx, n, d, d, s);
            A(false, d, Ad);
            double Add = dot<T>(ctx, m, Ad, Ad, s);
            return std::fabs(dd / (dDd + muDbeta * Add));
        }
        auto r = tau_bb<T>(ctx, n, g, gp, g2, g2p, uup, muDbeta, dg, dg2, s);
        return std::fabs(r[0] / r[1]);
    }
    void descend(T *U, double tau, bool force) {
        axpy3<T>(n, U, -tau, d, U, s);
        if (o->nonneg) nonneg<T>(n, U, s);
        launch_D<T>(dims, U, Ux, Uy, Uz, s);
        get_g(U, force);
    }
    void update_g(double alpha, T *U) {
        axpy3<T>(n, g, alpha, dg, gp, s);
        axpy3<T>(n, g2, alpha, dg2, g2p, s);
        axpy3<T>(n, U, alpha, dU, Up, s);
        axpy3<T>(m, Au, alpha, dAu, Aup, s);
        axpy3<T>(n, Ux, alpha, dUx, Uxp, s);
        axpy3<T>(n, Uy, alpha, dUy, Uyp, s);
        axpy3<T>(n, Uz, alpha, dUz, Uzp, s);
        get_lams();
        f = fval();
    }
    void gradient() { grad<T>(n, d, muDbeta, g, g2, DtsAtd, s); }
    void shrinkage() {
        lam1 = shrink_host<T>(ctx, n, o->tv_mode == TV3_TV_ISO, Ux, Uy, Uz, sx, sy, sz, beta, Wx,
                              Wy, Wz, s);
    }
    void update_W() {
        f -= (lam1 + beta / 2 * lam2 - lam4);
        shrinkage();
        auto a = lam24<T>(ctx, n, Ux, Uy, Uz, Wx, Wy, Wz, sx, sy, sz, Vx, Vy, Vz, s);
        lam2 = a[0]; lam4 = a[1];
        launch_Dt<T>(dims, Vx, Vy, Vz, g2, s);
        f += (lam1 + beta / 2 * lam2 - lam4);
    }
    void update_mlp() {
        f += (lam4 + lam5);
        lam4 = mlp_sigma<T>(ctx, n, Ux, Uy, Uz, Wx, Wy, Wz, beta, sx, sy, sz, s);
        lam5 = mlp_delta<T>(ctx, m, Au, bs, mu, delta, s);
        f -= (lam4 + lam5);
    }
};

// scaleb (3d/tval3_gpu_3d.cpp:225-266) on the host copy of b
template <typename T>
double scale_factor_of(const std::vector<T> &b) {
    if (b.empty()) return 1.0;
    T bmin = b[0], bmax = bmin, amin = std::fabs(bmin), amax = amin;
    for (T v : b) {
        T a = std::fabs(v);
        if (a < amin) { bmin = v; amin = a; }
        if (a > amax) { bmax = v; amax = a; }
    }
    T dif = std::fabs(bmax - bmin);
    T s = 1;
    if (dif < (T)0.5) s = (T)0.5 / dif;
    else if (dif > (T)1.5) s = (T)1.5 / dif;
    return (double)s;
}

}  // namespace

template <typename T>
void solve_plain(tv3_ctx *ctx, tv3_op *op, const tv3_opts &o, const int32_t dims[3],
                 const T *b_dev, T *U, const T *Ut_dev, tv3_info *info, tv3_history *hist) {
    hipStream_t s = ctx->stream;
    PlainState<T> S;
    S.ctx = ctx; S.op = op; S.o = &o; S.s = s;
    for (int a = 0; a < 3; ++a) S.dims[a] = dims[a];
    S.n = (int64_t)dims[0] * dims[1] * dims[2];
    S.m = op->m;
    const int64_t n = S.n, m = S.m;
    S.alloc();

    hipEvent_t ev0, ev1, evl0, evl1;
    TV3_HIP(hipEventCreate(&ev0)); TV3_HIP(hipEventCreate(&ev1));
    TV3_HIP(hipEventCreate(&evl0)); TV3_HIP(hipEventCreate(&evl1));
    TV3_HIP(hipEventRecord(ev0, s));

    // b: scale on the device by the factor found on a host copy (:1122-1124)
    std::vector<T> bh(m);
    TV3_HIP(hipMemcpyAsync(bh.data(), b_dev, sizeof(T) * m, hipMemcpyDeviceToHost, s));
    TV3_HIP(hipStreamSynchronize(s));
    for (T v : bh) TV3_REQUIRE(std::isfinite((double)v), TV3_E_NONFINITE, "b holds NaN or Inf");
    double scl = scale_factor_of<T>(bh);
    scale<T>(m, S.bs, scl, b_dev, s);

    S.A(true, S.bs, S.Atb);
    if (!o.warm_start) S.copy(U, S.Atb, n);
    else scale<T>(n, U, scl, U, s);
    S.mu = std::min(o.mu0, o.mu);
    S.beta = std::min(o.beta0, o.beta);
    S.muDbeta = S.mu / S.beta;
    S.copy(S.rcdU, U, n);
    launch_D<T>(dims, U, S.Ux, S.Uy, S.Uz, s);
    S.shrinkage();
    S.get_g(U, true);
    S.gradient();

    double Q = 1, Qp, C = S.f, gam = o.gam;
    int outer = 0, total = 0, backtracks = 0, fallbacks = 0;
    double RelChg = 0, RelChgOut = 0, nrmup = 0;
    bool nonfinite = false;
    TV3_HIP(hipEventRecord(evl0, s));

    while (outer < o.maxcnt && total < o.maxit) {
        bool fst = true;
        while (total < o.maxit) {
            double tau = S.get_tau(fst);
            S.copy(S.Up, U, n); S.copy(S.gp, S.g, n); S.copy(S.g2p, S.g2, n);
            S.copy(S.Aup, S.Au, m); S.copy(S.Uxp, S.Ux, n); S.copy(S.Uyp, S.Uy, n);
            S.copy(S.Uzp, S.Uz, n);
            S.descend(U, tau, false);
            double alpha = 1.0;
            const double f_trial = S.f;
            axpy3<T>(n, S.dU, -1.0, S.Up, U, s);
            double c_armij = dot<T>(ctx, n, S.d, S.d, s) * tau * o.c * S.beta;
            int cnt = 0;
            if (std::fabs(S.f) > std::fabs(C - alpha * c_armij)) {
                axpy3<T>(n, S.dg, -1.0, S.gp, S.g, s);
                axpy3<T>(n, S.dg2, -1.0, S.g2p, S.g2, s);
                axpy3<T>(m, S.dAu, -1.0, S.Aup, S.Au, s);
                axpy3<T>(n, S.dUx, -1.0, S.Uxp, S.Ux, s);
                axpy3<T>(n, S.dUy, -1.0, S.Uyp, S.Uy, s);
                axpy3<T>(n, S.dUz, -1.0, S.Uzp, S.Uz, s);
                while (std::fabs(S.f) > std::fabs(C - alpha * c_armij)) {
                    if (cnt == 5) {
                        gam *= o.rate_gam;
                        tau = S.get_tau(true);
                        S.copy(U, S.Up, n);
                        S.descend(U, tau, true);
                        cnt = 6;
                        ++fallbacks;
                        break;
                    }
                    alpha *= o.gamma;
                    S.update_g(alpha, U);
                    ++cnt;
                }
                backtracks += cnt > 5 ? 5 : cnt;
            }
            S.update_W();
            Qp = Q; Q = gam * Qp + 1; C = (gam * Qp * C + S.f) / Q;
            axpy3<T>(n, S.uup, -1.0, S.Up, U, s);
            S.gradient();
            ++total;
            nrmup = std::sqrt(dot<T>(ctx, n, S.Up, S.Up, s));
            RelChg = std::sqrt(dot<T>(ctx, n, S.uup, S.uup, s)) / nrmup;
            if (hist && hist->count < hist->cap) {
                int k = hist->count++;
                hist->f[k] = S.f; hist->lam1[k] = S.lam1; hist->lam2[k] = S.lam2;
                hist->lam3[k] = S.lam3; hist->lam4[k] = S.lam4; hist->lam5[k] = S.lam5;
                hist->tau[k] = tau; hist->alpha[k] = alpha; hist->relchg[k] = RelChg;
                hist->mu[k] = S.mu; hist->beta[k] = S.beta; hist->C[k] = C;
                hist->cnt[k] = cnt; hist->outer[k] = outer;
                if (hist->mulT) hist->mulT[k] = S.last_mul;
                if (hist->f_trial) hist->f_trial[k] = f_trial;
            }
            if (!std::isfinite(S.f)) { nonfinite = true; break; }
            if (RelChg < o.tol_inn && !fst) break;
            fst = false;
        }
        if (nonfinite) break;
        ++outer;
        auto dn = diff_norm<T>(ctx, n, U, S.rcdU, s);
        RelChgOut = std::sqrt(dn[0]) / nrmup;
        S.copy(S.rcdU, U, n);
        if (RelChgOut < o.tol) break;
        S.update_mlp();
        double beta0 = S.beta;
        S.beta *= o.rate_cnt; S.mu *= o.rate_cnt;
        if (S.beta > o.beta) S.beta = o.beta;
        if (S.mu > o.mu) S.mu = o.mu;
        S.muDbeta = S.mu / S.beta;
        S.f = S.fval();
        scale<T>(n, S.DtsAtd, beta0 / S.beta, S.d, s);
        S.gradient();
        gam = o.gam; Q = 1; C = S.f;
    }
    TV3_HIP(hipEventRecord(evl1, s));
    divide<T>(n, U, scl, U, s);    // :1275-1277
    TV3_HIP(hipEventRecord(ev1, s));
    TV3_HIP(hipStreamSynchronize(s));
    float ms = 0, msl = 0;
    TV3_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    TV3_HIP(hipEventElapsedTime(&msl, evl0, evl1));
    info->total_iters = total;
    info->outer_iters = outer;
    info->rel_chg = RelChgOut;
    info->secs = ms / 1000.0;
    info->loop_secs = msl / 1000.0;
    info->backtracks = backtracks;
    info->fallbacks = fallbacks;
    info->n_matvec_N = S.nN;
    info->n_matvec_T = S.nT;
    info->status = nonfinite ? TV3_E_NONFINITE : TV3_OK;
    info->rel_error = 0;
    if (Ut_dev) {
        auto e = diff_norm<T>(ctx, n, Ut_dev, U, s);
        info->rel_error = std::sqrt(e[0]) / std::sqrt(e[1]);
    }
    (void)hipEventDestroy(ev0); (void)hipEventDestroy(ev1);
    (void)hipEventDestroy(evl0); (void)hipEventDestroy(evl1);
}

template void solve_plain<float>(tv3_ctx *, tv3_op *, const tv3_opts &, const int32_t *,
                                 const float *, float *, const float *, tv3_info *, tv3_history *);
template void solve_plain<double>(tv3_ctx *, tv3_op *, const tv3_opts &, const int32_t *,
                                  const double *, double *, const double *, tv3_info *,
                                  tv3_history *);

}  // namespace tv3
