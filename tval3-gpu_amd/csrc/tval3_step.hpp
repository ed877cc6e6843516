// tval3_step.hpp -- the per-problem scalar step of the device-driven TVAL3
// solvers (the fused solver and the batched solver): the control block, the
// shrinkage of one voxel, the nonmonotone line search on f(alpha), the end of
// an inner iteration, the first-iteration step length, the initial scalars
// and the continuation step.  Reference: 3d/tval3_gpu_3d.cpp:1088-1345
// (formulas and order of decisions; the routines are cited where used).
#pragma once

#include <cmath>

#include "tv3_internal.hpp"

namespace tv3 {

// Control block of one problem (FP64 scalars, decision flags).
struct FCtl {
    // parameters
    double mu, beta, muDbeta, c, gamma, rate_gam, tol_inn, gam_opt, rate_cnt, mu_max, beta_max;
    int32_t maxit, nonneg, iso, hist_cap;
    // state
    double f, lam1, lam2, lam3, lam4, lam5, C, Q, gam, tau, alpha, dd, nrmup2, relchg;
    double ratio, mdb_old, relout;
    double beta_w;        // beta of the stored W while sigw_old is set (the old beta)
    double f_trial;       // f of the iteration's first trial (alpha = 1)
    double scl;           // scaleb's factor of b (:225-266), found on the device
    int32_t b_bad;        // b holds NaN or Inf
    int32_t total, fst, active, inner_done, need_fallback, nonfinite, sigw_old, cnt;
    int32_t backtracks, fallbacks, inner_loops;
    int32_t n_first;      // first-iteration step lengths taken (one A d product each)
    // batched solver only
    int32_t cur;          // copy (0/1) of U, g, Au holding the accepted iterate
    int32_t mode;         // 1: the next iteration accepts its trial (the fallback step)
    int32_t need_tau;     // the next iteration needs the first-iteration step length
    int32_t outer_flag;   // the outer update runs for this problem
    int32_t done;         // the problem has finished
    int32_t outer;        // outer iterations completed
    // A^T-skip heuristic (:398-468), fused solver only
    double skip_ratio, max_rel, cur_skip;
    int32_t n_init, cur_init;
    int32_t mul_T;        // the current get_g multiplies by A^T
    int32_t n_T;          // A^T products taken by get_g inside iterations
};

constexpr int HF = 16;   // history fields per iteration (14: A^T product taken, 15: f_trial)

// W = shrink(DU - sigma/beta) of one voxel (3d/tval3_gpu_3d_kernels.cu:1249-1300
// and the isotropic variant); returns the lam1 contribution.  ib = 1/beta:
// sigma/beta is formed as sigma * ib, the quotient itself when beta is a
// power of two (the default beta0 and continuation rate keep it one); the
// isotropic |b| and scale (|b| - ib)/|b| use the hardware reciprocal square
// root (about 1 ulp).  The plain solver keeps the quotients of the oracle.
__device__ __forceinline__ float shrink3(float ux, float uy, float uz, float sx, float sy, float sz,
                                        float ib, bool iso, float &wx, float &wy, float &wz) {
    const float bx = fmaf(-sx, ib, ux), by = fmaf(-sy, ib, uy), bz = fmaf(-sz, ib, uz);
    if (!iso) {
        const float tx = fmaxf(fabsf(bx) - ib, 0.f), ty = fmaxf(fabsf(by) - ib, 0.f),
                    tz = fmaxf(fabsf(bz) - ib, 0.f);
        wx = bx >= 0 ? tx : -tx;
        wy = by >= 0 ? ty : -ty;
        wz = bz >= 0 ? tz : -tz;
        return tx + ty + tz;
    }
    // 1/|b| of max(|b|^2, ib^2): below ib the shrink is 0 either way, and the
    // hardware reciprocal square root flushes denormal inputs to zero (an
    // infinite 1/|b| would turn W into Inf/NaN)
    const float n2 = bx * bx + by * by + bz * bz;
    const float r = __builtin_amdgcn_rsqf(fmaxf(n2, ib * ib));
    const float sh = n2 * r - ib;
    if (sh > 0.f) {
        const float f = sh * r;
        wx = bx * f; wy = by * f; wz = bz * f;
        return sh;
    }
    wx = wy = wz = 0.f;
    return 0.f;
}

__device__ __forceinline__ double fval(const FCtl *C, double l1, double l2, double l3, double l4,
                                      double l5) {
    return l1 + C->beta / 2 * l2 + C->mu / 2 * l3 - l4 - l5;
}

// Nonmonotone Armijo line search on f(alpha) (:839-890); mode 1 = accept.
// a = {|V_t|^2, <V_t,sigma>, <V_c,V_t>}, b = {|r_t|^2, <delta,r_t>, <r_c,r_t>}:
// U, Au, DU are affine in alpha, so lam2/lam3 are quadratic and lam4/lam5
// linear in alpha, and the backtracking sweeps of update_g (:706-818) are
// evaluated on these scalars.  keep_active: a failed search (fallback) sets
// need_fallback but leaves active to the iteration's end (kernels that run
// beside the search read active).
__device__ inline void line_search(FCtl *C, const double *a, const double *b, int mode,
                                   bool keep_active = false) {
    const double l2t = a[0], l4t = a[1], x01 = a[2], l3t = b[0], l5t = b[1], r01 = b[2];
    const double l2c = C->lam2, l4c = C->lam4, l3c = C->lam3, l5c = C->lam5;
    auto at = [&](double al, double *o) {
        const double bl = 1.0 - al;
        o[0] = bl * bl * l2c + 2 * al * bl * x01 + al * al * l2t;
        o[1] = bl * bl * l3c + 2 * al * bl * r01 + al * al * l3t;
        o[2] = bl * l4c + al * l4t;
        o[3] = bl * l5c + al * l5t;
    };
    double L[4] = {l2t, l3t, l4t, l5t};
    double f = fval(C, C->lam1, L[0], L[1], L[2], L[3]);
    if (mode != 1) C->f_trial = f;   // a fallback keeps the rejected trial's
    double alpha = 1.0;
    int cnt = 0;
    const double c_arm = C->dd * C->tau * C->c * C->beta;
    if (mode == 1) {
        cnt = 6;
    } else if (fabs(f) > fabs(C->C - alpha * c_arm)) {
        while (fabs(f) > fabs(C->C - alpha * c_arm)) {
            if (cnt == 5) {
                C->need_fallback = 1;
                C->need_tau = 1;
                if (!keep_active) C->active = 0;
                return;
            }
            alpha *= C->gamma;
            at(alpha, L);
            f = fval(C, C->lam1, L[0], L[1], L[2], L[3]);
            ++cnt;
        }
    }
    C->alpha = alpha;
    C->cnt = cnt;
    C->backtracks += cnt > 5 ? 5 : cnt;
    C->lam2 = L[0]; C->lam3 = L[1]; C->lam4 = L[2]; C->lam5 = L[3];
    C->f = f;
}

// f, C, Q, BB step, history and stopping tests after update_W (:965-1029,
// :1176-1197, :663); a = the 7 sums of the update_W kernel: lam1, lam2, lam4,
// |U - Up|^2, <U - Up, d - d_prev>, |Up|^2, |d|^2.
__device__ inline void iter_end(FCtl *C, const double *a, double *hist) {
    C->lam1 = a[0]; C->lam2 = a[1]; C->lam4 = a[2];
    const double ss = a[3], sy = a[4];
    C->nrmup2 = a[5];
    C->dd = a[6];
    C->f = fval(C, C->lam1, C->lam2, C->lam3, C->lam4, C->lam5);   // :973, :1028
    const double Qp = C->Q;
    C->Q = C->gam * Qp + 1;
    C->C = (C->gam * Qp * C->C + C->f) / C->Q;                      // :1176-1177
    C->relchg = sqrt(ss) / sqrt(C->nrmup2);                          // :1189-1197
    const int k = C->total;
    if (k < C->hist_cap) {
        double *h = hist + (int64_t)k * HF;
        h[0] = C->f; h[1] = C->lam1; h[2] = C->lam2; h[3] = C->lam3; h[4] = C->lam4;
        h[5] = C->lam5; h[6] = C->tau; h[7] = C->alpha; h[8] = C->relchg; h[9] = C->mu;
        h[10] = C->beta; h[11] = C->C; h[12] = C->cnt; h[13] = C->inner_loops - 1;
        h[14] = C->skip_ratio != 0.0 ? C->mul_T : 1;
        h[15] = C->f_trial;
    }
    C->total = k + 1;
    C->tau = fabs(ss / sy);                                          // BB step, :663
    C->sigw_old = 0;
    C->mode = 0;
    if (!isfinite(C->f)) { C->nonfinite = 1; C->active = 0; }
    if (C->relchg < C->tol_inn && !C->fst) { C->inner_done = 1; C->active = 0; }
    C->fst = 0;
    if (C->total >= C->maxit) C->active = 0;
}

// First-iteration step length tau = |dd / (|Dd|^2 + mu/beta |Ad|^2)|
// (:583-625); a fallback also shrinks gam (:877) and accepts its trial.
__device__ inline void tau_first_step(FCtl *C, double dDd, double Add, int fallback) {
    if (fallback) {
        C->gam *= C->rate_gam;
        C->fallbacks += 1;
        C->need_fallback = 0;
        C->active = 1;
        C->mode = 1;
    }
    C->need_tau = 0;
    C->n_first += 1;
    C->tau = fabs(C->dd / (dDd + C->muDbeta * Add));
}

// After the initial get_g / update (:1146-1153): a = {|r|^2, <delta,r>, .},
// w = the 7 sums of the update_W kernel.
__device__ inline void init_step(FCtl *C, const double *a, const double *w) {
    C->lam3 = a[0]; C->lam5 = a[1];
    C->lam1 = w[0]; C->lam2 = w[1]; C->lam4 = w[2]; C->dd = w[6];
    C->f = fval(C, C->lam1, C->lam2, C->lam3, C->lam4, C->lam5);
    C->Q = 1; C->C = C->f; C->gam = C->gam_opt;
    C->total = 0; C->fst = 1; C->active = 1; C->inner_done = 0; C->need_fallback = 0;
    C->nonfinite = 0; C->sigw_old = 0; C->cnt = 0; C->alpha = 1; C->backtracks = 0;
    C->fallbacks = 0; C->inner_loops = 1; C->mode = 0; C->need_tau = 1; C->n_first = 0;
}

// A^T-skip decision of get_g (:398-468) from |Au - last Au|^2 and |Au|^2:
// products when forced (the initial and the fallback call), for the first
// n_init calls, while the skip counter is below 1 (it then grows by the
// ratio) and when Au moved by more than max_rel since the last product;
// otherwise skip and count down.
__device__ inline void skip_decide(FCtl *C, double dn, double an, int force) {
    const double rel = sqrt(dn) / sqrt(an);
    int mul = 1;
    if (force) {
        C->cur_skip = 0.0;
    } else if (C->cur_init < C->n_init) {
        ++C->cur_init;
        C->cur_skip = 0.0;
    } else if (C->cur_skip < 1.0) {
        C->cur_skip += C->skip_ratio;
    } else if (rel > C->max_rel) {
        // Au moved too far: multiply
    } else {
        C->cur_skip -= 1.0;
        mul = 0;
    }
    C->mul_T = mul;
    C->n_T += mul;
}

// update_mlp bookkeeping and continuation (:1034-1083, :1222-1249)
__device__ inline void outer_step(FCtl *C, double l4, double l5) {
    C->lam4 = l4; C->lam5 = l5;
    const double beta0 = C->beta;
    C->beta *= C->rate_cnt;
    C->mu *= C->rate_cnt;
    if (C->beta > C->beta_max) C->beta = C->beta_max;
    if (C->mu > C->mu_max) C->mu = C->mu_max;
    C->mdb_old = C->muDbeta;
    C->muDbeta = C->mu / C->beta;
    C->ratio = beta0 / C->beta;
    C->f = fval(C, C->lam1, C->lam2, C->lam3, C->lam4, C->lam5);
    C->gam = C->gam_opt; C->Q = 1; C->C = C->f;
    C->beta_w = beta0;    // W of the last update_W: old sigma, old beta (:984-992)
    C->sigw_old = 1; C->fst = 1; C->inner_done = 0; C->active = 1; C->need_fallback = 0;
    C->inner_loops += 1;
    C->need_tau = 1;
}

}  // namespace tv3
