/*
 * tv3_oracle.h -- CPU restatement of the TVAL3 3-D hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP implementation's results are checked
 * against this library: it may be linked or called only from tests/, from
 * __graft_entry__.smoke() and from the cpu_baseline leg of bench.py.  The
 * product path never calls it.
 *
 * Every function is compiled twice: once with REAL=double (suffix _f64, the
 * FP64 oracle) and once with REAL=float (suffix _f32, the FP32 restatement
 * that the GPU FP32 path is compared against at 1e-4).  Inner products and
 * norms always accumulate in double, scalars (mu, beta, f, lam1..5, tau,
 * alpha, C, Q) are double in both builds.
 *
 * Reference files followed (paths relative to the reference root):
 *   3d/tval3_gpu_3d.cpp:204-1345   host loop, formulas, order of updates
 *   3d/tval3_gpu_3d_kernels.cu     :75-108 vec_add, :170-215 D, :479-510 Dt,
 *                                  :724-771 lam2/lam4 (corrected as in
 *                                  2d/tval3.cu:402-446), :876-919 lam3/lam5,
 *                                  :1017-1067 BB step, :1176-1188 nonneg,
 *                                  :1249-1300 shrinkage, :1612-1764 Bresenham
 *   3d/mat_vec_mul_sparse.h:1703-1780  CSR y=A x / CSC-as-CSR y=A^T x, beta=0
 *   3d/mat_vec_mul_dense.h:25-73       column-major gemv
 *   3d/mat_vec_mul_dynamic.h:501-545   compacted ray order
 *   2d/tval3_types.h:7-34              option names and defaults
 */
#ifndef TV3_ORACLE_H_
#define TV3_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* options / info / history                                            */
/* ------------------------------------------------------------------ */
typedef struct {
    double mu, mu0, beta, beta0, rate_cnt, tol, tol_inn;
    int maxcnt, maxit;
    double c, gamma, gam, rate_gam;
    int nonneg;
    int tv_iso;          /* 0 = anisotropic shrinkage (reference), 1 = isotropic */
    /* A^T-skip heuristic (3d/tval3_gpu_3d.cpp:398-468); skipMulRatio = 0 off */
    int numInitIterations;
    double skipMulRatio, maxRelativeChange;
} orc_opts;

typedef struct {
    int total_iters, outer_iters;
    double rel_chg, rel_error;
    int backtracks, fallbacks;
    int status;          /* 0 ok, 6 non-finite */
    int n_matvec_T;      /* A^T products of the loop (A^T b excluded) */
    double loop_secs;    /* wall time of the outer loop (setup excluded), as the
                          * reference's loop timer (3d/tval3_gpu_3d.cpp:1159,1254-1257) */
} orc_info;

/* per inner iteration record (capacity = cap entries, count written) */
typedef struct {
    int cap, count;
    double *f, *lam1, *lam2, *lam3, *lam4, *lam5, *tau, *alpha, *relchg, *mu, *beta, *C;
    int *cnt;            /* backtracking steps taken (6 = fallback) */
    int *outer;          /* outer index of the iteration */
    int *mulT;           /* 1: the iteration's get_g multiplied by A^T; NULL to omit */
    /* line search of min_u (3d/tval3_gpu_3d.cpp:839-890), each NULL to omit:
     * f of the first trial (alpha = 1), f of the accepted trial (before
     * update_W), the reference value C the Armijo test compared against,
     * c_armij = |d|^2 tau c beta, |d|^2 of the direction, and gam as used by
     * the C/Q update of the iteration (:1176-1177) */
    double *f_trial, *f_acc, *C_in, *c_armij, *dd, *gam;
    /* state at the end of the solve (scaled problem, REAL values widened),
     * each NULL to omit: sigma (3n: x, y, z components), delta (m), the
     * DtsAtd term of the gradient (n), the last gradient d (n) */
    double *sigma_out, *delta_out, *dts_out, *d_out;
} orc_history;

void orc_opts_default(orc_opts *o);

/* OpenMP threads of the parallel loops (n > 0 sets, n <= 0 queries); returns
 * the count now in effect.  Results do not depend on it: every reduction
 * adds 64 fixed chunks in order. */
int orc_set_threads(int n);

/* ------------------------------------------------------------------ */
/* operators                                                           */
/* ------------------------------------------------------------------ */
enum { ORC_OP_DENSE = 0, ORC_OP_CSR = 1, ORC_OP_BRESENHAM = 2, ORC_OP_SIDDON = 3 };

/* An operator description; all arrays are caller owned, host memory. */
typedef struct {
    int kind;
    int64_t m, n;
    /* dense: column-major m x n, leading dimension lda (values in double) */
    const double *dense; int64_t lda;
    /* CSR (A) and CSC (A^T as CSR); values in double */
    const int32_t *rowptr, *col; const double *val;
    const int32_t *colptr, *row; const double *cval;
    /* geometry (Bresenham / Siddon) */
    int num_emitters, num_receivers, num_mps;
    int rv_x, rv_y, rv_z;              /* rv_y fastest: idx = z*rv_x*rv_y + x*rv_y + y */
    double scale_factor;
    const int32_t *ex, *ey, *ez, *rx, *ry, *rz;   /* Bresenham voxel coordinates */
    const double *epos, *rpos;         /* Siddon world positions, 3 per transducer (x,y,z) */
    double origin[3], voxel[3];        /* world coords of the volume corner and voxel size (x,y,z) */
    const uint8_t *use_path;           /* E*R mask or NULL */
    /* compacted ray list (built by orc_build_paths) */
    int64_t num_paths; const int32_t *path_e, *path_r;
} orc_op;

/* Compacted path list in reference order (3d/mat_vec_mul_dynamic.h:501-545):
 * i runs over mp-major, emitter, receiver; pairs with use_path[e*R+r]!=0 kept.
 * path_e/path_r get global transducer indices (mp*E+e, mp*R+r).  Returns count. */
int64_t orc_build_paths(int E, int R, int mps, const uint8_t *use_path,
                        int32_t *path_e, int32_t *path_r);

/* Bresenham walk of 3d/tval3_gpu_3d_kernels.cu:1656-1759; writes the visited
 * voxel linear indices (coordinates clamped into the grid), returns count.
 * cap must be >= max(|d|)+1. */
int orc_bresenham_ray(int x0, int y0, int z0, int x1, int y1, int z1,
                      int rv_x, int rv_y, int rv_z, int32_t *vox, int cap);

/* Siddon exact radiological path of the segment p0->p1 through the box
 * [origin, origin+dims*voxel] (merged plane-crossing parameters, FP64).
 * Writes voxel indices and intersection lengths in traversal order,
 * segments shorter than min_len dropped.  Returns count (<= cap). */
int orc_siddon_ray(const double p0[3], const double p1[3], const double origin[3],
                   const double voxel[3], int rv_x, int rv_y, int rv_z,
                   int32_t *vox, double *len, int cap, double min_len);

/* Build the CSR of a Siddon geometry (rows = compacted paths). Call with
 * col==NULL to get rowptr only (count pass). */
int64_t orc_siddon_csr(const orc_op *op, int32_t *rowptr, int32_t *col, double *val);

/* Stable CSR -> CSC transpose (entries of a column ordered by row). */
void orc_csr_transpose(int64_t m, int64_t n, const int32_t *rowptr, const int32_t *col,
                       const double *val, int32_t *colptr, int32_t *row, double *cval);

/* ------------------------------------------------------------------ */
/* precision-dependent entry points                                    */
/* ------------------------------------------------------------------ */
#define ORC_DECL(REAL, SFX)                                                              \
void orc_matvec##SFX(const orc_op *op, int trans, const REAL *x, REAL *y);               \
void orc_D##SFX(int p, int q, int r, const REAL *U, REAL *Ux, REAL *Uy, REAL *Uz);        \
void orc_Dt##SFX(int p, int q, int r, const REAL *X, const REAL *Y, const REAL *Z,        \
                 REAL *res);                                                             \
double orc_shrink##SFX(int64_t n, int iso, const REAL *Ux, const REAL *Uy,                \
                       const REAL *Uz, const REAL *sx, const REAL *sy, const REAL *sz,   \
                       double beta, REAL *Wx, REAL *Wy, REAL *Wz);                       \
double orc_scaleb##SFX(int64_t m, REAL *b);                                              \
int orc_siddon_walk##SFX(const REAL p0[3], const REAL p1[3], const REAL org[3],           \
                         const REAL h[3], int rv_x, int rv_y, int rv_z, int32_t *vox,     \
                         REAL *len, int cap);                                            \
int64_t orc_siddon_walk_csr##SFX(const orc_op *op, int32_t *rowptr, int32_t *col,         \
                                 REAL *val);                                             \
int orc_solve##SFX(const orc_op *op, const orc_opts *o, int p, int q, int r,              \
                   const REAL *b, REAL *U, const REAL *Ut, orc_info *info,               \
                   orc_history *hist);

ORC_DECL(double, _f64)
ORC_DECL(float, _f32)

#ifdef __cplusplus
}
#endif
#endif
