/*
 * tval3_mi355x.h -- C ABI of the MI355X-native TVAL3 solver.
 *
 * Plain C, no exceptions, no HIP or torch types in the signatures: callable
 * from C, from C++ (include/tval3_gpu_3d.h restores the reference's C++
 * overload set on top of it) and from Python via ctypes.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference root).
 *
 * Conventions
 *   - Volumes are column-major with dim_y fastest: U[y + x*dim_y + z*dim_y*dim_x]
 *     (3d/container_device.h:21-23).  dims[3] = {dim_y, dim_x, dim_z}.
 *   - Ray order of dynamic operators is the reference's compacted order:
 *     measurement position, then emitter, then receiver, masked pairs skipped
 *     (3d/mat_vec_mul_dynamic.h:501-545).
 *   - "_dev" pointers are device pointers, "_host" pointers host memory.
 *   - Every function returns a tv3_status; tv3_last_error() gives the text.
 *   - A context is confined to one host thread (the reference keeps its
 *     handles and streams as process globals, 3d/tval3_gpu_3d.cpp:71-79).
 */
#ifndef TVAL3_MI355X_H_
#define TVAL3_MI355X_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TV3_ABI_VERSION 1

typedef enum {
    TV3_OK = 0,
    TV3_E_ARG = 1,        /* invalid argument (NULL, negative size, bad enum)          */
    TV3_E_SHAPE = 2,      /* sizes inconsistent (3d/tval3_gpu_3d.cpp:1369-1380)        */
    TV3_E_FORMAT = 3,     /* unsupported storage (3d/tval3_gpu_3d.cpp:1404-1406,1463)  */
    TV3_E_NOMEM = 4,
    TV3_E_HIP = 5,        /* HIP runtime error (reference: HANDLE_ERROR / cuda_exception) */
    TV3_E_NONFINITE = 6,  /* NaN/Inf in b, A or the iterates                            */
    TV3_E_INTERNAL = 7,
    TV3_E_UNSUPPORTED = 8 /* option the reference has but this library does not support  */
} tv3_status;

typedef enum { TV3_N = 0, TV3_T = 1 } tv3_trans;   /* CUBLAS_OP_N / CUBLAS_OP_T */

typedef enum { TV3_F32 = 0, TV3_F16 = 1, TV3_BF16 = 2, TV3_F64 = 3 } tv3_dtype;

typedef enum {
    TV3_WEIGHT_BRESENHAM = 0, /* integer 3-D walk, constant weight (3d/tval3_gpu_3d_kernels.cu:1612-1764) */
    TV3_WEIGHT_SIDDON = 1     /* exact intersection lengths (Siddon 1985)                                  */
} tv3_weight_mode;

typedef enum { TV3_TV_ANISO = 0, TV3_TV_ISO = 1 } tv3_tv_mode;

typedef enum {
    TV3_SOLVER_PLAIN = 0,  /* reference-shaped host loop: one kernel per vector op, host scalars */
    TV3_SOLVER_FUSED = 1   /* fused stencil kernels, device FP64 scalars, hipGraph iterations     */
} tv3_solver_kind;

typedef struct tv3_ctx tv3_ctx; /* owns the stream, workspace, graphs and tuning cache */
typedef struct tv3_op tv3_op;   /* owns device copies of a system matrix / geometry   */

/* Options: names and defaults of tval3_options (2d/tval3_types.h:7-34) plus
 * the 3-D skip parameters (3d/tval3_gpu_3d.cpp:151-157) and build-specific
 * fields.  Always initialise with tv3_opts_default(). */
typedef struct {
    double mu, mu0, beta, beta0, rate_cnt, tol, tol_inn;
    int32_t maxcnt, maxit;
    double c, gamma, gam, rate_gam;
    int32_t isreal, nonneg;
    /* A^T-skip heuristic (3d/tval3_gpu_3d.cpp:398-468), plain, fused and
     * batched solves (a decision per problem): skipMulRatio = 0 turns it off.  A skipped A^T product leaves g as
     * A^T(last Au) - Atb (the reference subtracts Atb again from the stale g). */
    uint32_t numInitIterations;
    double skipMulRatio, maxRelativeChange;
    int32_t tv_mode;              /* tv3_tv_mode                                    */
    int32_t solver;               /* tv3_solver_kind                                */
    int32_t warm_start;           /* 1: U on entry is the initial guess, else A^T b */
    int32_t graph_iters;          /* fused: inner iterations per graph launch, 0 = auto */
    int32_t verbose;
    int32_t tiles_per_wg;         /* fused, 8x8x8 tiles: tiles per stencil workgroup,
                                   * 0 = auto (as many workgroups as are resident)  */
} tv3_opts;

/* tval3_info (2d/tval3_types.h:37-43; 3-D adds outer_secs, 3d/tval3_gpu_3d.cpp:1507) */
typedef struct {
    int32_t total_iters, outer_iters;
    double rel_chg, rel_error;
    double secs;          /* GPU time of the whole solve (events), as info.secs        */
    double outer_secs;    /* host wall time of the entry point including uploads        */
    double loop_secs;     /* GPU time of the outer loop only                            */
    int32_t backtracks, fallbacks;
    int64_t n_matvec_N, n_matvec_T;
    int32_t status;
} tv3_info;

/* Optional per-inner-iteration history; arrays are caller owned, length cap. */
typedef struct {
    int32_t cap, count;
    double *f, *lam1, *lam2, *lam3, *lam4, *lam5, *tau, *alpha, *relchg, *mu, *beta, *C;
    int32_t *cnt;     /* backtracking steps; 6 = the cnt==5 fallback */
    int32_t *outer;
    int32_t *mulT;    /* 1: the iteration's get_g took the A^T product, 0: the
                       * A^T-skip heuristic skipped it; NULL to omit */
    double *f_trial;  /* f of the iteration's first trial (alpha = 1) that the
                       * Armijo test of min_u saw (3d/tval3_gpu_3d.cpp:848); NULL
                       * to omit */
} tv3_history;

/* Transducer geometry of a dynamic operator (geometry_host / geometry_device,
 * 3d/container_device.h:103-172).  Volume: rv_x * rv_y * rv_z voxels, rv_y
 * fastest, so dims = {rv_y, rv_x, rv_z}. */
typedef struct {
    int32_t num_emitters, num_receivers, num_mps;
    int32_t rv_x, rv_y, rv_z;
    float scale_factor;                      /* Bresenham weight per voxel           */
    const int32_t *x_emitters, *y_emitters, *z_emitters;    /* num_mps*E, Bresenham */
    const int32_t *x_receivers, *y_receivers, *z_receivers; /* num_mps*R, Bresenham */
    const float *pos_emitters;               /* num_mps*E*3 world (x,y,z), Siddon     */
    const float *pos_receivers;              /* num_mps*R*3 world (x,y,z), Siddon     */
    float origin[3];                         /* world (x,y,z) of the volume corner    */
    float voxel[3];                          /* voxel size (x,y,z)                    */
    const uint8_t *use_path;                 /* E*R mask, NULL = all pairs            */
} tv3_geometry;

/* ---- version / defaults ------------------------------------------------- */
int32_t tv3_abi_version(void);
void tv3_opts_default(tv3_opts *opts);             /* 2d/tval3_types.h:29-34 */

/* ---- context ------------------------------------------------------------ */
/* Replaces the process globals of 3d/tval3_gpu_3d.cpp:71-79 (streams, handles,
 * majorRevision) and the cublasCreate/cudaStreamCreate calls at :1110-1113,1473. */
tv3_status tv3_ctx_create(int32_t device, tv3_ctx **out);
void tv3_ctx_destroy(tv3_ctx *ctx);
const char *tv3_last_error(const tv3_ctx *ctx);    /* NULL ctx: last create error */
void *tv3_ctx_stream(tv3_ctx *ctx);                /* the context's hipStream_t    */

/* ---- operators ---------------------------------------------------------- */
/* Sparse operator: replaces sparse_mm_multGPU(A, 1, {dev}, dev)
 * (3d/mat_vec_mul_sparse.h:79-694, single device) and sparse_mat_device
 * (3d/container_device.h:48-98).  Host arrays; values in val_type (F32, F16,
 * BF16 or F64).  csc_* may be NULL: the CSC (used for A^T) is then built on
 * the device by a stable transpose.  The reference requires sparse_mat_both
 * (3d/tval3_gpu_3d.cpp:1463-1465). */
tv3_status tv3_op_csr_create(tv3_ctx *ctx, int64_t m, int64_t n, int64_t nnz,
                             const int32_t *rowptr, const int32_t *col, const void *val,
                             tv3_dtype val_type, const int32_t *csc_colptr,
                             const int32_t *csc_row, const void *csc_val, tv3_op **out);

/* Dense operator: replaces dense_mm (3d/mat_vec_mul_dense.h:13-20).
 * Column-major m x n with leading dimension lda >= m, host memory. */
tv3_status tv3_op_dense_create(tv3_ctx *ctx, int64_t m, int64_t n, const void *A, int64_t lda,
                               tv3_dtype val_type, tv3_op **out);

/* Matrix-free operator: replaces geom_mm_multGPU(A, 1, {dev}, dev, 16)
 * (3d/mat_vec_mul_dynamic.h:330-637, entry at 3d/tval3_gpu_3d.cpp:1570).
 * Ray weights are generated inside every product. */
tv3_status tv3_op_dynamic_create(tv3_ctx *ctx, const tv3_geometry *geom, tv3_weight_mode mode,
                                 tv3_op **out);

/* Hybrid operator: replaces sparse_geom_mm(hA_geo, hA_mat, ...)
 * (3d/mat_vec_mul.h:104-120, products :732-756; entry 3d/tval3_gpu_3d.cpp:
 * 1603-1668): A x by the dynamic traversal of geom, A^T x by the caller's
 * precomputed matrix (num_paths x voxels, host arrays, values in val_type).
 * The CSC (csc_*) is used for A^T as given; a CSR alone is transposed on the
 * device; with neither (all pointers NULL) the matrix of the geometry itself
 * is built on the device. */
tv3_status tv3_op_hybrid_create(tv3_ctx *ctx, const tv3_geometry *geom, tv3_weight_mode mode,
                                int64_t nnz, const int32_t *rowptr, const int32_t *col,
                                const void *val, tv3_dtype val_type, const int32_t *csc_colptr,
                                const int32_t *csc_row, const void *csc_val, tv3_op **out);

/* Precompute the sparse matrix of a dynamic operator on the device (CSR by a
 * count pass, a scan and a fill pass; CSC by a stable transpose).  Values are
 * stored as val_type (F32, F16, BF16).  No reference counterpart (the
 * reference expects host-built matrices). */
tv3_status tv3_op_materialize(tv3_op *dynamic_op, tv3_dtype val_type, tv3_op **out_csr);

/* Frees the operator's device memory after the work queued on its device
 * has finished.  It does not touch the context: an operator may be destroyed
 * before or after the context that created it. */
void tv3_op_destroy(tv3_op *op);
tv3_status tv3_op_info(const tv3_op *op, int64_t *m, int64_t *n, int64_t *nnz);

/* Copy a sparse operator's CSR to host memory (values converted to double).
 * Pass NULL arrays to query nnz only. */
tv3_status tv3_op_csr_download(tv3_op *op, int32_t *rowptr, int32_t *col, double *val);

/* ---- products ----------------------------------------------------------- */
/* mat_vec_mul(trans, A, x, y, stream, mulChain) overload set
 * (3d/mat_vec_mul_sparse.h:1703-1780, 3d/mat_vec_mul_dense.h:25-73,
 * 3d/mat_vec_mul_dynamic.h:682-967, 3d/mat_vec_mul.h:732-756).
 * x, y device pointers; y is fully overwritten; asynchronous on stream
 * (a hipStream_t; NULL is the null stream, as for the reference's
 * stream = 0 default; pass tv3_ctx_stream(ctx) for the context stream).
 * mulChain is a multi-GPU hint and has no counterpart here. */
tv3_status tv3_matvec(tv3_op *op, tv3_trans trans, const float *x_dev, float *y_dev,
                      void *stream);
tv3_status tv3_matvec_f64(tv3_op *op, tv3_trans trans, const double *x_dev, double *y_dev,
                          void *stream);
/* tv3_matvec with no device allocation.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The only scratch a
 * product needs, the split-K partial sums of a dense operator's A x, lives in
 * a static device array of 16 MiB per GPU (set up by tv3_ctx_create) instead
 * of the context's workspace; a dense A x that needs more (rows times column
 * slabs, min(128, ceil(cols/64)) slabs, 4 bytes each) returns
 * TV3_E_UNSUPPORTED.  Every other product uses no scratch.  Results equal
 * tv3_matvec's bit for bit.  All calls on one GPU share that array: calls
 * that overlap on one GPU (two streams, two host threads) race. */
tv3_status tv3_matvec_noalloc(tv3_op *op, tv3_trans trans, const float *x_dev, float *y_dev,
                              void *stream);
/* Several products at once, Y[:, j] = op X[:, j] (or op^T X[:, j]) for
 * j < nrhs; X and Y column-major device arrays (column j at X + j ldx).  The
 * reference multiplies a dense matrix by several vectors as one GEMM for its
 * multi-layer case (2d/mat_vec_mul.h:91-99).  A dense BF16 operator whose
 * sizes are multiples of 256 runs them on the matrix cores, 64 columns per
 * product (each FP32 value of X split into three BF16 terms, hi + mid + lo,
 * FP32 accumulation; the first such call builds copies of A and A^T in the
 * matrix cores' operand order, 2 m n bytes each, kept with the operator);
 * any other operator applies tv3_matvec per column.  Asynchronous on
 * stream. */
tv3_status tv3_matvec_multi(tv3_op *op, tv3_trans trans, int32_t nrhs, const float *X_dev,
                            int64_t ldx, float *Y_dev, int64_t ldy, void *stream);

/* ---- solver ------------------------------------------------------------- */
/* tval3_gpu_3d(U, A, b, opts, Ut, pagelocked, mainGPU, numGPUs, gpuArray)
 * (3d/tval3_gpu_3d.cpp:1673-1745, core loop :1088-1345) for float data.
 * b_host: m floats; U_host: n floats (output; input too if opts->warm_start);
 * Ut_host: n floats or NULL (info->rel_error); hist may be NULL. */
tv3_status tv3_solve(tv3_ctx *ctx, tv3_op *op, const tv3_opts *opts, const int32_t dims[3],
                     const float *b_host, float *U_host, const float *Ut_host, tv3_info *info,
                     tv3_history *hist);

/* Same with device-resident b, U, Ut (no host transfers inside). */
tv3_status tv3_solve_device(tv3_ctx *ctx, tv3_op *op, const tv3_opts *opts,
                            const int32_t dims[3], const float *b_dev, float *U_dev,
                            const float *Ut_dev, tv3_info *info, tv3_history *hist);

/* double-precision overloads (tval3_info<double> tval3_gpu_3d(...), :1710-1745):
 * the plain solver in FP64 on the device. */
tv3_status tv3_solve_f64(tv3_ctx *ctx, tv3_op *op, const tv3_opts *opts, const int32_t dims[3],
                         const double *b_host, double *U_host, const double *Ut_host,
                         tv3_info *info, tv3_history *hist);

/* ---- batched solves ------------------------------------------------------ */
/* No reference counterpart (the reference solves one problem per call).
 * nprob independent problems, each with its sparse operator (CSR, one
 * weight type for the sparse ones) or a dense BF16 operator that several
 * problems may share (their products then run as one matrix-core product
 * with up to 64 right-hand sides; rows and columns multiples of 256), dims
 * (3 per problem: dim_y, dim_x, dim_z), b and result U (device pointers,
 * float), solved together with one
 * set of options: every kernel of an iteration covers all problems, the
 * iterations are graph-captured, each problem keeps its own scalars,
 * decisions and stopping (its iterates equal a separate fused-schedule solve
 * of that problem up to FP32 summation order).  Ut_dev (array or NULL,
 * entries may be NULL) gives infos[k].rel_error; hists (array of nprob or
 * NULL) receives each problem's history.  opts->solver is ignored; the
 * A^T-skip heuristic decides per problem.  binfo may be
 * NULL; when it is not, lanes_N and lanes_T > 0 on entry fix the sparse
 * products' launch shapes and splits_N, splits_T > 0 the dense products' K
 * splits instead of timing them (out: the shapes used, the first group's). */
typedef struct tv3_batch_info {
    int32_t lanes_N, lanes_T;   /* shapes of the block-diagonal products (in/out): lanes
                                   per row, or -NB for row blocks of NB entries */
    int32_t graph_iters;        /* iterations per graph launch                          */
    int32_t graph_launches;
    double secs;                /* GPU time from scaling b to the unscaled results      */
    double loop_secs;           /* GPU time of the iterations                           */
    double host_secs;           /* host wall time of the call, set-up included          */
    int64_t total_iters;        /* inner iterations summed over the problems            */
    int64_t nnz, voxels, rays;  /* batch totals (per problem, voxels padded to 512 and
                                   rays to 256)                                          */
    int32_t dense_groups;       /* shared dense operators (matrix-core products)        */
    int32_t splits_N, splits_T; /* K splits of the dense products (in/out, like lanes)  */
    int32_t panels_T;           /* A^T of the sparse problems in panel form (tiles of
                                   512 voxels, each tile's rays staged in LDS; every
                                   problem's voxel range is padded to a multiple of 512);
                                   in: 1024, 2048 or 4096 uses it with row blocks of that
                                   many entries, -1 never, 0 timed against the row
                                   kernels;
                                   out: the block size used, 0 when not used */
    int64_t dense_values;       /* entries of the shared dense operators                */
    int32_t panels_A;           /* A of the sparse problems in two-pass panel form
                                   (segment sums per panel, then per ray); in and out
                                   like panels_T (4096 too with panel_box_A = 8)        */
    int32_t panel_box_A;        /* tiles per panel of that form: 1 (8^3 voxels) or 8
                                   (16^3: fewer, longer ray segments; every sparse
                                   problem on 8^3 tiles with even tile counts); in: 0
                                   lets the timing choose when panels_A is 0 and means
                                   1 otherwise; out: the size used, 0 without panels */
} tv3_batch_info;

tv3_status tv3_solve_batched(tv3_ctx *ctx, int32_t nprob, tv3_op *const *ops,
                             const int32_t *dims, const tv3_opts *opts,
                             const float *const *b_dev, float *const *U_dev,
                             const float *const *Ut_dev, tv3_info *infos, tv3_history *hists,
                             tv3_batch_info *binfo);

/* ---- device building blocks (exposed for tests and autograd bindings) ---- */
/* D (3d/tval3_gpu_3d.cpp:271-314): periodic forward differences.
 * Ux along dim_x (stride dim_y), Uy along dim_y (stride 1), Uz along dim_z. */
tv3_status tv3_D(tv3_ctx *ctx, const int32_t dims[3], const float *U_dev, float *Ux_dev,
                 float *Uy_dev, float *Uz_dev, void *stream);
/* D^T (3d/tval3_gpu_3d.cpp:319-361). */
tv3_status tv3_Dt(tv3_ctx *ctx, const int32_t dims[3], const float *X_dev, const float *Y_dev,
                  const float *Z_dev, float *res_dev, void *stream);
/* shrinkage_gpu (3d/tval3_gpu_3d.cpp:924-960): W = shrink(DU - sigma/beta),
 * returns lam1 = sum |W| (aniso) or sum |W|_2 (iso). Synchronous. */
tv3_status tv3_shrink(tv3_ctx *ctx, int64_t n, tv3_tv_mode mode, const float *Ux,
                      const float *Uy, const float *Uz, const float *sx, const float *sy,
                      const float *sz, double beta, float *Wx, float *Wy, float *Wz,
                      double *lam1);
/* D and D^T in double precision (the double instantiation of the same
 * kernels, as the reference's double overloads use). */
tv3_status tv3_D_f64(tv3_ctx *ctx, const int32_t dims[3], const double *U_dev, double *Ux_dev,
                     double *Uy_dev, double *Uz_dev, void *stream);
tv3_status tv3_Dt_f64(tv3_ctx *ctx, const int32_t dims[3], const double *X_dev,
                      const double *Y_dev, const double *Z_dev, double *res_dev, void *stream);
/* The shrinkage of shrinkage_gpu (3d/tval3_gpu_3d.cpp:924-960) as a map of its
 * argument V (= DU - sigma/beta there) at threshold t (= 1/beta there):
 *   aniso  W_c = sign(V_c) max(|V_c| - t, 0),  s = sum_c max(|V_c| - t, 0)
 *   iso    W = max(|V|_2 - t, 0) V/|V|_2,       s = max(|V|_2 - t, 0)
 * per voxel; lam1 = sum(s).  s may be NULL.  dtype TV3_F32 or TV3_F64 (all
 * pointers of that type, device memory); asynchronous on stream.
 * mask_out (n bytes, may be NULL) receives the active set: aniso bit c set
 * when |V_c| > t, iso bit 0 when |V|_2 > t.  mask_in (may be NULL) replaces
 * those comparisons by a recorded active set: the smooth branch it selects
 * is evaluated (active: the formulas above without the max; inactive: 0). */
tv3_status tv3_shrink_apply(tv3_ctx *ctx, int64_t n, tv3_tv_mode mode, tv3_dtype dtype,
                            const void *Vx, const void *Vy, const void *Vz, double t, void *Wx,
                            void *Wy, void *Wz, void *s, const uint8_t *mask_in,
                            uint8_t *mask_out, void *stream);
/* Its vector-Jacobian product: gV = (dW/dV)^T gW + (ds/dV)^T gs (gs may be
 * NULL), the inactive side taken on the threshold surface.  The reference
 * has no adjoint; this is the backward of the autograd binding. */
tv3_status tv3_shrink_adjoint(tv3_ctx *ctx, int64_t n, tv3_tv_mode mode, tv3_dtype dtype,
                              const void *Vx, const void *Vy, const void *Vz, double t,
                              const void *gWx, const void *gWy, const void *gWz, const void *gs,
                              void *gVx, void *gVy, void *gVz, void *stream);

/* ---- the fused solver's operator layout (no reference counterpart) ------ */
typedef struct tv3_layout_info {
    int32_t brick;          /* voxel numbering: 8 (8^3 tiles of 4^3 bricks), 4 (bricks), 1 */
    int32_t panels;         /* bit 0: A x in panel form (per-tile ray segments, then a
                               per-ray sum); bit 1: A^T r in panel form (per-tile rays
                               staged in LDS); bit 2 / bit 3: A's / A^T's panels
                               span two tiles; bit 4 / bit 5: A's panels span
                               2 x 2 x 1 / 2 x 2 x 2 tiles; bit 8: their entries
                               packed (12-bit indices, FP16/BF16 weights); bit 9: A^T's
                               entries packed (panels of at most 4096 rays) */
    int32_t code_N, code_T; /* launch shapes as tv3_op_autotune codes; for a panel form its
                               row block size as -NB */
    int64_t npan, nseg;     /* tiles (panels) and ray segments (tile, ray) of the plan */
    int32_t max_seg;        /* most rays crossing one panel */
    int32_t reserved;
} tv3_layout_info;

/* Build the fused solver's layout of a sparse operator for a volume shape:
 * voxels and rays renumbered for locality, A and its CSC in that numbering,
 * the CSC's panels, launch shapes timed on this GPU.  A FUSED solve builds it
 * on first use; this call takes that setup out of the first solve.  Kept
 * until the operator is destroyed or prepared for another shape. */
tv3_status tv3_op_prepare(tv3_op *op, const int32_t dims[3], tv3_layout_info *info);

/* Override the layout's tuned launch shapes (tests, experiments).  panels:
 * bit 0 puts A x, bit 1 puts A^T r in panel form, rebuilt with panels of at
 * most pan_nnz entries per tile (bit 2 / bit 3: A's / A^T's panels of two
 * tiles; bit 4 / bit 5: A's panels of 2 x 2 x 1 / 2 x 2 x 2 tiles, which need
 * 8^3 tiles and an even tile count along every axis; at most one of bits 2, 4
 * and 5; bit 8 with bit 4 or 5: their entries packed, 12-bit local indices
 * and 16-bit weights in 28 bytes per 8 entries, rows padded to pairs, FP16
 * or BF16 operators and blocks of 2048 or 4096; bit 9 with bit 1: A^T's entries
 * packed the same way, rows in groups of 4, its panels split to at most
 * 4096 rays) and row blocks of -code_N / -code_T
 * entries (-1024, -2048 or -4096);
 * a direction not in panel form takes code_N / code_T as a tv3_op_set_lanes
 * code (0 = keep).  A layout not yet prepared for dims is built here without
 * timing any launch shape (a caller replaying shapes it stored, such as the
 * Python binding's JSON tuning cache, skips the tuner's timings). */
tv3_status tv3_op_layout_force(tv3_op *op, const int32_t dims[3], int32_t code_N, int32_t code_T,
                               int32_t pan_nnz, int32_t panels, tv3_layout_info *info);

/* y = A x (TV3_N) or A^T x (TV3_T) by the fused solver's product kernels in
 * its layout; x and y in natural order (permuted in and out), on stream. */
tv3_status tv3_op_layout_matvec(tv3_op *op, const int32_t dims[3], tv3_trans trans,
                                const float *x_dev, float *y_dev, void *stream);

/* Average duration (microseconds) of one launch of the layout's A (TV3_N) or
 * A^T (TV3_T) product (x = 1; for the panel form of A both of its kernels),
 * timed with HIP events on the context stream around each of `reps`
 * launches.  Before each launch a read of a 512 MiB buffer evicts the
 * operator from the caches, as the solver's other kernels do between two
 * products. */
tv3_status tv3_op_layout_time(tv3_op *op, const int32_t dims[3], tv3_trans trans, int32_t reps,
                              double *us);

/* Achievable HBM bandwidth on the context's device (no reference
 * counterpart; the measurement the roofline fractions are also quoted
 * against): a streaming copy of `bytes` (GB/s counts bytes read plus bytes
 * written) and a streaming read of `bytes`, each the median of `reps`
 * launches timed with HIP events on the context stream.  Either output may
 * be NULL. */
tv3_status tv3_bandwidth_probe(tv3_ctx *ctx, int64_t bytes, int32_t reps, double *copy_GBps,
                               double *read_GBps);

/* ---- tuning ------------------------------------------------------------- */
/* Time the sparse kernels of op over their launch shapes and keep the
 * fastest: the row-blocked kernel with 1024/2048/4096-entry blocks (returned
 * as -1024/-2048/-4096) or the row-group kernel with L lanes per row
 * (returned as L).  tv3_op_set_lanes accepts the same codes (0 = keep). */
tv3_status tv3_op_autotune(tv3_op *op, int32_t *lanes_N, int32_t *lanes_T);
tv3_status tv3_op_set_lanes(tv3_op *op, int32_t lanes_N, int32_t lanes_T);

#ifdef __cplusplus
}
#endif
#endif
