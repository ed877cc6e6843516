/*
 * tv3_oracle.c -- CPU restatement of the TVAL3 3-D hot path (for tests only).
 *
 * TEST INFRASTRUCTURE ONLY: linked or called only by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.
 *
 * This file holds the precision-independent parts (ray list, Bresenham walk,
 * Siddon traversal, CSR transpose) and instantiates tv3_oracle_t.inc for
 * REAL = double (_f64) and REAL = float (_f32).
 *
 * Pinning status (each function is checked in tests/test_oracle.py):
 *   orc_build_paths     ray order of 3d/mat_vec_mul_dynamic.h:501-545, checked
 *                       by brute-force enumeration on small E,R,mps with masks.
 *   orc_bresenham_ray   integer walk of 3d/tval3_gpu_3d_kernels.cu:1656-1759,
 *                       checked by known answers (axis rays, diagonals, the
 *                       >= tie cascade of :1679,1687) and a brute-force
 *                       "nearest voxel to the line" property.
 *   orc_siddon_ray      Siddon, Med. Phys. 12(2) 1985: sum of lengths equals
 *                       the slab-clipped chord, axis rays through voxel centres
 *                       give exactly one voxel size per voxel, supersampled
 *                       line integral on 4^3..8^3 grids.
 *   orc_csr_transpose   transpose round trip, brute force against dense.
 */
#include "tv3_oracle.h"

#include <math.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void orc_opts_default(orc_opts *o) {
    /* names and defaults of 2d/tval3_types.h:29-34 */
    o->mu = 256.0; o->mu0 = 256.0; o->beta = 32.0; o->beta0 = 32.0;
    o->rate_cnt = 2.0; o->tol = 1e-6; o->tol_inn = 1e-3;
    o->maxcnt = 10; o->maxit = 1025;
    o->c = 1e-5; o->gamma = 0.6; o->gam = 0.9995; o->rate_gam = 0.9;
    o->nonneg = 1; o->tv_iso = 0;
    o->numInitIterations = 0; o->skipMulRatio = 0.0; o->maxRelativeChange = 0.0;
}

int64_t orc_build_paths(int E, int R, int mps, const uint8_t *use_path,
                        int32_t *path_e, int32_t *path_r) {
    /* 3d/mat_vec_mul_dynamic.h:501-545: i over numMPs*E*R signals,
     * x = i / R (global emitter), y = (i / (E*R))*R + i % R (global receiver) */
    int64_t total = (int64_t)E * R * mps, per_mp = (int64_t)E * R, k = 0;
    for (int64_t i = 0; i < total; ++i) {
        if (use_path && !use_path[i % per_mp]) continue;
        if (path_e) {
            path_e[k] = (int32_t)(i / R);
            path_r[k] = (int32_t)((i / per_mp) * R + i % R);
        }
        ++k;
    }
    return k;
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

int orc_bresenham_ray(int x0, int y0, int z0, int x1, int y1, int z1,
                      int rv_x, int rv_y, int rv_z, int32_t *vox, int cap) {
    int c[3] = {x0, y0, z0};          /* x, y, z pixel */
    int cnt = 0;
#define ORC_VISIT()                                                                    \
    do {                                                                               \
        int xx = clampi(c[0], 0, rv_x - 1), yy = clampi(c[1], 0, rv_y - 1),            \
            zz = clampi(c[2], 0, rv_z - 1);                                            \
        if (cnt < cap) vox[cnt] = (int32_t)((int64_t)zz * rv_x * rv_y + (int64_t)xx * rv_y + yy); \
        ++cnt;                                                                         \
    } while (0)
    ORC_VISIT();
    int d1 = x1 - c[0], d2 = y1 - c[1], d3 = z1 - c[2];
    int m1 = abs(d1), m2 = abs(d2), m3 = abs(d3);
    int a1, a2, a3;                   /* which coordinate is fast / 2nd / 3rd */
    if (m1 >= m2 && m1 >= m3) {
        a1 = 0; a2 = 1; a3 = 2;
    } else if (m2 >= m1 && m2 >= m3) {
        int t = d1; d1 = d2; d2 = t;
        a1 = 1; a2 = 0; a3 = 2;
    } else {
        int t = d1; d1 = d3; d3 = d2; d2 = t;
        a1 = 2; a2 = 0; a3 = 1;
    }
    int inc1 = d1 < 0 ? -1 : 1, inc2 = d2 < 0 ? -1 : 1, inc3 = d3 < 0 ? -1 : 1;
    m1 = abs(d1); m2 = abs(d2); m3 = abs(d3);
    int e1 = 2 * m2 - m1, e2 = 2 * m3 - m1;
    for (int j = 1; j < m1 + 1; ++j) {
        if (e1 > 0) { c[a2] += inc2; e1 -= 2 * m1; }
        if (e2 > 0) { c[a3] += inc3; e2 -= 2 * m1; }
        e1 += 2 * m2; e2 += 2 * m3;
        c[a1] += inc1;
        ORC_VISIT();
    }
#undef ORC_VISIT
    return cnt;
}

static int cmp_double(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int orc_siddon_ray(const double p0[3], const double p1[3], const double origin[3],
                   const double voxel[3], int rv_x, int rv_y, int rv_z,
                   int32_t *vox, double *len, int cap, double min_len) {
    const int N[3] = {rv_x, rv_y, rv_z};
    double dir[3], amin = 0.0, amax = 1.0;
    for (int a = 0; a < 3; ++a) dir[a] = p1[a] - p0[a];
    double L = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (L == 0.0) return 0;
    for (int a = 0; a < 3; ++a) {
        double lo = origin[a], hi = origin[a] + N[a] * voxel[a];
        if (dir[a] == 0.0) {
            if (p0[a] <= lo || p0[a] >= hi) return 0;
            continue;
        }
        double t0 = (lo - p0[a]) / dir[a], t1 = (hi - p0[a]) / dir[a];
        if (t0 > t1) { double t = t0; t0 = t1; t1 = t; }
        if (t0 > amin) amin = t0;
        if (t1 < amax) amax = t1;
    }
    if (!(amin < amax)) return 0;
    int cnt_a = 2 + N[0] + N[1] + N[2] + 3;
    double *al = (double *)malloc(sizeof(double) * cnt_a);
    int k = 0;
    al[k++] = amin; al[k++] = amax;
    for (int a = 0; a < 3; ++a) {
        if (dir[a] == 0.0) continue;
        for (int i = 0; i <= N[a]; ++i) {
            double t = (origin[a] + i * voxel[a] - p0[a]) / dir[a];
            if (t > amin && t < amax) al[k++] = t;
        }
    }
    qsort(al, k, sizeof(double), cmp_double);
    int cnt = 0;
    for (int j = 0; j + 1 < k; ++j) {
        double seg = (al[j + 1] - al[j]) * L;
        if (!(seg > min_len)) continue;
        double mid = 0.5 * (al[j] + al[j + 1]);
        int ic[3];
        for (int a = 0; a < 3; ++a) {
            double pos = p0[a] + mid * dir[a];
            int i = (int)floor((pos - origin[a]) / voxel[a]);
            ic[a] = clampi(i, 0, N[a] - 1);
        }
        if (cnt < cap) {
            vox[cnt] = (int32_t)((int64_t)ic[2] * rv_x * rv_y + (int64_t)ic[0] * rv_y + ic[1]);
            len[cnt] = seg;
        }
        ++cnt;
    }
    free(al);
    return cnt;
}

int64_t orc_siddon_csr(const orc_op *op, int32_t *rowptr, int32_t *col, double *val) {
    int cap = 2 * (op->rv_x + op->rv_y + op->rv_z) + 8;
    int64_t P = op->num_paths;
    int32_t *cnts = (int32_t *)malloc(sizeof(int32_t) * (P > 0 ? P : 1));
#pragma omp parallel
    {
        int32_t *vb = (int32_t *)malloc(sizeof(int32_t) * cap);
        double *lb = (double *)malloc(sizeof(double) * cap);
#pragma omp for schedule(dynamic, 256)
        for (int64_t k = 0; k < P; ++k) {
            const double *a = op->epos + 3 * (int64_t)op->path_e[k];
            const double *b = op->rpos + 3 * (int64_t)op->path_r[k];
            int c = orc_siddon_ray(a, b, op->origin, op->voxel, op->rv_x, op->rv_y, op->rv_z,
                                   vb, lb, cap, 0.0);
            cnts[k] = c;
            if (col) {
                int64_t base = rowptr[k];
                for (int j = 0; j < c; ++j) { col[base + j] = vb[j]; val[base + j] = lb[j]; }
            }
        }
        free(vb); free(lb);
    }
    int64_t nnz = 0;
    if (!col) {
        rowptr[0] = 0;
        for (int64_t k = 0; k < P; ++k) { nnz += cnts[k]; rowptr[k + 1] = (int32_t)nnz; }
    } else {
        nnz = rowptr[P];
    }
    free(cnts);
    return nnz;
}

void orc_csr_transpose(int64_t m, int64_t n, const int32_t *rowptr, const int32_t *col,
                       const double *val, int32_t *colptr, int32_t *row, double *cval) {
    int64_t *next = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t k = 0; k < rowptr[m]; ++k) next[col[k] + 1]++;
    for (int64_t j = 0; j < n; ++j) next[j + 1] += next[j];
    for (int64_t j = 0; j <= n; ++j) colptr[j] = (int32_t)next[j];
    for (int64_t i = 0; i < m; ++i)
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            int64_t dst = next[col[k]]++;
            row[dst] = (int32_t)i;
            cval[dst] = val[k];
        }
    free(next);
}

/* ---- precision instantiations ---- */
#define REAL double
#define RSQRT sqrt
#define RFMA fma
#define SFX(name) name##_f64
#include "tv3_oracle_t.inc"
#undef REAL
#undef RSQRT
#undef RFMA
#undef SFX

#define REAL float
#define RSQRT sqrtf
#define RFMA fmaf
#define SFX(name) name##_f32
#include "tv3_oracle_t.inc"
#undef REAL
#undef RSQRT
#undef RFMA
#undef SFX
