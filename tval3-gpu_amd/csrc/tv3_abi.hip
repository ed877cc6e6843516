// tv3_abi.hip -- the C ABI (include/tval3_mi355x.h): argument validation,
// object lifetimes, dispatch to the kernels, exception -> status mapping.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>

#include "vec.hpp"

using namespace tv3;

namespace {

thread_local std::string g_create_err;

template <typename F>
tv3_status guard(tv3_ctx *ctx, F &&f) {
    // a HIP error left in the thread's last-error slot by a call outside the
    // library (its status already returned to that caller) is not this
    // call's: clear it so launch checks below report only their own errors
    (void)hipGetLastError();
    try {
        f();
        return TV3_OK;
    } catch (const tv3_error &e) {
        (ctx ? ctx->err : g_create_err) = e.what();
        return e.status;
    } catch (const std::bad_alloc &) {
        (ctx ? ctx->err : g_create_err) = "host allocation failed";
        return TV3_E_NOMEM;
    } catch (const std::exception &e) {
        (ctx ? ctx->err : g_create_err) = std::string("internal error: ") + e.what();
        return TV3_E_INTERNAL;
    }
}

void set_device(tv3_ctx *ctx) { TV3_HIP(hipSetDevice(ctx->device)); }

template <typename T>
void upload(DevBuf<T> &d, const T *h, size_t count, hipStream_t s) {
    d.alloc(count > 0 ? count : 1);
    if (count) TV3_HIP(hipMemcpyAsync(d.p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
}

void upload_bytes(DevBuf<uint8_t> &d, const void *h, size_t bytes, hipStream_t s) {
    d.alloc(bytes > 0 ? bytes : 1);
    if (bytes) TV3_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, s));
}

void check_csr_host(int64_t rows, int64_t cols, int64_t nnz, const int32_t *ptr,
                    const int32_t *ind, const char *what) {
    TV3_REQUIRE(ptr && (nnz == 0 || ind), TV3_E_ARG, std::string(what) + ": NULL array");
    TV3_REQUIRE(ptr[0] == 0 && ptr[rows] == nnz, TV3_E_FORMAT,
                std::string(what) + ": pointer array must start at 0 and end at nnz");
    for (int64_t i = 0; i < rows; ++i)
        TV3_REQUIRE(ptr[i + 1] >= ptr[i], TV3_E_FORMAT,
                    std::string(what) + ": pointer array not monotone");
    for (int64_t k = 0; k < nnz; ++k)
        TV3_REQUIRE(ind[k] >= 0 && ind[k] < cols, TV3_E_FORMAT,
                    std::string(what) + ": index out of range");
}

bool finite_host(const void *v, tv3_dtype t, int64_t count) {
    if (t == TV3_F32) {
        const float *p = static_cast<const float *>(v);
        for (int64_t k = 0; k < count; ++k) if (!std::isfinite(p[k])) return false;
    } else if (t == TV3_F64) {
        const double *p = static_cast<const double *>(v);
        for (int64_t k = 0; k < count; ++k) if (!std::isfinite(p[k])) return false;
    } else {
        // F16 / BF16: exponent all ones means Inf or NaN
        const uint16_t *p = static_cast<const uint16_t *>(v);
        const uint16_t mask = t == TV3_F16 ? 0x7C00 : 0x7F80;
        for (int64_t k = 0; k < count; ++k) if ((p[k] & mask) == mask) return false;
    }
    return true;
}

double to_double(const uint8_t *p, tv3_dtype t, int64_t k) {
    switch (t) {
    case TV3_F32: { float v; std::memcpy(&v, p + 4 * k, 4); return v; }
    case TV3_F64: { double v; std::memcpy(&v, p + 8 * k, 8); return v; }
    case TV3_F16: {
        uint16_t h; std::memcpy(&h, p + 2 * k, 2);
        int sgn = h >> 15, e = (h >> 10) & 31, mant = h & 1023;
        double v = e == 0 ? std::ldexp((double)mant, -24)
                          : (e == 31 ? INFINITY : std::ldexp((double)(mant | 1024), e - 25));
        return sgn ? -v : v;
    }
    case TV3_BF16: {
        uint16_t h; std::memcpy(&h, p + 2 * k, 2);
        uint32_t u = (uint32_t)h << 16; float v; std::memcpy(&v, &u, 4); return v;
    }
    }
    return 0.0;
}

void validate_opts(const tv3_opts *o) {
    TV3_REQUIRE(o, TV3_E_ARG, "opts is NULL");
    TV3_REQUIRE(o->skipMulRatio >= 0.0 && o->maxRelativeChange >= 0.0, TV3_E_ARG,
                "skip heuristic parameters must be non-negative");
    TV3_REQUIRE(o->maxit >= 0 && o->maxcnt >= 0, TV3_E_ARG, "maxit/maxcnt must be >= 0");
    TV3_REQUIRE(o->tiles_per_wg >= 0, TV3_E_ARG, "tiles_per_wg must be >= 0");
    TV3_REQUIRE(o->mu > 0 && o->mu0 > 0 && o->beta > 0 && o->beta0 > 0, TV3_E_ARG,
                "mu, mu0, beta, beta0 must be positive");
    TV3_REQUIRE(o->tv_mode == TV3_TV_ANISO || o->tv_mode == TV3_TV_ISO, TV3_E_ARG, "bad tv_mode");
    TV3_REQUIRE(o->solver == TV3_SOLVER_PLAIN || o->solver == TV3_SOLVER_FUSED, TV3_E_ARG,
                "bad solver");
}

int64_t dims_n(const int32_t dims[3]) {
    TV3_REQUIRE(dims && dims[0] > 0 && dims[1] > 0 && dims[2] > 0, TV3_E_ARG,
                "dims must be three positive sizes");
    return (int64_t)dims[0] * dims[1] * dims[2];
}

}  // namespace

namespace tv3 {

void op_apply(tv3_op *op, bool trans, const float *x, float *y, hipStream_t s, bool acc64) {
    switch (op->kind) {
    case OpKind::CSR: spmv_csr(trans ? op->AT : op->A, x, y, s, 0, acc64); break;
    case OpKind::DENSE: dense_gemv(op, trans, x, y, s, acc64); break;
    case OpKind::DYNAMIC: dynamic_apply(op, trans, x, y, s, acc64); break;
    case OpKind::HYBRID:
        if (trans) spmv_csr(op->AT, x, y, s, 0, acc64);
        else dynamic_apply(op, false, x, y, s, acc64);
        break;
    }
}

void op_apply(tv3_op *op, bool trans, const double *x, double *y, hipStream_t s, bool) {
    switch (op->kind) {
    case OpKind::CSR: spmv_csr(trans ? op->AT : op->A, x, y, s); break;
    case OpKind::DENSE: dense_gemv(op, trans, x, y, s); break;
    case OpKind::DYNAMIC: dynamic_apply(op, trans, x, y, s); break;
    case OpKind::HYBRID:
        TV3_REQUIRE(op->AT.vtype == TV3_F32 || op->AT.vtype == TV3_F64, TV3_E_FORMAT,
                    "FP64 products of a hybrid operator need F32 weights");
        if (trans) spmv_csr(op->AT, x, y, s);
        else dynamic_apply(op, false, x, y, s);
        break;
    }
}

}  // namespace tv3

template <typename T>
static tv3_status solve_host(tv3_ctx *ctx, tv3_op *op, const tv3_opts *opts,
                             const int32_t dims[3], const T *b, T *U, const T *Ut,
                             tv3_info *info, tv3_history *hist) {
    if (!ctx) return TV3_E_ARG;
    auto t0 = std::chrono::steady_clock::now();
    tv3_status st = guard(ctx, [&] {
        validate_opts(opts);
        TV3_REQUIRE(op && b && U && info, TV3_E_ARG, "NULL argument");
        int64_t n = dims_n(dims);
        TV3_REQUIRE(n == op->n, TV3_E_SHAPE, "length of U must equal the columns of A");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        DevBuf<T> bd(op->m), Ud(n), Utd;
        TV3_HIP(hipMemcpyAsync(bd.p, b, sizeof(T) * op->m, hipMemcpyHostToDevice, s));
        if (opts->warm_start)
            TV3_HIP(hipMemcpyAsync(Ud.p, U, sizeof(T) * n, hipMemcpyHostToDevice, s));
        if (Ut) {
            Utd.alloc(n);
            TV3_HIP(hipMemcpyAsync(Utd.p, Ut, sizeof(T) * n, hipMemcpyHostToDevice, s));
        }
        std::memset(info, 0, sizeof(*info));
        if (hist) hist->count = 0;
        if constexpr (std::is_same<T, float>::value) {
            if (opts->solver == TV3_SOLVER_FUSED)
                solve_fused(ctx, op, *opts, dims, bd.p, Ud.p, Ut ? Utd.p : nullptr, info, hist);
            else
                solve_plain<float>(ctx, op, *opts, dims, bd.p, Ud.p, Ut ? Utd.p : nullptr, info,
                                   hist);
        } else {
            solve_plain<double>(ctx, op, *opts, dims, bd.p, Ud.p, Ut ? Utd.p : nullptr, info,
                                hist);
        }
        TV3_HIP(hipMemcpyAsync(U, Ud.p, sizeof(T) * n, hipMemcpyDeviceToHost, s));
        TV3_HIP(hipStreamSynchronize(s));
    });
    if (info) {
        info->outer_secs =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (st == TV3_OK && info->status != TV3_OK) {
            ctx->err = "iterates became non-finite";
            return (tv3_status)info->status;
        }
    }
    return st;
}

extern "C" {

int32_t tv3_abi_version(void) { return TV3_ABI_VERSION; }

void tv3_opts_default(tv3_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    // 2d/tval3_types.h:29-34
    o->mu = 256.0; o->mu0 = 256.0; o->beta = 32.0; o->beta0 = 32.0;
    o->rate_cnt = 2.0; o->tol = 1e-6; o->tol_inn = 1e-3;
    o->maxcnt = 10; o->maxit = 1025;
    o->c = 1e-5; o->gamma = 0.6; o->gam = 0.9995; o->rate_gam = 0.9;
    o->isreal = 1; o->nonneg = 1;
    o->numInitIterations = 0; o->skipMulRatio = 0.0; o->maxRelativeChange = 0.0;
    o->tv_mode = TV3_TV_ANISO;
    o->solver = TV3_SOLVER_PLAIN;
    o->warm_start = 0;
    o->graph_iters = 0;
    o->verbose = 0;
}

tv3_status tv3_ctx_create(int32_t device, tv3_ctx **out) {
    return guard(nullptr, [&] {
        TV3_REQUIRE(out, TV3_E_ARG, "out is NULL");
        *out = nullptr;
        int count = 0;
        TV3_HIP(hipGetDeviceCount(&count));
        TV3_REQUIRE(count > 0, TV3_E_HIP, "no HIP device");
        TV3_REQUIRE(device >= 0 && device < count, TV3_E_ARG, "device index out of range");
        TV3_HIP(hipSetDevice(device));
        std::unique_ptr<tv3_ctx> c(new tv3_ctx());
        c->device = device;
        TV3_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        TV3_HIP(hipGetDeviceProperties(&prop, device));
        c->num_cus = prop.multiProcessorCount;
        c->lds_per_block = prop.sharedMemPerBlock;
        noalloc_setup(device);           // device current: its scratch array's address
        *out = c.release();
    });
}

void tv3_ctx_destroy(tv3_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->batch_session.reset();
    ctx->red_partials.release();
    ctx->scratch.release();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *tv3_last_error(const tv3_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

void *tv3_ctx_stream(tv3_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

// host rows (CSR of A, or its CSC as the CSR of A^T) to the device, with the
// default launch shape and 2048-entry row blocks
static void upload_rows(tv3_ctx *ctx, SparseRows &R, int64_t rows, int64_t cols, int64_t nnz,
                        const int32_t *ptr, const int32_t *ind, const void *val, tv3_dtype vt) {
    hipStream_t s = ctx->stream;
    R.rows = rows; R.cols = cols; R.nnz = nnz; R.vtype = vt;
    upload(R.ptr, ptr, rows + 1, s);
    upload(R.ind, ind, nnz, s);
    upload_bytes(R.val, val, nnz * dtype_size(vt), s);
    R.lanes = default_lanes(rows, nnz);
    build_row_blocks(ctx, R, 2048);
}

tv3_status tv3_op_csr_create(tv3_ctx *ctx, int64_t m, int64_t n, int64_t nnz,
                             const int32_t *rowptr, const int32_t *col, const void *val,
                             tv3_dtype vt, const int32_t *cptr, const int32_t *crow,
                             const void *cval, tv3_op **out) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        TV3_REQUIRE(out, TV3_E_ARG, "out is NULL");
        *out = nullptr;
        TV3_REQUIRE(m > 0 && n > 0 && nnz >= 0 && nnz <= INT32_MAX, TV3_E_SHAPE,
                    "bad sparse sizes");
        TV3_REQUIRE(dtype_size(vt) > 0, TV3_E_ARG, "bad value type");
        TV3_REQUIRE(nnz == 0 || val, TV3_E_ARG, "val is NULL");
        check_csr_host(m, n, nnz, rowptr, col, "CSR");
        TV3_REQUIRE(finite_host(val, vt, nnz), TV3_E_NONFINITE, "A holds NaN or Inf");
        const bool has_csc = cptr != nullptr;
        if (has_csc) {
            TV3_REQUIRE(nnz == 0 || (crow && cval), TV3_E_ARG, "partial CSC arrays");
            check_csr_host(n, m, nnz, cptr, crow, "CSC");
        }
        set_device(ctx);
        std::unique_ptr<tv3_op> op(new tv3_op());
        op->ctx = ctx;
        op->device = ctx->device;
        op->kind = OpKind::CSR;
        op->m = m; op->n = n;
        upload_rows(ctx, op->A, m, n, nnz, rowptr, col, val, vt);
        if (has_csc) upload_rows(ctx, op->AT, n, m, nnz, cptr, crow, cval, vt);
        else build_transpose(ctx, op->A, op->AT);
        TV3_HIP(hipStreamSynchronize(ctx->stream));
        *out = op.release();
    });
}

tv3_status tv3_op_dense_create(tv3_ctx *ctx, int64_t m, int64_t n, const void *A, int64_t lda,
                               tv3_dtype vt, tv3_op **out) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        TV3_REQUIRE(out, TV3_E_ARG, "out is NULL");
        *out = nullptr;
        TV3_REQUIRE(m > 0 && n > 0 && lda >= m, TV3_E_SHAPE, "bad dense sizes (lda >= m)");
        TV3_REQUIRE(A, TV3_E_ARG, "A is NULL");
        TV3_REQUIRE(dtype_size(vt) > 0, TV3_E_ARG, "bad value type");
        TV3_REQUIRE(finite_host(A, vt, lda * n), TV3_E_NONFINITE, "A holds NaN or Inf");
        set_device(ctx);
        std::unique_ptr<tv3_op> op(new tv3_op());
        op->ctx = ctx;
        op->device = ctx->device;
        op->kind = OpKind::DENSE;
        op->m = m; op->n = n; op->lda = lda; op->dtype = vt;
        upload_bytes(op->dense, A, (size_t)lda * n * dtype_size(vt), ctx->stream);
        TV3_HIP(hipStreamSynchronize(ctx->stream));
        *out = op.release();
    });
}

static void fill_geometry(tv3_ctx *ctx, const tv3_geometry *g, tv3_weight_mode mode,
                          tv3_op *op) {
    TV3_REQUIRE(g, TV3_E_ARG, "geometry is NULL");
    TV3_REQUIRE(g->num_emitters > 0 && g->num_receivers > 0 && g->num_mps > 0, TV3_E_SHAPE,
                "geometry needs emitters, receivers and measurement positions");
    TV3_REQUIRE(g->rv_x > 0 && g->rv_y > 0 && g->rv_z > 0, TV3_E_SHAPE, "bad volume size");
    TV3_REQUIRE(mode == TV3_WEIGHT_BRESENHAM || mode == TV3_WEIGHT_SIDDON, TV3_E_ARG,
                "bad weight mode");
    const int E = g->num_emitters, R = g->num_receivers, mps = g->num_mps;
    const int64_t ne = (int64_t)E * mps, nr = (int64_t)R * mps;
    DynGeom &G = op->geo;
    G.E = E; G.R = R; G.mps = mps;
    G.rv_x = g->rv_x; G.rv_y = g->rv_y; G.rv_z = g->rv_z;
    G.scale = g->scale_factor;
    G.mode = mode;
    hipStream_t s = ctx->stream;
    if (mode == TV3_WEIGHT_BRESENHAM) {
        TV3_REQUIRE(g->x_emitters && g->y_emitters && g->z_emitters && g->x_receivers &&
                        g->y_receivers && g->z_receivers,
                    TV3_E_ARG, "Bresenham mode needs integer transducer coordinates");
        std::vector<int32_t> ec(3 * ne), rc(3 * nr);
        for (int64_t k = 0; k < ne; ++k) {
            ec[3 * k] = g->x_emitters[k]; ec[3 * k + 1] = g->y_emitters[k];
            ec[3 * k + 2] = g->z_emitters[k];
        }
        for (int64_t k = 0; k < nr; ++k) {
            rc[3 * k] = g->x_receivers[k]; rc[3 * k + 1] = g->y_receivers[k];
            rc[3 * k + 2] = g->z_receivers[k];
        }
        upload(G.ecoord, ec.data(), ec.size(), s);
        upload(G.rcoord, rc.data(), rc.size(), s);
        TV3_HIP(hipStreamSynchronize(s));
        G.epos_h.assign(ec.begin(), ec.end());
        G.rpos_h.assign(rc.begin(), rc.end());
    } else {
        TV3_REQUIRE(g->pos_emitters && g->pos_receivers, TV3_E_ARG,
                    "Siddon mode needs transducer positions");
        TV3_REQUIRE(g->voxel[0] > 0 && g->voxel[1] > 0 && g->voxel[2] > 0, TV3_E_ARG,
                    "voxel sizes must be positive");
        for (int64_t k = 0; k < 3 * ne; ++k)
            TV3_REQUIRE(std::isfinite(g->pos_emitters[k]), TV3_E_NONFINITE, "emitter position");
        for (int64_t k = 0; k < 3 * nr; ++k)
            TV3_REQUIRE(std::isfinite(g->pos_receivers[k]), TV3_E_NONFINITE, "receiver position");
        upload(G.epos, g->pos_emitters, 3 * ne, s);
        upload(G.rpos, g->pos_receivers, 3 * nr, s);
        G.epos_h.assign(g->pos_emitters, g->pos_emitters + 3 * ne);
        G.rpos_h.assign(g->pos_receivers, g->pos_receivers + 3 * nr);
        for (int a = 0; a < 3; ++a) { G.origin[a] = g->origin[a]; G.voxel[a] = g->voxel[a]; }
        TV3_HIP(hipStreamSynchronize(s));
    }
    // compacted ray list, 3d/mat_vec_mul_dynamic.h:501-545
    std::vector<int32_t> path;
    const int64_t per = (int64_t)E * R, tot = per * mps;
    path.reserve(2 * tot);
    for (int64_t i = 0; i < tot; ++i) {
        if (g->use_path && !g->use_path[i % per]) continue;
        path.push_back((int32_t)(i / R));
        path.push_back((int32_t)((i / per) * R + i % R));
    }
    G.num_paths = (int64_t)path.size() / 2;
    if (mode == TV3_WEIGHT_BRESENHAM) {
        int32_t mw = 1;
        for (int64_t k = 0; k < G.num_paths; ++k) {
            int64_t e = path[2 * k], r = path[2 * k + 1];
            int32_t dx = std::abs(g->x_receivers[r] - g->x_emitters[e]);
            int32_t dy = std::abs(g->y_receivers[r] - g->y_emitters[e]);
            int32_t dz = std::abs(g->z_receivers[r] - g->z_emitters[e]);
            mw = std::max(mw, std::max(dx, std::max(dy, dz)) + 1);
        }
        G.max_walk = mw;
    }
    TV3_REQUIRE(G.num_paths > 0, TV3_E_SHAPE, "use_path masks every ray");
    upload(G.path, path.data(), path.size(), s);
    TV3_HIP(hipStreamSynchronize(s));
    G.path_h = std::move(path);
    op->m = G.num_paths;
    op->n = (int64_t)g->rv_x * g->rv_y * g->rv_z;
}

tv3_status tv3_op_dynamic_create(tv3_ctx *ctx, const tv3_geometry *geom, tv3_weight_mode mode,
                                 tv3_op **out) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        TV3_REQUIRE(out, TV3_E_ARG, "out is NULL");
        *out = nullptr;
        set_device(ctx);
        std::unique_ptr<tv3_op> op(new tv3_op());
        op->ctx = ctx;
        op->device = ctx->device;
        op->kind = OpKind::DYNAMIC;
        fill_geometry(ctx, geom, mode, op.get());
        op->fused = build_dyn_layout(op.get());   // Siddon, dims multiples of 4: tile products
        *out = op.release();
    });
}

tv3_status tv3_op_hybrid_create(tv3_ctx *ctx, const tv3_geometry *geom, tv3_weight_mode mode,
                                int64_t nnz, const int32_t *rowptr, const int32_t *col,
                                const void *val, tv3_dtype vt, const int32_t *cptr,
                                const int32_t *crow, const void *cval, tv3_op **out) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        TV3_REQUIRE(out, TV3_E_ARG, "out is NULL");
        *out = nullptr;
        TV3_REQUIRE(dtype_size(vt) > 0, TV3_E_ARG, "bad value type");
        set_device(ctx);
        std::unique_ptr<tv3_op> op(new tv3_op());
        op->ctx = ctx;
        op->device = ctx->device;
        op->kind = OpKind::HYBRID;
        fill_geometry(ctx, geom, mode, op.get());
        const int64_t m = op->m, n = op->n;
        SparseRows A;                        // freed after the stream has drained
        if (cptr) {                          // the caller's CSC is A^T's rows
            TV3_REQUIRE(nnz >= 0 && nnz <= INT32_MAX, TV3_E_SHAPE, "bad sparse sizes");
            TV3_REQUIRE(nnz == 0 || (crow && cval), TV3_E_ARG, "partial CSC arrays");
            check_csr_host(n, m, nnz, cptr, crow, "CSC");
            TV3_REQUIRE(finite_host(cval, vt, nnz), TV3_E_NONFINITE, "A holds NaN or Inf");
            upload_rows(ctx, op->AT, n, m, nnz, cptr, crow, cval, vt);
        } else if (rowptr) {                 // the caller's CSR, transposed on the device
            TV3_REQUIRE(nnz >= 0 && nnz <= INT32_MAX, TV3_E_SHAPE, "bad sparse sizes");
            TV3_REQUIRE(nnz == 0 || (col && val), TV3_E_ARG, "partial CSR arrays");
            check_csr_host(m, n, nnz, rowptr, col, "CSR");
            TV3_REQUIRE(finite_host(val, vt, nnz), TV3_E_NONFINITE, "A holds NaN or Inf");
            upload_rows(ctx, A, m, n, nnz, rowptr, col, val, vt);
            build_transpose(ctx, A, op->AT);
        } else {                             // no matrix given: the geometry's own
            dynamic_materialize(op.get(), vt, A);
            build_transpose(ctx, A, op->AT);
        }
        TV3_HIP(hipStreamSynchronize(ctx->stream));
        *out = op.release();
    });
}

tv3_status tv3_op_materialize(tv3_op *dyn, tv3_dtype vt, tv3_op **out) {
    if (!dyn) return TV3_E_ARG;
    tv3_ctx *ctx = dyn->ctx;
    return guard(ctx, [&] {
        TV3_REQUIRE(out, TV3_E_ARG, "out is NULL");
        *out = nullptr;
        TV3_REQUIRE(dyn->kind == OpKind::DYNAMIC || dyn->kind == OpKind::HYBRID, TV3_E_ARG,
                    "materialize needs a dynamic operator");
        set_device(ctx);
        std::unique_ptr<tv3_op> op(new tv3_op());
        op->ctx = ctx;
        op->device = ctx->device;
        op->kind = OpKind::CSR;
        op->m = dyn->m; op->n = dyn->n;
        dynamic_materialize(dyn, vt, op->A);
        build_transpose(ctx, op->A, op->AT);
        op->ray_order = geometry_ray_order(dyn->geo);
        *out = op.release();
    });
}

void tv3_op_destroy(tv3_op *op) {
    if (!op) return;
    (void)hipSetDevice(op->device);
    (void)hipDeviceSynchronize();
    delete op;
}

tv3_status tv3_op_info(const tv3_op *op, int64_t *m, int64_t *n, int64_t *nnz) {
    if (!op) return TV3_E_ARG;
    if (m) *m = op->m;
    if (n) *n = op->n;
    if (nnz) {
        switch (op->kind) {
        case OpKind::CSR: *nnz = op->A.nnz; break;
        case OpKind::HYBRID: *nnz = op->AT.nnz; break;
        case OpKind::DENSE: *nnz = op->m * op->n; break;
        case OpKind::DYNAMIC: *nnz = -1; break;
        }
    }
    return TV3_OK;
}

tv3_status tv3_op_csr_download(tv3_op *op, int32_t *rowptr, int32_t *col, double *val) {
    if (!op) return TV3_E_ARG;
    return guard(op->ctx, [&] {
        TV3_REQUIRE(op->kind == OpKind::CSR, TV3_E_ARG, "not a CSR operator");
        set_device(op->ctx);
        const SparseRows &A = op->A;
        hipStream_t s = op->ctx->stream;
        if (rowptr)
            TV3_HIP(hipMemcpyAsync(rowptr, A.ptr.p, sizeof(int32_t) * (A.rows + 1),
                                   hipMemcpyDeviceToHost, s));
        if (col && A.nnz)
            TV3_HIP(hipMemcpyAsync(col, A.ind.p, sizeof(int32_t) * A.nnz, hipMemcpyDeviceToHost, s));
        std::vector<uint8_t> raw;
        if (val && A.nnz) {
            raw.resize(A.nnz * dtype_size(A.vtype));
            TV3_HIP(hipMemcpyAsync(raw.data(), A.val.p, raw.size(), hipMemcpyDeviceToHost, s));
        }
        TV3_HIP(hipStreamSynchronize(s));
        if (val)
            for (int64_t k = 0; k < A.nnz; ++k) val[k] = to_double(raw.data(), A.vtype, k);
    });
}

tv3_status tv3_matvec(tv3_op *op, tv3_trans trans, const float *x, float *y, void *stream) {
    if (!op) return TV3_E_ARG;
    return guard(op->ctx, [&] {
        TV3_REQUIRE(x && y, TV3_E_ARG, "x or y is NULL");
        TV3_REQUIRE(trans == TV3_N || trans == TV3_T, TV3_E_ARG, "bad trans");
        set_device(op->ctx);
        hipStream_t s = (hipStream_t)stream;   // NULL: the null stream
        op_apply(op, trans == TV3_T, x, y, s);
        TV3_HIP(hipGetLastError());
    });
}

tv3_status tv3_matvec_f64(tv3_op *op, tv3_trans trans, const double *x, double *y, void *stream) {
    if (!op) return TV3_E_ARG;
    return guard(op->ctx, [&] {
        TV3_REQUIRE(x && y, TV3_E_ARG, "x or y is NULL");
        TV3_REQUIRE(trans == TV3_N || trans == TV3_T, TV3_E_ARG, "bad trans");
        set_device(op->ctx);
        hipStream_t s = (hipStream_t)stream;   // NULL: the null stream
        op_apply(op, trans == TV3_T, x, y, s);
        TV3_HIP(hipGetLastError());
    });
}

tv3_status tv3_matvec_multi(tv3_op *op, tv3_trans trans, int32_t nrhs, const float *X,
                            int64_t ldx, float *Y, int64_t ldy, void *stream) {
    if (!op) return TV3_E_ARG;
    return guard(op->ctx, [&] {
        TV3_REQUIRE(trans == TV3_N || trans == TV3_T, TV3_E_ARG, "bad trans");
        TV3_REQUIRE(nrhs >= 0 && (nrhs == 0 || (X && Y)), TV3_E_ARG, "NULL argument");
        const bool T = trans == TV3_T;
        const int64_t xl = T ? op->m : op->n, yl = T ? op->n : op->m;
        TV3_REQUIRE(ldx >= xl && ldy >= yl, TV3_E_ARG, "leading dimension below the vector length");
        set_device(op->ctx);
        hipStream_t s = (hipStream_t)stream;
        const int64_t R = yl, K = xl;
        const bool mfma = op->kind == OpKind::DENSE && op->dtype == TV3_BF16 && R % 256 == 0 &&
                          K % 64 == 0 && op->lda % 8 == 0 && ldx % 4 == 0;
        if (!mfma) {
            for (int32_t j = 0; j < nrhs; ++j) op_apply(op, T, X + j * ldx, Y + j * ldy, s);
            TV3_HIP(hipGetLastError());
            return;
        }
        // K splits: enough workgroups to fill the device, K a multiple of 64 splits
        int splits = 1;
        while ((R / gemm_rhs_rows()) * splits < 1024 && gemm_rhs_shape_ok(R, K, splits * 2))
            splits *= 2;
        const size_t need = gemm_rhs_scratch(R, splits) * sizeof(float);
        if (op->ctx->scratch.bytes() < need) op->ctx->scratch.alloc(need);
        float *part = reinterpret_cast<float *>(op->ctx->scratch.p);
        dense_operand_order(op, s);
        const uint16_t *M = (const uint16_t *)(T ? op->dense_oT.p : op->dense_oN.p);
        for (int32_t c = 0; c < nrhs; c += 64)
            gemm_rhs(M, R, K, X + c * ldx, ldx, std::min(64, nrhs - c), Y + c * ldy, ldy, nullptr,
                     splits, part, s);
        TV3_HIP(hipGetLastError());
    });
}

static void solve_device_impl(tv3_ctx *ctx, tv3_op *op, const tv3_opts *opts,
                              const int32_t dims[3], const float *b, float *U, const float *Ut,
                              tv3_info *info, tv3_history *hist) {
    validate_opts(opts);
    TV3_REQUIRE(op && b && U && info, TV3_E_ARG, "NULL argument");
    TV3_REQUIRE(dims_n(dims) == op->n, TV3_E_SHAPE, "length of U must equal the columns of A");
    std::memset(info, 0, sizeof(*info));
    if (hist) hist->count = 0;
    set_device(ctx);
    if (opts->solver == TV3_SOLVER_FUSED)
        solve_fused(ctx, op, *opts, dims, b, U, Ut, info, hist);
    else
        solve_plain<float>(ctx, op, *opts, dims, b, U, Ut, info, hist);
}

tv3_status tv3_solve_device(tv3_ctx *ctx, tv3_op *op, const tv3_opts *opts, const int32_t dims[3],
                            const float *b, float *U, const float *Ut, tv3_info *info,
                            tv3_history *hist) {
    if (!ctx) return TV3_E_ARG;
    tv3_status st = guard(ctx, [&] { solve_device_impl(ctx, op, opts, dims, b, U, Ut, info, hist); });
    if (st == TV3_OK && info && info->status != TV3_OK) {
        ctx->err = "iterates became non-finite";
        return (tv3_status)info->status;
    }
    return st;
}

tv3_status tv3_solve(tv3_ctx *ctx, tv3_op *op, const tv3_opts *opts, const int32_t dims[3],
                     const float *b, float *U, const float *Ut, tv3_info *info,
                     tv3_history *hist) {
    return solve_host<float>(ctx, op, opts, dims, b, U, Ut, info, hist);
}

tv3_status tv3_solve_f64(tv3_ctx *ctx, tv3_op *op, const tv3_opts *opts, const int32_t dims[3],
                         const double *b, double *U, const double *Ut, tv3_info *info,
                         tv3_history *hist) {
    return solve_host<double>(ctx, op, opts, dims, b, U, Ut, info, hist);
}

tv3_status tv3_solve_batched(tv3_ctx *ctx, int32_t nprob, tv3_op *const *ops,
                             const int32_t *dims, const tv3_opts *opts,
                             const float *const *b_dev, float *const *U_dev,
                             const float *const *Ut_dev, tv3_info *infos, tv3_history *hists,
                             tv3_batch_info *binfo) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        validate_opts(opts);
        TV3_REQUIRE(nprob > 0 && ops && dims && b_dev && U_dev && infos, TV3_E_ARG,
                    "NULL argument or empty batch");
        for (int32_t k = 0; k < nprob; ++k) {
            TV3_REQUIRE(ops[k] && b_dev[k] && U_dev[k], TV3_E_ARG, "NULL entry in the batch");
            TV3_REQUIRE(ops[k]->ctx == ctx, TV3_E_ARG, "operator of another context");
        }
        if (binfo && (binfo->lanes_N != 0 || binfo->lanes_T != 0))
            for (int32_t L : {binfo->lanes_N, binfo->lanes_T})
                TV3_REQUIRE((L >= 2 && L <= 64 && (L & (L - 1)) == 0) || L == -1024 ||
                                L == -2048 || L == -4096,
                            TV3_E_ARG,
                            "launch shapes: lanes per row (a power of two in [2, 64]) or "
                            "-1024, -2048, -4096 (row blocks of that many entries)");
        set_device(ctx);
        solve_batched(ctx, nprob, ops, dims, *opts, b_dev, U_dev, Ut_dev, infos, hists, binfo);
    });
}

tv3_status tv3_D(tv3_ctx *ctx, const int32_t dims[3], const float *U, float *Ux, float *Uy,
                 float *Uz, void *stream) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        dims_n(dims);
        TV3_REQUIRE(U && Ux && Uy && Uz, TV3_E_ARG, "NULL argument");
        set_device(ctx);
        launch_D<float>(dims, U, Ux, Uy, Uz, (hipStream_t)stream);
        TV3_HIP(hipGetLastError());
    });
}

tv3_status tv3_Dt(tv3_ctx *ctx, const int32_t dims[3], const float *X, const float *Y,
                  const float *Z, float *res, void *stream) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        dims_n(dims);
        TV3_REQUIRE(X && Y && Z && res, TV3_E_ARG, "NULL argument");
        set_device(ctx);
        launch_Dt<float>(dims, X, Y, Z, res, (hipStream_t)stream);
        TV3_HIP(hipGetLastError());
    });
}

tv3_status tv3_shrink(tv3_ctx *ctx, int64_t n, tv3_tv_mode mode, const float *Ux, const float *Uy,
                      const float *Uz, const float *sx, const float *sy, const float *sz,
                      double beta, float *Wx, float *Wy, float *Wz, double *lam1) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        TV3_REQUIRE(n > 0 && Ux && Uy && Uz && sx && sy && sz && Wx && Wy && Wz, TV3_E_ARG,
                    "NULL argument");
        TV3_REQUIRE(beta > 0, TV3_E_ARG, "beta must be positive");
        set_device(ctx);
        double l = shrink_host<float>(ctx, n, mode == TV3_TV_ISO, Ux, Uy, Uz, sx, sy, sz, beta, Wx,
                                      Wy, Wz, ctx->stream);
        if (lam1) *lam1 = l;
    });
}

tv3_status tv3_D_f64(tv3_ctx *ctx, const int32_t dims[3], const double *U, double *Ux,
                     double *Uy, double *Uz, void *stream) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        dims_n(dims);
        TV3_REQUIRE(U && Ux && Uy && Uz, TV3_E_ARG, "NULL argument");
        set_device(ctx);
        launch_D<double>(dims, U, Ux, Uy, Uz, (hipStream_t)stream);
        TV3_HIP(hipGetLastError());
    });
}

tv3_status tv3_Dt_f64(tv3_ctx *ctx, const int32_t dims[3], const double *X, const double *Y,
                      const double *Z, double *res, void *stream) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        dims_n(dims);
        TV3_REQUIRE(X && Y && Z && res, TV3_E_ARG, "NULL argument");
        set_device(ctx);
        launch_Dt<double>(dims, X, Y, Z, res, (hipStream_t)stream);
        TV3_HIP(hipGetLastError());
    });
}

tv3_status tv3_shrink_apply(tv3_ctx *ctx, int64_t n, tv3_tv_mode mode, tv3_dtype dtype,
                            const void *Vx, const void *Vy, const void *Vz, double t, void *Wx,
                            void *Wy, void *Wz, void *s, const uint8_t *mask_in,
                            uint8_t *mask_out, void *stream) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        TV3_REQUIRE(n >= 0, TV3_E_ARG, "n must be non-negative");
        TV3_REQUIRE(n == 0 || (Vx && Vy && Vz && Wx && Wy && Wz), TV3_E_ARG, "NULL argument");
        TV3_REQUIRE(mode == TV3_TV_ANISO || mode == TV3_TV_ISO, TV3_E_ARG, "bad tv mode");
        TV3_REQUIRE(t >= 0 && std::isfinite(t), TV3_E_ARG, "threshold must be finite and >= 0");
        set_device(ctx);
        hipStream_t st = (hipStream_t)stream;
        int iso = mode == TV3_TV_ISO;
        if (dtype == TV3_F32)
            shrink_apply<float>(n, iso, t, (const float *)Vx, (const float *)Vy,
                                (const float *)Vz, (float *)Wx, (float *)Wy, (float *)Wz,
                                (float *)s, mask_in, mask_out, st);
        else if (dtype == TV3_F64)
            shrink_apply<double>(n, iso, t, (const double *)Vx, (const double *)Vy,
                                 (const double *)Vz, (double *)Wx, (double *)Wy, (double *)Wz,
                                 (double *)s, mask_in, mask_out, st);
        else
            TV3_REQUIRE(false, TV3_E_ARG, "dtype must be TV3_F32 or TV3_F64");
        TV3_HIP(hipGetLastError());
    });
}

tv3_status tv3_shrink_adjoint(tv3_ctx *ctx, int64_t n, tv3_tv_mode mode, tv3_dtype dtype,
                              const void *Vx, const void *Vy, const void *Vz, double t,
                              const void *gWx, const void *gWy, const void *gWz, const void *gs,
                              void *gVx, void *gVy, void *gVz, void *stream) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        TV3_REQUIRE(n >= 0, TV3_E_ARG, "n must be non-negative");
        TV3_REQUIRE(n == 0 || (Vx && Vy && Vz && gWx && gWy && gWz && gVx && gVy && gVz),
                    TV3_E_ARG, "NULL argument");
        TV3_REQUIRE(mode == TV3_TV_ANISO || mode == TV3_TV_ISO, TV3_E_ARG, "bad tv mode");
        TV3_REQUIRE(t >= 0 && std::isfinite(t), TV3_E_ARG, "threshold must be finite and >= 0");
        set_device(ctx);
        hipStream_t st = (hipStream_t)stream;
        int iso = mode == TV3_TV_ISO;
        if (dtype == TV3_F32)
            shrink_adjoint<float>(n, iso, t, (const float *)Vx, (const float *)Vy,
                                  (const float *)Vz, (const float *)gWx, (const float *)gWy,
                                  (const float *)gWz, (const float *)gs, (float *)gVx,
                                  (float *)gVy, (float *)gVz, st);
        else if (dtype == TV3_F64)
            shrink_adjoint<double>(n, iso, t, (const double *)Vx, (const double *)Vy,
                                   (const double *)Vz, (const double *)gWx, (const double *)gWy,
                                   (const double *)gWz, (const double *)gs, (double *)gVx,
                                   (double *)gVy, (double *)gVz, st);
        else
            TV3_REQUIRE(false, TV3_E_ARG, "dtype must be TV3_F32 or TV3_F64");
        TV3_HIP(hipGetLastError());
    });
}

static float time_spmv(tv3_ctx *ctx, const SparseRows &S, int L, const float *x, float *y) {
    hipStream_t s = ctx->stream;
    hipEvent_t a, b;
    TV3_HIP(hipEventCreate(&a)); TV3_HIP(hipEventCreate(&b));
    spmv_csr(S, x, y, s, L);               // warm up
    TV3_HIP(hipEventRecord(a, s));
    for (int r = 0; r < 5; ++r) spmv_csr(S, x, y, s, L);
    TV3_HIP(hipEventRecord(b, s));
    TV3_HIP(hipEventSynchronize(b));
    float ms = 0;
    TV3_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a); (void)hipEventDestroy(b);
    return ms / 5;
}

tv3_status tv3_op_autotune(tv3_op *op, int32_t *lanes_N, int32_t *lanes_T) {
    if (!op) return TV3_E_ARG;
    return guard(op->ctx, [&] {
        TV3_REQUIRE(op->kind == OpKind::CSR || op->kind == OpKind::HYBRID, TV3_E_ARG,
                    "autotune needs a sparse operator");
        set_device(op->ctx);
        DevBuf<float> x(op->n), y(op->m > op->n ? op->m : op->n);
        TV3_HIP(hipMemsetAsync(x.p, 0, x.bytes(), op->ctx->stream));
        DevBuf<float> xm(op->m);
        TV3_HIP(hipMemsetAsync(xm.p, 0, xm.bytes(), op->ctx->stream));
        // candidates: row-blocked stream kernel with 1024/2048/4096-entry
        // blocks (reported as -NB) and the row-group kernel with L lanes
        auto best = [&](SparseRows &S, const float *xx) {
            int choice = 0;
            float bt = 1e30f;
            for (int nb : {1024, 2048, 4096}) {
                build_row_blocks(op->ctx, S, nb);
                float t = time_spmv(op->ctx, S, 0, xx, y.p);
                if (t < bt) { bt = t; choice = -nb; }
            }
            for (int L : {4, 8, 16, 32, 64}) {
                float t = time_spmv(op->ctx, S, L, xx, y.p);
                if (t < bt) { bt = t; choice = L; }
            }
            if (choice < 0) {
                build_row_blocks(op->ctx, S, -choice);
            } else {
                S.lanes = choice;
                S.blk_nnz = 0;
            }
            return choice;
        };
        int ln = op->kind == OpKind::CSR ? best(op->A, x.p) : 0;
        int lt = best(op->AT, xm.p);
        if (lanes_N) *lanes_N = ln;
        if (lanes_T) *lanes_T = lt;
    });
}

static void fill_layout_info(const FusedLayout &F, tv3_layout_info *info) {
    if (!info) return;
    std::memset(info, 0, sizeof(*info));
    auto code = [](const SparseRows &S) { return S.blk_nnz ? -S.blk_nnz : S.lanes; };
    info->brick = F.B;
    info->panels = (F.panels_A ? 1 : 0) | (F.panels_T ? 2 : 0) |
                   (F.panels_A && F.unit_A == 1024 ? 4 : 0) |
                   (F.panels_A && F.unit_A == 2048 ? 16 : 0) |
                   (F.panels_A && F.unit_A == 4096 ? 32 : 0) |
                   (F.panels_A && F.PA.fmt == 12 ? 256 : 0) |
                   (F.panels_T && F.PT.fmt == 12 ? 512 : 0) |
                   (F.panels_T && F.unit_T != 512 ? 8 : 0);
    info->code_N = F.panels_A ? -F.PA.blk_nnz : code(F.A);
    info->code_T = F.panels_T ? -F.PT.blk_nnz : code(F.AT);
    info->npan = F.npan;
    info->nseg = F.nseg;
    info->max_seg = F.max_seg;
}

static void check_dims(const tv3_op *op, const int32_t dims[3]) {
    TV3_REQUIRE(dims && dims[0] > 0 && dims[1] > 0 && dims[2] > 0, TV3_E_ARG, "bad dims");
    TV3_REQUIRE((int64_t)dims[0] * dims[1] * dims[2] == op->n, TV3_E_ARG,
                "dims do not match the operator's column count");
}

tv3_status tv3_op_prepare(tv3_op *op, const int32_t dims[3], tv3_layout_info *info) {
    if (!op) return TV3_E_ARG;
    return guard(op->ctx, [&] {
        check_dims(op, dims);
        set_device(op->ctx);
        fill_layout_info(fused_layout(op, dims), info);
    });
}

tv3_status tv3_op_layout_force(tv3_op *op, const int32_t dims[3], int32_t code_N, int32_t code_T,
                               int32_t pan_nnz, int32_t panels, tv3_layout_info *info) {
    if (!op) return TV3_E_ARG;
    auto ok = [](int L) {
        return L == 0 || (L >= 2 && L <= 64 && (L & (L - 1)) == 0) || L == -1024 || L == -2048 ||
               L == -4096;
    };
    return guard(op->ctx, [&] {
        TV3_REQUIRE(ok(code_N) && ok(code_T), TV3_E_ARG,
                    "launch codes are 0, a power of two in [2, 64], -1024, -2048 or -4096");
        TV3_REQUIRE(pan_nnz >= 0 && panels >= 0 && panels <= 1023 && (panels & 192) == 0,
                    TV3_E_ARG, "pan_nnz must be >= 0 and panels bits 0-5, 8 and 9");
        TV3_REQUIRE(!(panels & 256) || (panels & (16 | 32)), TV3_E_ARG,
                    "packed entries (bit 8) go with A's panels of 4 or 8 tiles (bit 4 or 5)");
        TV3_REQUIRE(!(panels & 512) || (panels & 2), TV3_E_ARG,
                    "packed A^T entries (bit 9) go with A^T's panel form (bit 1)");
        const int ua = panels & (4 | 16 | 32);
        TV3_REQUIRE((ua & (ua - 1)) == 0, TV3_E_ARG,
                    "at most one of the panel bits 2, 4 and 5 (A's panels of 2, 4 or 8 tiles)");
        check_dims(op, dims);
        set_device(op->ctx);
        layout_force(op, dims, code_N, code_T, pan_nnz, panels);
        fill_layout_info(fused_layout(op, dims), info);
    });
}

tv3_status tv3_op_layout_matvec(tv3_op *op, const int32_t dims[3], tv3_trans trans,
                                const float *x, float *y, void *stream) {
    if (!op) return TV3_E_ARG;
    return guard(op->ctx, [&] {
        TV3_REQUIRE(x && y, TV3_E_ARG, "x or y is NULL");
        TV3_REQUIRE(trans == TV3_N || trans == TV3_T, TV3_E_ARG, "bad trans");
        check_dims(op, dims);
        set_device(op->ctx);
        layout_matvec(op, dims, trans == TV3_T, x, y, (hipStream_t)stream);
    });
}

tv3_status tv3_bandwidth_probe(tv3_ctx *ctx, int64_t bytes, int32_t reps, double *copy_GBps,
                               double *read_GBps) {
    if (!ctx) return TV3_E_ARG;
    return guard(ctx, [&] {
        TV3_REQUIRE(bytes > 0 && reps > 0, TV3_E_ARG, "bytes and reps must be positive");
        set_device(ctx);
        bandwidth_probe(ctx, bytes, reps, copy_GBps, read_GBps);
    });
}

tv3_status tv3_op_layout_time(tv3_op *op, const int32_t dims[3], tv3_trans trans, int32_t reps,
                              double *us) {
    if (!op || !us) return TV3_E_ARG;
    return guard(op->ctx, [&] {
        TV3_REQUIRE(trans == TV3_N || trans == TV3_T, TV3_E_ARG, "bad trans");
        check_dims(op, dims);
        set_device(op->ctx);
        *us = layout_time(op, dims, trans == TV3_T, reps);
    });
}

tv3_status tv3_op_set_lanes(tv3_op *op, int32_t lanes_N, int32_t lanes_T) {
    if (!op) return TV3_E_ARG;
    auto ok = [](int L) {
        return L == 0 || (L >= 2 && L <= 64 && (L & (L - 1)) == 0) || L == -1024 || L == -2048 ||
               L == -4096;
    };
    if (!ok(lanes_N) || !ok(lanes_T)) return TV3_E_ARG;
    return guard(op->ctx, [&] {
        auto apply = [&](SparseRows &S, int L) {
            if (L > 0) { S.lanes = L; S.blk_nnz = 0; }
            else if (L < 0) build_row_blocks(op->ctx, S, -L);
        };
        if (op->A.rows) apply(op->A, lanes_N);
        if (op->AT.rows) apply(op->AT, lanes_T);
    });
}

}  // extern "C"
