// tval3_gpu_3d.hpp -- header-only C++ drop-in for the reference's 3-D entry
// points, on top of the C ABI in tval3_mi355x.h.
//
// It re-creates the exported overload set of 3d/tval3_gpu_3d.cpp:23-69
//     tval3_info<T> tval3_gpu_3d(mat_host<T>& U, A, const mat_host<T>& b,
//                                const tval3_options<T>& opts,
//                                const mat_host<T>& Ut, bool pagelocked,
//                                int mainGPU, int numGPUs, int* gpuArray)
// for A = sparse_mat_host<T>, mat_host<T> (dense, column-major),
// geometry_host (matrix-free) and geometry_host + sparse_mat_host<T>
// (hybrid, 3d/tval3_gpu_3d.cpp:1603-1668), and the operator plug-in call
//     mat_vec_mul(trans, A, x, y, stream, mulChain)
// (3d/mat_vec_mul_dense.h:26, 3d/mat_vec_mul_sparse.h:1704,
// 3d/mat_vec_mul_dynamic.h:683, 3d/mat_vec_mul.h:733) for device pointers.
// Errors throw tval3_gpu_exception (2d/tval3_types.h:45-49), as the
// reference's HANDLE_ERROR and check_params do.
//
// The reference's 3-D host container and option headers are not part of its
// published sources; the types below follow their use in 3d/tval3_gpu_3d.cpp
// (U.dim_y/dim_x/dim_z/len, b.len, A.rows/cols/nnz/ptr()/ind()/val(), the
// opts fields at :1088-1345) and the 2-D versions (2d/container_host.h,
// 2d/tval3_types.h).  One GPU: numGPUs must be 0 or 1 and gpuArray NULL or
// {mainGPU}.  pagelocked is accepted and has no effect (transfers go
// through the library's own staging).
#pragma once

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tval3_mi355x.h"

enum T_mat_format { mat_col_major = 0, mat_row_major = 1 };
enum T_sparse_mat_format { sparse_mat_csc = 0, sparse_mat_csr = 1 };

class tval3_gpu_exception : public std::runtime_error {
public:
    explicit tval3_gpu_exception(const std::string &message) : std::runtime_error(message) {}
};

// Host volume or vector: dim_y fastest (column-major), len = product.
template <class T>
class mat_host {
    std::vector<T> own_;
    T *p_;

public:
    const int dim_y, dim_x, dim_z, len;
    const T_mat_format format;
    const bool pagelocked;

    mat_host(int l_y, int l_x, int l_z, bool init = true, bool mem_pagelocked = false)
        : own_((size_t)l_y * l_x * l_z), p_(own_.data()), dim_y(l_y), dim_x(l_x), dim_z(l_z),
          len(l_y * l_x * l_z), format(mat_col_major), pagelocked(mem_pagelocked) {
        (void)init;   // storage is value-initialised either way
    }
    explicit mat_host(int num_elements, bool init = true, bool mem_pagelocked = false)
        : mat_host(num_elements, 1, 1, init, mem_pagelocked) {}
    // wraps caller memory (not copied, not freed)
    mat_host(int l_y, int l_x, int l_z, T *buffer, bool mem_pagelocked = false)
        : p_(buffer), dim_y(l_y), dim_x(l_x), dim_z(l_z), len(l_y * l_x * l_z),
          format(mat_col_major), pagelocked(mem_pagelocked) {}
    mat_host(int num_elements, T *buffer, bool mem_pagelocked = false)
        : mat_host(num_elements, 1, 1, buffer, mem_pagelocked) {}
    mat_host(const mat_host &m)
        : own_(m.p_, m.p_ + m.len), p_(own_.data()), dim_y(m.dim_y), dim_x(m.dim_x),
          dim_z(m.dim_z), len(m.len), format(m.format), pagelocked(m.pagelocked) {}
    mat_host &operator=(const mat_host &m) {
        if (m.len != len) throw tval3_gpu_exception("mat_host assignment: size mismatch");
        std::memcpy(p_, m.p_, sizeof(T) * (size_t)len);
        return *this;
    }
    T *data() { return p_; }
    const T *data() const { return p_; }
    T &operator[](int i) { return p_[i]; }
    const T &operator[](int i) const { return p_[i]; }
};

// Host sparse matrix, CSC (the reference's default) or CSR.
template <class T>
class sparse_mat_host {
    std::vector<int> ptr_, ind_;
    std::vector<T> val_;

public:
    const int rows, cols, nnz;
    const T_sparse_mat_format format;

    sparse_mat_host(int num_rows, int num_cols, int n_nonzero,
                    T_sparse_mat_format storage_format = sparse_mat_csc)
        : ptr_((storage_format == sparse_mat_csc ? num_cols : num_rows) + 1), ind_(n_nonzero),
          val_(n_nonzero), rows(num_rows), cols(num_cols), nnz(n_nonzero),
          format(storage_format) {}
    sparse_mat_host(int num_rows, int num_cols, int n_nonzero, const T *buff_val,
                    const int *buff_ptr, const int *buff_ind,
                    T_sparse_mat_format storage_format = sparse_mat_csc)
        : sparse_mat_host(num_rows, num_cols, n_nonzero, storage_format) {
        std::memcpy(ptr_.data(), buff_ptr, sizeof(int) * ptr_.size());
        std::memcpy(ind_.data(), buff_ind, sizeof(int) * (size_t)n_nonzero);
        std::memcpy(val_.data(), buff_val, sizeof(T) * (size_t)n_nonzero);
    }
    int *ptr() { return ptr_.data(); }
    const int *ptr() const { return ptr_.data(); }
    int *ind() { return ind_.data(); }
    const int *ind() const { return ind_.data(); }
    T *val() { return val_.data(); }
    const T *val() const { return val_.data(); }
};

// Transducer geometry (3d/container_device.h:103-172 geometry_device);
// arrays are caller-owned host memory.  Integer coordinates select the
// Bresenham weights of the reference (a constant scale_factor per visited
// voxel); world positions (pos_*) select exact Siddon path lengths.
struct geometry_host {
    int num_emitters = 0, num_receivers = 0, num_mps = 1;
    int rv_x = 0, rv_y = 0, rv_z = 0;
    float scale_factor = 1.f;
    const int *x_emitters = nullptr, *y_emitters = nullptr, *z_emitters = nullptr;
    const int *x_receivers = nullptr, *y_receivers = nullptr, *z_receivers = nullptr;
    const float *pos_emitters = nullptr, *pos_receivers = nullptr;
    float origin[3] = {0, 0, 0}, voxel[3] = {1, 1, 1};
    const uint8_t *use_path = nullptr;
};

// 2d/tval3_types.h:7-35 plus the 3-D fields read at 3d/tval3_gpu_3d.cpp:398-468;
// tv_mode, solver and graph_iters select this library's variants.
template <class T>
struct tval3_options {
    T mu = 256, mu0 = 256, beta = 32, beta0 = 32, rate_cnt = 2, tol = T(1e-6), tol_inn = T(1e-3);
    int maxcnt = 10, maxit = 1025;
    T c = T(1e-5), gamma = T(0.6), gam = T(0.9995), rate_gam = T(0.9);
    bool isreal = true, nonneg = true;
    int numInitIterations = 0;
    T skipMulRatio = 0, maxRelativeChange = 0;
    tv3_tv_mode tv_mode = TV3_TV_ANISO;
    tv3_solver_kind solver = TV3_SOLVER_PLAIN;
    int graph_iters = 0;
    int tiles_per_wg = 0;
};

template <class T>
struct tval3_info {
    int total_iters = 0, outer_iters = 0;
    T rel_chg = 0, rel_error = 0;
    double secs = 0;
};

namespace tval3_detail {

inline void check(tv3_status st, const tv3_ctx *ctx) {
    if (st != TV3_OK) {
        const char *m = tv3_last_error(ctx);
        throw tval3_gpu_exception(std::string("tval3: ") + (m ? m : "error") + " (status " +
                                  std::to_string((int)st) + ")");
    }
}

// one context per device for the process, as the reference keeps its
// handles and streams in process globals (3d/tval3_gpu_3d.cpp:71-79)
inline tv3_ctx *context(int device) {
    static std::map<int, std::unique_ptr<tv3_ctx, void (*)(tv3_ctx *)>> ctxs;
    auto it = ctxs.find(device);
    if (it != ctxs.end()) return it->second.get();
    tv3_ctx *c = nullptr;
    check(tv3_ctx_create(device, &c), nullptr);
    ctxs.emplace(device, std::unique_ptr<tv3_ctx, void (*)(tv3_ctx *)>(c, tv3_ctx_destroy));
    return c;
}

inline int device_of(int mainGPU, int numGPUs, const int *gpuArray) {
    if (numGPUs > 1 || (gpuArray && numGPUs == 1 && gpuArray[0] != (mainGPU < 0 ? 0 : mainGPU)))
        throw tval3_gpu_exception("tval3: this library runs on one GPU (numGPUs <= 1)");
    return mainGPU < 0 ? 0 : mainGPU;
}

template <class T>
tv3_opts to_opts(const tval3_options<T> &o) {
    tv3_opts r;
    tv3_opts_default(&r);
    r.mu = o.mu; r.mu0 = o.mu0; r.beta = o.beta; r.beta0 = o.beta0; r.rate_cnt = o.rate_cnt;
    r.tol = o.tol; r.tol_inn = o.tol_inn; r.maxcnt = o.maxcnt; r.maxit = o.maxit;
    r.c = o.c; r.gamma = o.gamma; r.gam = o.gam; r.rate_gam = o.rate_gam;
    r.isreal = o.isreal; r.nonneg = o.nonneg;
    r.numInitIterations = o.numInitIterations; r.skipMulRatio = o.skipMulRatio;
    r.maxRelativeChange = o.maxRelativeChange;
    r.tv_mode = o.tv_mode; r.solver = o.solver; r.graph_iters = o.graph_iters;
    r.tiles_per_wg = o.tiles_per_wg;
    return r;
}

template <class T> constexpr tv3_dtype dtype_of();
template <> constexpr tv3_dtype dtype_of<float>() { return TV3_F32; }
template <> constexpr tv3_dtype dtype_of<double>() { return TV3_F64; }

struct op_deleter {
    void operator()(tv3_op *p) const { tv3_op_destroy(p); }
};
using op_ptr = std::unique_ptr<tv3_op, op_deleter>;

template <class T>
op_ptr make_sparse(tv3_ctx *ctx, const sparse_mat_host<T> &A) {
    tv3_op *op = nullptr;
    if (A.format == sparse_mat_csr) {
        check(tv3_op_csr_create(ctx, A.rows, A.cols, A.nnz, A.ptr(), A.ind(), A.val(),
                                dtype_of<T>(), nullptr, nullptr, nullptr, &op), ctx);
        return op_ptr(op);
    }
    // CSC given: build the CSR on the host (stable), pass the CSC for A^T
    std::vector<int> rp(A.rows + 1, 0), ci(A.nnz);
    std::vector<T> cv(A.nnz);
    for (int k = 0; k < A.nnz; ++k) ++rp[A.ind()[k] + 1];
    for (int i = 0; i < A.rows; ++i) rp[i + 1] += rp[i];
    std::vector<int> pos(rp.begin(), rp.end() - 1);
    for (int j = 0; j < A.cols; ++j)
        for (int k = A.ptr()[j]; k < A.ptr()[j + 1]; ++k) {
            const int q = pos[A.ind()[k]]++;
            ci[q] = j;
            cv[q] = A.val()[k];
        }
    check(tv3_op_csr_create(ctx, A.rows, A.cols, A.nnz, rp.data(), ci.data(), cv.data(),
                            dtype_of<T>(), A.ptr(), A.ind(), A.val(), &op), ctx);
    return op_ptr(op);
}

template <class T>
op_ptr make_dense(tv3_ctx *ctx, const mat_host<T> &A, int rows) {
    if (rows <= 0 || A.len % rows) throw tval3_gpu_exception("tval3: dense A shape");
    tv3_op *op = nullptr;
    check(tv3_op_dense_create(ctx, rows, A.len / rows, A.data(), rows, dtype_of<T>(), &op), ctx);
    return op_ptr(op);
}

inline tv3_geometry to_geometry(const geometry_host &g) {
    tv3_geometry r;
    std::memset(&r, 0, sizeof(r));
    r.num_emitters = g.num_emitters; r.num_receivers = g.num_receivers; r.num_mps = g.num_mps;
    r.rv_x = g.rv_x; r.rv_y = g.rv_y; r.rv_z = g.rv_z; r.scale_factor = g.scale_factor;
    r.x_emitters = g.x_emitters; r.y_emitters = g.y_emitters; r.z_emitters = g.z_emitters;
    r.x_receivers = g.x_receivers; r.y_receivers = g.y_receivers; r.z_receivers = g.z_receivers;
    r.pos_emitters = g.pos_emitters; r.pos_receivers = g.pos_receivers;
    for (int a = 0; a < 3; ++a) { r.origin[a] = g.origin[a]; r.voxel[a] = g.voxel[a]; }
    r.use_path = g.use_path;
    return r;
}

inline tv3_weight_mode mode_of(const geometry_host &g) {
    return g.pos_emitters && g.pos_receivers ? TV3_WEIGHT_SIDDON : TV3_WEIGHT_BRESENHAM;
}

// the geometry's traversal for A x, A_mat (CSC as given, CSR transposed on
// the device) for A^T x
template <class T>
op_ptr make_hybrid(tv3_ctx *ctx, const geometry_host &geo, const sparse_mat_host<T> &A) {
    const tv3_geometry g = to_geometry(geo);
    tv3_op *op = nullptr;
    const bool csc = A.format == sparse_mat_csc;
    check(tv3_op_hybrid_create(ctx, &g, mode_of(geo), A.nnz, csc ? nullptr : A.ptr(),
                               csc ? nullptr : A.ind(), csc ? nullptr : A.val(), dtype_of<T>(),
                               csc ? A.ptr() : nullptr, csc ? A.ind() : nullptr,
                               csc ? A.val() : nullptr, &op),
          ctx);
    op_ptr r(op);
    int64_t m = 0, n = 0;
    check(tv3_op_info(op, &m, &n, nullptr), ctx);
    if (m != A.rows || n != A.cols)
        throw tval3_gpu_exception("tval3: A_mat does not have the geometry's shape");
    return r;
}

template <class T>
tval3_info<T> solve(tv3_ctx *ctx, tv3_op *op, mat_host<T> &U, const mat_host<T> &b,
                    const tval3_options<T> &opts, const mat_host<T> &Ut) {
    const int32_t dims[3] = {U.dim_y, U.dim_x, U.dim_z};
    tv3_opts o = to_opts(opts);
    tv3_info info;
    const T *ut = Ut.len == U.len ? Ut.data() : nullptr;   // Ut of another size: no error
    tv3_status st;
    if (dtype_of<T>() == TV3_F64)
        st = tv3_solve_f64(ctx, op, &o, dims, reinterpret_cast<const double *>(b.data()),
                           reinterpret_cast<double *>(U.data()),
                           reinterpret_cast<const double *>(ut), &info, nullptr);
    else
        st = tv3_solve(ctx, op, &o, dims, reinterpret_cast<const float *>(b.data()),
                       reinterpret_cast<float *>(U.data()), reinterpret_cast<const float *>(ut),
                       &info, nullptr);
    check(st, ctx);
    tval3_info<T> r;
    r.total_iters = info.total_iters;
    r.outer_iters = info.outer_iters;
    r.rel_chg = (T)info.rel_chg;
    r.rel_error = (T)info.rel_error;
    r.secs = info.secs;
    return r;
}

}  // namespace tval3_detail

// ---- tval3_gpu_3d overload set (3d/tval3_gpu_3d.cpp:23-69) ----------------
template <class T>
const tval3_info<T> tval3_gpu_3d(mat_host<T> &U, const sparse_mat_host<T> &A,
                                 const mat_host<T> &b, const tval3_options<T> &opts,
                                 const mat_host<T> &Ut, bool pagelocked = true, int mainGPU = 0,
                                 int numGPUs = 1, int *gpuArray = nullptr) {
    (void)pagelocked;
    tv3_ctx *ctx = tval3_detail::context(tval3_detail::device_of(mainGPU, numGPUs, gpuArray));
    if (A.rows != b.len || A.cols != U.len)
        throw tval3_gpu_exception("tval3: dimensions of A, b and U do not match");
    auto op = tval3_detail::make_sparse(ctx, A);
    return tval3_detail::solve(ctx, op.get(), U, b, opts, Ut);
}

// dense A: column-major b.len x U.len
template <class T>
const tval3_info<T> tval3_gpu_3d(mat_host<T> &U, const mat_host<T> &A, const mat_host<T> &b,
                                 const tval3_options<T> &opts, const mat_host<T> &Ut,
                                 bool pagelocked = true, int mainGPU = 0, int numGPUs = 1,
                                 int *gpuArray = nullptr) {
    (void)pagelocked;
    tv3_ctx *ctx = tval3_detail::context(tval3_detail::device_of(mainGPU, numGPUs, gpuArray));
    if ((int64_t)b.len * U.len != A.len)
        throw tval3_gpu_exception("tval3: dimensions of A, b and U do not match");
    auto op = tval3_detail::make_dense(ctx, A, b.len);
    return tval3_detail::solve(ctx, op.get(), U, b, opts, Ut);
}

// matrix-free A from the transducer geometry (3d/tval3_gpu_3d.cpp:1550-1600)
template <class T>
const tval3_info<T> tval3_gpu_3d(mat_host<T> &U, const geometry_host &A, const mat_host<T> &b,
                                 const tval3_options<T> &opts, const mat_host<T> &Ut,
                                 bool pagelocked = true, int mainGPU = 0, int numGPUs = 1,
                                 int *gpuArray = nullptr) {
    (void)pagelocked;
    tv3_ctx *ctx = tval3_detail::context(tval3_detail::device_of(mainGPU, numGPUs, gpuArray));
    const tv3_geometry g = tval3_detail::to_geometry(A);
    tv3_op *op = nullptr;
    tval3_detail::check(tv3_op_dynamic_create(ctx, &g, tval3_detail::mode_of(A), &op), ctx);
    tval3_detail::op_ptr hold(op);
    return tval3_detail::solve(ctx, op, U, b, opts, Ut);
}

// hybrid: A x by the geometry, A^T x by A_mat, the caller's precomputed
// matrix (sparse_geom_mm, 3d/mat_vec_mul.h:104-120,732-756; entry
// 3d/tval3_gpu_3d.cpp:1603-1668)
template <class T>
const tval3_info<T> tval3_gpu_3d(mat_host<T> &U, const geometry_host &A_geo,
                                 const sparse_mat_host<T> &A_mat, const mat_host<T> &b,
                                 const tval3_options<T> &opts, const mat_host<T> &Ut,
                                 bool pagelocked = true, int mainGPU = 0, int numGPUs = 1,
                                 int *gpuArray = nullptr) {
    (void)pagelocked;
    tv3_ctx *ctx = tval3_detail::context(tval3_detail::device_of(mainGPU, numGPUs, gpuArray));
    if (A_mat.rows != b.len || A_mat.cols != U.len)
        throw tval3_gpu_exception("tval3: dimensions of A, b and U do not match");
    auto op = tval3_detail::make_hybrid(ctx, A_geo, A_mat);
    return tval3_detail::solve(ctx, op.get(), U, b, opts, Ut);
}

// ---- operator plug-in: mat_vec_mul(trans, A, x, y, stream, mulChain) ------
// A device operator built once from any of the host descriptions above.
class tval3_operator {
    tv3_ctx *ctx_;
    tval3_detail::op_ptr op_;

public:
    template <class T>
    tval3_operator(const sparse_mat_host<T> &A, int device = 0)
        : ctx_(tval3_detail::context(device)), op_(tval3_detail::make_sparse(ctx_, A)) {}
    tval3_operator(const geometry_host &g, int device = 0) : ctx_(tval3_detail::context(device)) {
        const tv3_geometry t = tval3_detail::to_geometry(g);
        tv3_op *op = nullptr;
        tval3_detail::check(tv3_op_dynamic_create(ctx_, &t, tval3_detail::mode_of(g), &op), ctx_);
        op_.reset(op);
    }
    // sparse_geom_mm: A x by the geometry, A^T x by A_mat
    template <class T>
    tval3_operator(const geometry_host &g, const sparse_mat_host<T> &A_mat, int device = 0)
        : ctx_(tval3_detail::context(device)), op_(tval3_detail::make_hybrid(ctx_, g, A_mat)) {}
    tv3_op *get() const { return op_.get(); }
    tv3_ctx *context() const { return ctx_; }
};

// trans: 0 = N, 1 = T (the values of cublasOperation_t's N and T); x, y
// device pointers; y overwritten; asynchronous on stream (a hipStream_t, NULL =
// the null stream).  mulChain is a multi-GPU hint without effect here.
inline int mat_vec_mul(int trans, const tval3_operator &A, const float *x, float *y,
                       void *stream = nullptr, bool mulChain = false) {
    (void)mulChain;
    tval3_detail::check(tv3_matvec(A.get(), trans ? TV3_T : TV3_N, x, y, stream), A.context());
    return 0;
}
inline int mat_vec_mul(int trans, const tval3_operator &A, const double *x, double *y,
                       void *stream = nullptr, bool mulChain = false) {
    (void)mulChain;
    tval3_detail::check(tv3_matvec_f64(A.get(), trans ? TV3_T : TV3_N, x, y, stream),
                        A.context());
    return 0;
}
