// dropin_run.cpp -- runs the C++ drop-in (include/tval3_gpu_3d.hpp) the way a
// caller of the reference's exported overloads would (3d/tval3_gpu_3d.cpp:
// 23-69): every tval3_gpu_3d overload and mat_vec_mul on device pointers.
// Built and run by tests/test_gpu_dropin.py, which writes the inputs into
// the directory given as argv[1] and checks the outputs against the oracle.
//
// Inputs (raw little-endian arrays): meta.txt "p q r m nnz md E R ox oy oz
// hx hy hz"; b.f32 (m); rowptr.i32, col.i32, val.f32 (CSR of A); colptr.i32,
// row.i32, cval.f32 (its CSC); dense.f32 (md x n, column-major), bd.f32 (md);
// epos.f32 (E x 3), rpos.f32 (R x 3); x.f32 (n), r.f32 (m).
// Outputs: U_<case>.f32 / .f64, yN.f32, yT.f32, yN64.f64, info.txt.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "tval3_gpu_3d.hpp"

static std::string dir;

template <class T>
static std::vector<T> load(const std::string &name, size_t count) {
    std::vector<T> v(count);
    std::ifstream f(dir + "/" + name, std::ios::binary);
    if (!f.read(reinterpret_cast<char *>(v.data()), sizeof(T) * count))
        throw std::runtime_error("cannot read " + name);
    return v;
}

template <class T>
static void save(const std::string &name, const T *p, size_t count) {
    std::ofstream f(dir + "/" + name, std::ios::binary);
    f.write(reinterpret_cast<const char *>(p), sizeof(T) * count);
}

#define HIPCHK(x)                                                               \
    do {                                                                        \
        if ((x) != hipSuccess) throw std::runtime_error("HIP error at " #x);    \
    } while (0)

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    dir = argv[1];
    try {
        int p, q, r, m, nnz, md, E, R;
        float org[3], vox[3];
        {
            std::ifstream f(dir + "/meta.txt");
            f >> p >> q >> r >> m >> nnz >> md >> E >> R >> org[0] >> org[1] >> org[2] >>
                vox[0] >> vox[1] >> vox[2];
            if (!f) throw std::runtime_error("bad meta.txt");
        }
        const int n = p * q * r;
        auto b = load<float>("b.f32", m);
        auto rowptr = load<int>("rowptr.i32", m + 1);
        auto col = load<int>("col.i32", nnz);
        auto val = load<float>("val.f32", nnz);
        auto colptr = load<int>("colptr.i32", n + 1);
        auto row = load<int>("row.i32", nnz);
        auto cval = load<float>("cval.f32", nnz);
        auto dense = load<float>("dense.f32", (size_t)md * n);
        auto bd = load<float>("bd.f32", md);
        auto epos = load<float>("epos.f32", 3 * E);
        auto rpos = load<float>("rpos.f32", 3 * R);
        auto x = load<float>("x.f32", n);
        auto rr = load<float>("r.f32", m);

        tval3_options<float> o;
        o.maxit = 40; o.tol_inn = 1e-3f; o.mu = o.mu0 = 512.f; o.tv_mode = TV3_TV_ISO;
        tval3_options<double> od;
        od.maxit = 40; od.tol_inn = 1e-3; od.mu = od.mu0 = 512.0; od.tv_mode = TV3_TV_ISO;

        mat_host<float> bh(m, b.data()), Ut(p, q, r);
        std::ofstream info(dir + "/info.txt");
        auto report = [&](const char *tag, int iters, int outer) {
            info << tag << " " << iters << " " << outer << "\n";
        };
        // 1. sparse_mat_host in CSR and in CSC (the reference's default)
        sparse_mat_host<float> Acsr(m, n, nnz, val.data(), rowptr.data(), col.data(),
                                    sparse_mat_csr);
        sparse_mat_host<float> Acsc(m, n, nnz, cval.data(), colptr.data(), row.data(),
                                    sparse_mat_csc);
        {
            mat_host<float> U(p, q, r);
            auto i = tval3_gpu_3d(U, Acsr, bh, o, Ut, true, 0, 1, nullptr);
            save("U_csr.f32", U.data(), n);
            report("csr", i.total_iters, i.outer_iters);
        }
        {
            mat_host<float> U(p, q, r);
            auto i = tval3_gpu_3d(U, Acsc, bh, o, Ut);
            save("U_csc.f32", U.data(), n);
            report("csc", i.total_iters, i.outer_iters);
        }
        // 2. dense column-major md x n
        {
            mat_host<float> U(p, q, r), Ad(md * n, dense.data()), bdh(md, bd.data());
            auto i = tval3_gpu_3d(U, Ad, bdh, o, Ut);
            save("U_dense.f32", U.data(), n);
            report("dense", i.total_iters, i.outer_iters);
        }
        // 3. geometry (matrix-free Siddon) and 4. geometry + matrix (hybrid)
        geometry_host g;
        g.num_emitters = E; g.num_receivers = R; g.num_mps = 1;
        g.rv_x = q; g.rv_y = p; g.rv_z = r;
        g.pos_emitters = epos.data(); g.pos_receivers = rpos.data();
        for (int a = 0; a < 3; ++a) { g.origin[a] = org[a]; g.voxel[a] = vox[a]; }
        {
            mat_host<float> U(p, q, r);
            auto i = tval3_gpu_3d(U, g, bh, o, Ut);
            save("U_geo.f32", U.data(), n);
            report("geo", i.total_iters, i.outer_iters);
        }
        {
            mat_host<float> U(p, q, r);
            auto i = tval3_gpu_3d(U, g, Acsc, bh, o, Ut);
            save("U_hyb.f32", U.data(), n);
            report("hyb", i.total_iters, i.outer_iters);
        }
        // double overload
        {
            std::vector<double> v64(val.begin(), val.end()), b64(b.begin(), b.end());
            sparse_mat_host<double> A64(m, n, nnz, v64.data(), rowptr.data(), col.data(),
                                        sparse_mat_csr);
            mat_host<double> U(p, q, r), bh64(m, b64.data()), Ut64(p, q, r);
            auto i = tval3_gpu_3d(U, A64, bh64, od, Ut64);
            save("U_csr64.f64", U.data(), n);
            report("csr64", i.total_iters, i.outer_iters);
        }
        // 5. mat_vec_mul(trans, A, x, y, stream, mulChain) on device pointers
        {
            tval3_operator op(Acsc);
            float *dx, *dy, *dr, *dz;
            HIPCHK(hipMalloc(&dx, sizeof(float) * n));
            HIPCHK(hipMalloc(&dz, sizeof(float) * n));
            HIPCHK(hipMalloc(&dr, sizeof(float) * m));
            HIPCHK(hipMalloc(&dy, sizeof(float) * m));
            HIPCHK(hipMemcpy(dx, x.data(), sizeof(float) * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(dr, rr.data(), sizeof(float) * m, hipMemcpyHostToDevice));
            mat_vec_mul(0, op, dx, dy, nullptr, false);
            mat_vec_mul(1, op, dr, dz, nullptr, true);
            HIPCHK(hipDeviceSynchronize());
            std::vector<float> yN(m), yT(n);
            HIPCHK(hipMemcpy(yN.data(), dy, sizeof(float) * m, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(yT.data(), dz, sizeof(float) * n, hipMemcpyDeviceToHost));
            save("yN.f32", yN.data(), m);
            save("yT.f32", yT.data(), n);
            std::vector<double> v64(val.begin(), val.end()), x64(x.begin(), x.end());
            sparse_mat_host<double> A64(m, n, nnz, v64.data(), rowptr.data(), col.data(),
                                        sparse_mat_csr);
            tval3_operator op64(A64);
            double *dx64, *dy64;
            HIPCHK(hipMalloc(&dx64, sizeof(double) * n));
            HIPCHK(hipMalloc(&dy64, sizeof(double) * m));
            HIPCHK(hipMemcpy(dx64, x64.data(), sizeof(double) * n, hipMemcpyHostToDevice));
            mat_vec_mul(0, op64, dx64, dy64);
            HIPCHK(hipDeviceSynchronize());
            std::vector<double> yN64(m);
            HIPCHK(hipMemcpy(yN64.data(), dy64, sizeof(double) * m, hipMemcpyDeviceToHost));
            save("yN64.f64", yN64.data(), m);
            for (void *pp : {(void *)dx, (void *)dy, (void *)dr, (void *)dz, (void *)dx64,
                             (void *)dy64})
                HIPCHK(hipFree(pp));
        }
        // errors cross the boundary as tval3_gpu_exception
        try {
            mat_host<float> U(p, q, r + 1);
            tval3_gpu_3d(U, Acsr, bh, o, Ut);
            info << "shape_error none\n";
        } catch (const tval3_gpu_exception &e) {
            info << "shape_error thrown\n";
        }
        info.close();
        std::printf("dropin_run: ok\n");
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "dropin_run: %s\n", e.what());
        return 1;
    }
}
