"""The C++ drop-in (include/tval3_gpu_3d.hpp) executed on the GPU: a C++
program (tests/cpp/dropin_run.cpp) calls every tval3_gpu_3d overload the
reference exports (3d/tval3_gpu_3d.cpp:23-69: sparse in CSR and CSC form,
dense, matrix-free geometry, geometry + matrix, float and double) and
mat_vec_mul on device pointers; its outputs are checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import tv3_oracle as orc
from tval3_amd import api, workload as wl
from tval3_amd import _lib as L

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_cpp_dropin_runs_every_overload(tmp_path):
    import torch
    assert torch.cuda.is_available()
    N, E, R, md = 12, 24, 48, 400
    dims = (N, N, N)
    n = N ** 3
    u = wl.phantom_3d(dims=dims, seed=11)
    g = wl.transducer_geometry(E, R, dims=dims, seed=12)
    ctx = api.Context(0)
    A32 = api.DynamicOperator(ctx, g, "siddon").materialize("f32")
    rp, col, val = A32.download_csr()
    m, nnz = len(rp) - 1, int(rp[-1])
    val32 = val.astype(np.float32)
    oop = orc.csr_op(m, n, rp, col, val32.astype(np.float64))
    b = wl.add_noise(orc.matvec(oop, 0, u), 0.005, seed=13).astype(np.float32)
    cp, rw, cv = orc.csr_transpose(m, n, rp, col, val32.astype(np.float64))
    Ad = wl.dense_gaussian(md, n, seed=14)
    Ad32 = np.asfortranarray(Ad.astype(np.float32))
    bd = wl.add_noise(Ad32.astype(np.float64) @ u, 0.005, seed=15).astype(np.float32)
    rng = np.random.default_rng(16)
    x = rng.uniform(size=n).astype(np.float32)
    r = rng.normal(size=m).astype(np.float32)
    d = str(tmp_path)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(" ".join(str(v) for v in (N, N, N, m, nnz, md, E, R, *g["origin"], *g["voxel"])))
    files = {"b.f32": b, "rowptr.i32": rp.astype(np.int32), "col.i32": col.astype(np.int32),
             "val.f32": val32, "colptr.i32": cp, "row.i32": rw, "cval.f32": cv.astype(np.float32),
             "dense.f32": Ad32.reshape(-1, order="F"), "bd.f32": bd,
             "epos.f32": np.asarray(g["e_pos"], np.float32),
             "rpos.f32": np.asarray(g["r_pos"], np.float32), "x.f32": x, "r.f32": r}
    for k, a in files.items():
        np.ascontiguousarray(a).tofile(os.path.join(d, k))
    exe = os.path.join(d, "dropin_run")
    libdir = os.path.dirname(L.LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__",
                           "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                           os.path.join(ROOT, "tests", "cpp", "dropin_run.cpp"), "-o", exe,
                           "-L" + libdir, "-ltval3_mi355x", "-L/opt/rocm/lib", "-lamdhip64",
                           "-Wl,-rpath," + libdir + ":/opt/rocm/lib"])
    out = subprocess.run([exe, d], capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    info = {}
    for line in open(os.path.join(d, "info.txt")):
        k, *v = line.split()
        info[k] = v
    ld = lambda name, dt: np.fromfile(os.path.join(d, name), dt)
    o = dict(mu=512.0, mu0=512.0, tol_inn=1e-3, maxit=40, tv_iso=1)
    # sparse, both storage forms: the reference-shaped FP32 loop reproduces
    # the FP32 oracle (same iteration count, U within 1e-4)
    Uo, io, _, _ = orc.solve(oop, dims, b, orc.default_opts(**o), dtype=np.float32)
    Ucsr, Ucsc = ld("U_csr.f32", np.float32), ld("U_csc.f32", np.float32)
    print(f"csr: iters {info['csr'][0]} oracle {io.total_iters}; rel {rel(Ucsr, Uo):.2e}")
    assert int(info["csr"][0]) == io.total_iters and rel(Ucsr, Uo) < 1e-4
    assert (Ucsc == Ucsr).all()                       # same matrix, other storage
    # double overload against the FP64 oracle
    U64o, i64, _, _ = orc.solve(oop, dims, b.astype(np.float64), orc.default_opts(**o))
    U64 = ld("U_csr64.f64", np.float64)
    print(f"csr64: rel {rel(U64, U64o):.2e}")
    assert int(info["csr64"][0]) == i64.total_iters and rel(U64, U64o) < 1e-10
    # dense
    Udo, ido, _, _ = orc.solve(orc.dense_op(Ad32.astype(np.float64)), dims, bd,
                               orc.default_opts(**o), dtype=np.float32)
    Ud = ld("U_dense.f32", np.float32)
    print(f"dense: iters {info['dense'][0]} oracle {ido.total_iters}; rel {rel(Ud, Udo):.2e}")
    assert int(info["dense"][0]) == ido.total_iters and rel(Ud, Udo) < 1e-4
    # matrix-free and hybrid: the walk's products equal the CSR's up to FP32
    # summation order, which the BB step amplifies over 40 iterations
    for k in ("geo", "hyb"):
        Ug = ld(f"U_{k}.f32", np.float32)
        print(f"{k}: iters {info[k][0]}; rel to oracle {rel(Ug, Uo):.2e}")
        assert rel(Ug, Uo) < 2e-3
    # mat_vec_mul: FP32 products within the summation bound, FP64 to 1e-15
    yN, yT = ld("yN.f32", np.float32), ld("yT.f32", np.float32)
    wN = orc.matvec(oop, 0, x.astype(np.float64))
    wT = orc.matvec(oop, 1, r.astype(np.float64))
    absA = orc.csr_op(m, n, rp, col, np.abs(val32.astype(np.float64)))
    bN = orc.matvec(absA, 0, np.abs(x).astype(np.float64))
    bT = orc.matvec(absA, 1, np.abs(r).astype(np.float64))
    gam = 64 * 2.0 ** -24
    assert (np.abs(yN - wN) <= gam * bN + 1e-30).all()
    assert (np.abs(yT - wT) <= gam * bT + 1e-30).all()
    assert np.abs(ld("yN64.f64", np.float64) - wN).max() <= 1e-14 * np.abs(wN).max()
    assert info["shape_error"] == ["thrown"]
