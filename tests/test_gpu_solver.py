"""Whole-solve parity on the GPU against the oracle (through the C ABI).

The FP32 plain path must reproduce the oracle's FP32 iterates (same number of
inner/outer iterations, final U within 1e-4 relative L2); the FP64 plain path
must match the FP64 oracle to 1e-9.  Every result also states its error
against the FP64 oracle."""
import numpy as np
import pytest

import tv3_oracle as orc
from tval3_amd import api, workload as wl
from tval3_amd._lib import TV3Error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available()
    return api.Context(0)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def oracle_opts(o):
    return orc.default_opts(mu=o.mu, mu0=o.mu0, beta=o.beta, beta0=o.beta0, tol=o.tol,
                            tol_inn=o.tol_inn, maxit=o.maxit, maxcnt=o.maxcnt,
                            tv_iso=int(o.tv_mode == 1))


def problem_2d(N=32, m_frac=0.4, seed=1):
    u = wl.phantom_2d(N, 6, seed=seed)
    A = wl.dense_gaussian(int(m_frac * N * N), N * N, seed=seed + 3)
    b = wl.add_noise(A @ u, 1e-3, seed=seed + 2)
    return u, A, b


def problem_3d(ctx, N=16, E=24, R=64, seed=1):
    u = wl.phantom_3d(N, 4, seed=seed)
    g = wl.transducer_geometry(E, R, N=N, seed=seed + 1)
    dyn = api.DynamicOperator(ctx, g, "siddon")
    A = dyn.materialize("f32")
    rp, c, v = A.download_csr()
    oop = orc.csr_op(A.m, A.n, rp, c, v)
    b = wl.add_noise(orc.matvec(oop, 0, u), 0.005, seed=seed + 2)
    return u, g, dyn, A, oop, b


def test_plain_fp32_dense_2d_matches_oracle(ctx):
    N = 32
    u, A, b = problem_2d(N)
    op = api.DenseOperator(ctx, A)
    o = api.default_opts(maxit=120, tol=1e-4, tol_inn=1e-3)
    U, info, hist = api.tval3_gpu_3d(ctx, op, b.astype(np.float32), (N, N, 1), o, Ut=u)
    A32 = A.astype(np.float32).astype(np.float64)
    oop = orc.dense_op(A32)
    Uo, io, ho, _ = orc.solve(oop, (N, N, 1), b.astype(np.float32), oracle_opts(o), Ut=u,
                              dtype=np.float32)
    U64, i64, _, _ = orc.solve(oop, (N, N, 1), b, oracle_opts(o), Ut=u)
    print(f"2d: iters {info['total_iters']}/{io.total_iters} rel(gpu,orc32)={rel(U, Uo):.2e} "
          f"rel(gpu,orc64)={rel(U, U64):.2e} err={info['rel_error']:.3e}")
    assert info["total_iters"] == io.total_iters and info["outer_iters"] == io.outer_iters
    assert (hist["cnt"] == ho["cnt"]).all()
    assert rel(U, Uo) < 1e-4
    assert np.allclose(hist["f"], ho["f"], rtol=1e-4)


def test_plain_fp32_csr_3d_iso_matches_oracle(ctx):
    N = 16
    u, g, dyn, A, oop, b = problem_3d(ctx, N)
    o = api.default_opts(maxit=80, tol_inn=1e-3, mu=2.0 ** 9, mu0=2.0 ** 9, tv_mode="iso")
    U, info, hist = api.tval3_gpu_3d(ctx, A, b.astype(np.float32), (N, N, N), o, Ut=u)
    Uo, io, ho, _ = orc.solve(oop, (N, N, N), b.astype(np.float32), oracle_opts(o), Ut=u,
                              dtype=np.float32)
    U64, i64, _, _ = orc.solve(oop, (N, N, N), b, oracle_opts(o), Ut=u)
    print(f"3d: iters {info['total_iters']}/{io.total_iters} rel(gpu,orc32)={rel(U, Uo):.2e} "
          f"rel(gpu,orc64)={rel(U, U64):.2e} err={info['rel_error']:.3e}/{i64.rel_error:.3e}")
    assert info["total_iters"] == io.total_iters
    assert rel(U, Uo) < 1e-4
    assert info["n_matvec_N"] > info["total_iters"]


@pytest.mark.parametrize("skip", [(0.5, 2, 0.0), (2.0, 3, 1e30), (1.0, 0, 0.05)])
def test_plain_skip_heuristic_matches_oracle(ctx, skip):
    """A^T-skip rule (3d/tval3_gpu_3d.cpp:398-468, skipped g kept as
    A^T(last Au) - Atb): the FP32 plain solver takes the same skip decisions
    as the FP32 oracle (same count of A^T products) and the same iterates."""
    ratio, ninit, maxrel = skip
    N = 32
    u, A, b = problem_2d(N, seed=4)
    op = api.DenseOperator(ctx, A)
    kw = dict(maxit=100, tol=1e-4, tol_inn=1e-3, skipMulRatio=ratio, numInitIterations=ninit,
              maxRelativeChange=maxrel)
    o = api.default_opts(**kw)
    U, info, hist = api.tval3_gpu_3d(ctx, op, b.astype(np.float32), (N, N, 1), o, Ut=u)
    oop = orc.dense_op(A.astype(np.float32).astype(np.float64))
    oo = oracle_opts(o)
    oo.skipMulRatio, oo.numInitIterations, oo.maxRelativeChange = ratio, ninit, maxrel
    Uo, io, ho, _ = orc.solve(oop, (N, N, 1), b.astype(np.float32), oo, Ut=u, dtype=np.float32)
    print(f"skip {skip}: A^T products gpu {info['n_matvec_T']} oracle {io.n_matvec_T} of "
          f"{io.total_iters} iterations; rel(gpu,orc32)={rel(U, Uo):.2e} err {info['rel_error']:.3e}")
    assert info["total_iters"] == io.total_iters
    # the GPU count also holds the product A^T b
    assert info["n_matvec_T"] - 1 == io.n_matvec_T
    if maxrel == 0.0:
        assert io.n_matvec_T == io.total_iters + 1 + io.fallbacks    # nothing skipped
    else:
        assert io.n_matvec_T < io.total_iters                       # skips happened
    assert (hist["cnt"] == ho["cnt"]).all()
    assert rel(U, Uo) < 1e-4


def test_plain_fp64_matches_fp64_oracle(ctx):
    N = 12
    u, g, dyn, A, oop, b = problem_3d(ctx, N, E=16, R=40)
    rp, c, v = A.download_csr()
    A64 = api.SparseOperator(ctx, A.m, A.n, rp, c, v, dtype="f64")
    o = api.default_opts(maxit=60, tol_inn=1e-3, mu=2.0 ** 9, mu0=2.0 ** 9)
    U, info, hist = api.tval3_gpu_3d(ctx, A64, b.astype(np.float64), (N, N, N), o, Ut=u)
    U64, i64, h64, _ = orc.solve(oop, (N, N, N), b, oracle_opts(o), Ut=u)
    print(f"fp64: iters {info['total_iters']}/{i64.total_iters} rel={rel(U, U64):.2e}")
    assert info["total_iters"] == i64.total_iters
    assert rel(U, U64) < 1e-9


def test_plain_dynamic_siddon_solve(ctx):
    N = 12
    u, g, dyn, A, oop, b = problem_3d(ctx, N, E=16, R=40)
    o = api.default_opts(maxit=40, tol_inn=1e-3, mu=2.0 ** 9, mu0=2.0 ** 9)
    Ud, infod, _ = api.tval3_gpu_3d(ctx, dyn, b.astype(np.float32), (N, N, N), o, Ut=u)
    Uc, infoc, _ = api.tval3_gpu_3d(ctx, A, b.astype(np.float32), (N, N, N), o, Ut=u)
    # same weights, only the A^T summation order differs (atomics vs CSC)
    assert rel(Ud, Uc) < 1e-3
    assert abs(infod["rel_error"] - infoc["rel_error"]) < 1e-3


def test_device_solve_and_fallback_branch(ctx):
    import torch
    N = 8
    u, A, b = problem_2d(N, 0.6, seed=3)
    b = wl.add_noise(A @ u, 0.05, seed=14)
    op = api.DenseOperator(ctx, A)
    o = api.default_opts(maxit=30, gamma=1.0, c=0.0, gam=0.0, tol_inn=1e-12)
    bt = torch.from_numpy(b.astype(np.float32)).cuda()
    U, info, hist = api.tval3_gpu_3d(ctx, op, bt, (N, N, 1), o)
    oo = orc.default_opts(maxit=30, gamma=1.0, c=0.0, gam=0.0, tol_inn=1e-12)
    Uo, io, ho, _ = orc.solve(orc.dense_op(A.astype(np.float32).astype(np.float64)), (N, N, 1),
                              b.astype(np.float32), oo, dtype=np.float32)
    assert info["total_iters"] == 30
    assert set(hist["cnt"].tolist()) <= {0, 6}
    assert info["fallbacks"] == io.fallbacks > 0
    assert rel(U.cpu().numpy(), Uo) < 1e-3


def test_solver_error_paths(ctx):
    N = 8
    u, A, b = problem_2d(N)
    op = api.DenseOperator(ctx, A)
    with pytest.raises(TV3Error) as e:
        api.tval3_gpu_3d(ctx, op, b.astype(np.float32), (N, N, 2))
    assert e.value.status == 2
    bb = b.astype(np.float32).copy()
    bb[0] = np.inf
    with pytest.raises(TV3Error) as e:
        api.tval3_gpu_3d(ctx, op, bb, (N, N, 1))
    assert e.value.status == 6
    with pytest.raises(TV3Error) as e:      # the fused solver takes CSR operators
        api.tval3_gpu_3d(ctx, op, b.astype(np.float32), (N, N, 1),
                         api.default_opts(solver="fused"))
    assert e.value.status == 8
    with pytest.raises(TV3Error) as e:
        api.tval3_gpu_3d(ctx, op, b.astype(np.float32), (N, N, 1),
                         api.default_opts(skipMulRatio=-1.0))
    assert e.value.status == 1
    U, info, _ = api.tval3_gpu_3d(ctx, op, b.astype(np.float32), (N, N, 1),
                                  api.default_opts(maxit=0))
    assert info["total_iters"] == 0
