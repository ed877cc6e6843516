"""GPU parity of the operators and stencil building blocks against the oracle,
through the C ABI.  Integer and index work is compared bit-exactly; FP32
products within the tolerance written in each test."""
import numpy as np
import pytest

import tv3_oracle as orc
from tval3_amd import api, workload as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available()
    return api.Context(0)


def T(a, dt=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dt is not None:
        t = t.to(dt)
    return t.cuda()


def H(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rand_csr(m, n, density, seed, integer=False, empty_rows=True):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(m, n)) * (rng.uniform(size=(m, n)) < density)
    if integer:
        A = np.round(A * 3)
    if empty_rows:
        A[rng.integers(0, m, max(1, m // 10))] = 0
    rows, cols = np.nonzero(A)
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))]).astype(np.int32)
    return A, rowptr, cols.astype(np.int32), A[rows, cols]


def test_csr_integer_bit_exact_all_lanes(ctx):
    import torch
    m, n = 700, 531
    A, rp, c, v = rand_csr(m, n, 0.08, 1, integer=True)
    op = api.SparseOperator(ctx, m, n, rp, c, v)
    x = np.random.default_rng(2).integers(-4, 5, n).astype(np.float32)
    r = np.random.default_rng(3).integers(-4, 5, m).astype(np.float32)
    # row-group kernel (L lanes per row) and row-blocked kernel (-NB entries)
    for L in (2, 4, 8, 16, 32, 64, -1024, -2048, -4096):
        op.set_lanes(L, L)
        y = H(op.mat_vec_mul(T(x)))
        z = H(op.mat_vec_mul(T(r), trans=True))
        assert (y == (A @ x).astype(np.float32)).all(), L
        assert (z == (A.T @ r).astype(np.float32)).all(), L


def test_csr_stream_long_rows_and_empty(ctx):
    # rows longer than a block, empty rows and a block of many short rows
    rng = np.random.default_rng(21)
    m, n = 40, 9000
    lens = np.array([0, 5000, 3, 0, 1200, 4100] + [7] * 34)
    rowptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = np.concatenate([rng.choice(n, l, replace=False) for l in lens]).astype(np.int32)
    val = rng.integers(-3, 4, len(col)).astype(np.float64)
    A = np.zeros((m, n))
    for i in range(m):
        A[i, col[rowptr[i]:rowptr[i + 1]]] = val[rowptr[i]:rowptr[i + 1]]
    op = api.SparseOperator(ctx, m, n, rowptr, col, val)
    x = rng.integers(-4, 5, n).astype(np.float32)
    r = rng.integers(-4, 5, m).astype(np.float32)
    for L in (-1024, -2048, -4096, 64):
        op.set_lanes(L, L)
        assert (H(op.mat_vec_mul(T(x))) == (A @ x).astype(np.float32)).all(), L
        assert (H(op.mat_vec_mul(T(r), trans=True)) == (A.T @ r).astype(np.float32)).all(), L


def test_csr_float_matches_oracle_and_given_csc(ctx):
    m, n = 1200, 900
    A, rp, c, v = rand_csr(m, n, 0.05, 4)
    oop = orc.csr_op(m, n, rp, c, v)
    op_dev = api.SparseOperator(ctx, m, n, rp, c, v)                  # CSC built on device
    cp, rw, cv = orc.csr_transpose(m, n, rp, c, v)
    op_host = api.SparseOperator(ctx, m, n, rp, c, v, csc=(cp, rw, cv))  # CSC from host
    x = np.random.default_rng(5).normal(size=n).astype(np.float32)
    r = np.random.default_rng(6).normal(size=m).astype(np.float32)
    y = H(op_dev.mat_vec_mul(T(x)))
    ref = orc.matvec(oop, 0, x)
    assert np.abs(y - ref).max() <= 1e-5 * np.abs(ref).max()
    z1 = H(op_dev.mat_vec_mul(T(r), trans=True))
    z2 = H(op_host.mat_vec_mul(T(r), trans=True))
    assert (z1 == z2).all()          # stable device transpose == host transpose, bitwise
    ref = orc.matvec(oop, 1, r)
    assert np.abs(z1 - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("dt,tol", [("f16", 2e-3), ("bf16", 1e-2)])
def test_csr_reduced_precision_weights(ctx, dt, tol):
    m, n = 800, 640
    A, rp, c, v = rand_csr(m, n, 0.05, 7)
    op = api.SparseOperator(ctx, m, n, rp, c, v, dtype=dt)
    x = np.random.default_rng(8).normal(size=n).astype(np.float32)
    y = H(op.mat_vec_mul(T(x)))
    ref = A @ x
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < tol


def test_csr_f64_adjoint(ctx):
    import torch
    m, n = 500, 400
    A, rp, c, v = rand_csr(m, n, 0.06, 9)
    op = api.SparseOperator(ctx, m, n, rp, c, v, dtype="f64")
    x = np.random.default_rng(10).normal(size=n)
    r = np.random.default_rng(11).normal(size=m)
    y = H(op.mat_vec_mul(T(x)))
    z = H(op.mat_vec_mul(T(r), trans=True))
    assert np.allclose(y, A @ x, rtol=1e-13, atol=1e-13)
    assert abs(y @ r - x @ z) < 1e-12 * abs(y @ r)


def test_csr_rejects_bad_input(ctx):
    from tval3_amd._lib import TV3Error
    rp = np.array([0, 2, 1], np.int32)
    with pytest.raises(TV3Error) as e:
        api.SparseOperator(ctx, 2, 3, rp, np.array([0, 1], np.int32), np.ones(2))
    assert e.value.status == 3
    rp = np.array([0, 1, 2], np.int32)
    with pytest.raises(TV3Error) as e:
        api.SparseOperator(ctx, 2, 3, rp, np.array([0, 5], np.int32), np.ones(2))
    assert e.value.status == 3
    with pytest.raises(TV3Error) as e:
        api.SparseOperator(ctx, 2, 3, rp, np.array([0, 1], np.int32), np.array([1.0, np.nan]))
    assert e.value.status == 6


def test_dense_gemv(ctx):
    m, n = 333, 1000
    A = wl.dense_gaussian(m, n, seed=12)
    op = api.DenseOperator(ctx, A)
    x = np.random.default_rng(13).normal(size=n).astype(np.float32)
    r = np.random.default_rng(14).normal(size=m).astype(np.float32)
    y = H(op.mat_vec_mul(T(x)))
    z = H(op.mat_vec_mul(T(r), trans=True))
    A32 = A.astype(np.float32).astype(np.float64)
    assert np.abs(y - A32 @ x).max() < 1e-5 * np.abs(A32 @ x).max()
    assert np.abs(z - A32.T @ r).max() < 1e-5 * np.abs(A32.T @ r).max()


def _geom(E, R, N, seed=2):
    return wl.transducer_geometry(E, R, N=N, seed=seed)


def test_bresenham_bit_exact(ctx):
    N = 20
    g = _geom(16, 24, N)
    oop = orc.bresenham_op(g)
    op = api.DynamicOperator(ctx, g, "bresenham")
    assert op.m == 16 * 24 and op.n == N ** 3
    rng = np.random.default_rng(15)
    x = rng.integers(-3, 4, op.n).astype(np.float32)
    r = rng.integers(-3, 4, op.m).astype(np.float32)
    # scale_factor = 2/N is a power of two times an integer for N = 20? use exact scale 0.25
    g2 = dict(g, scale=0.25)
    oop = orc.bresenham_op(g2)
    op = api.DynamicOperator(ctx, g2, "bresenham")
    y = H(op.mat_vec_mul(T(x)))
    z = H(op.mat_vec_mul(T(r), trans=True))
    assert (y == orc.matvec(oop, 0, x.astype(np.float64)).astype(np.float32)).all()
    assert (z == orc.matvec(oop, 1, r.astype(np.float64)).astype(np.float32)).all()
    # materialized CSR: same visited voxels (bit-exact), weights == scale
    A = op.materialize("f32")
    rp, c, v = A.download_csr()
    for k in (0, 7, 200, op.m - 1):
        e, rr = divmod(k, 24)
        vox = orc.bresenham_ray(g2["e_vox"][e], g2["r_vox"][rr], g2["rv"])
        assert (c[rp[k]:rp[k + 1]] == vox).all()
    assert (v == 0.25).all()


def _dense_of(rp, c, v, m, n):
    M = np.zeros((m, n))
    for i in range(m):
        np.add.at(M[i], c[rp[i]:rp[i + 1]], v[rp[i]:rp[i + 1]])
    return M


def test_siddon_materialize_matches_oracle(ctx):
    N = 10
    g = _geom(12, 20, N)
    oop = orc.siddon_op(g)
    orp, oc, ov = orc.siddon_csr(oop)
    op = api.DynamicOperator(ctx, g, "siddon")
    A = op.materialize("f32")
    rp, c, v = A.download_csr()
    M_gpu = _dense_of(rp, c, v, op.m, op.n)
    M_orc = _dense_of(orp, oc, ov, op.m, op.n)
    # FP32 traversal: per-entry error bounded by a few ulp of the ray length
    assert np.abs(M_gpu - M_orc).max() < 2e-5
    # row sums are the chord lengths
    assert np.allclose(M_gpu.sum(1), M_orc.sum(1), rtol=1e-5, atol=1e-6)
    # no spurious entries beyond tiny segments
    assert abs(A.nnz - len(oc)) <= max(3, len(oc) // 1000)


def test_siddon_dynamic_products_and_adjoint(ctx):
    N = 16
    g = _geom(24, 40, N)
    oop = orc.siddon_op(g)
    op = api.DynamicOperator(ctx, g, "siddon")
    rng = np.random.default_rng(16)
    x = rng.uniform(size=op.n).astype(np.float32)
    r = rng.normal(size=op.m).astype(np.float32)
    y = H(op.mat_vec_mul(T(x)))
    z = H(op.mat_vec_mul(T(r), trans=True))
    yr = orc.matvec(oop, 0, x.astype(np.float64))
    zr = orc.matvec(oop, 1, r.astype(np.float64))
    assert np.abs(y - yr).max() < 1e-5 * np.abs(yr).max()
    assert np.abs(z - zr).max() < 1e-5 * np.abs(zr).max()
    # adjoint test in FP64 on the device traversal
    y64 = H(op.mat_vec_mul(T(x.astype(np.float64))))
    z64 = H(op.mat_vec_mul(T(r.astype(np.float64)), trans=True))
    lhs, rhs = y64 @ r, x.astype(np.float64) @ z64
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_hybrid_operator(ctx):
    N = 12
    g = _geom(16, 24, N)
    hyb = api.HybridOperator(ctx, g, "siddon", "f32")
    dyn = api.DynamicOperator(ctx, g, "siddon")
    rng = np.random.default_rng(17)
    x = rng.uniform(size=dyn.n).astype(np.float32)
    r = rng.normal(size=dyn.m).astype(np.float32)
    # the same traversal; the hybrid sums each ray in one pass, the
    # matrix-free operator's tile products by segments (another FP32 order)
    y1, y2 = H(hyb.mat_vec_mul(T(x))), H(dyn.mat_vec_mul(T(x)))
    assert np.abs(y1 - y2).max() <= 1e-6 * np.abs(y2).max()
    z1 = H(hyb.mat_vec_mul(T(r), trans=True))
    z2 = H(dyn.mat_vec_mul(T(r), trans=True))
    assert np.abs(z1 - z2).max() < 1e-5 * np.abs(z2).max()


def test_D_Dt_shrink_match_oracle(ctx):
    dims = (17, 9, 6)
    n = int(np.prod(dims))
    rng = np.random.default_rng(18)
    U = rng.normal(size=n).astype(np.float32)
    Ux, Uy, Uz = (H(t) for t in ctx.D(T(U), dims))
    ox, oy, oz = orc.D(U, dims, np.float32)
    assert (Ux == ox).all() and (Uy == oy).all() and (Uz == oz).all()
    X, Y, Z = (rng.normal(size=n).astype(np.float32) for _ in range(3))
    r = H(ctx.Dt(T(X), T(Y), T(Z), dims))
    assert (r == orc.Dt(X, Y, Z, dims, np.float32)).all()
    S = [rng.normal(size=n).astype(np.float32) for _ in range(3)]
    for iso in (False, True):
        W = ctx.shrink(T(Ux), T(Uy), T(Uz), *[T(s) for s in S], 4.0, iso=iso)
        Wo = orc.shrink(Ux, Uy, Uz, *S, 4.0, iso=iso, dtype=np.float32)
        for a, b in zip(W[:3], Wo[:3]):
            assert np.abs(H(a) - b).max() <= 2e-7 * max(1, np.abs(b).max())
        assert np.isclose(W[3], Wo[3], rtol=1e-12)


@pytest.mark.parametrize("given", ["csr", "csc"])
def test_hybrid_transpose_follows_callers_matrix(ctx, given):
    """sparse_geom_mm(hA_geo, hA_mat) runs A x by the geometry and A^T x on
    hA_mat (3d/mat_vec_mul.h:104-120,732-756).  Here the geometry walks
    Bresenham rays while hA_mat holds other weights of the same shape
    (Siddon lengths, scaled and rounded to integers): A x must be the
    Bresenham product and A^T r the product with hA_mat, both bit for bit
    (integer weights and inputs make every sum exact)."""
    N = 12
    g = _geom(16, 24, N)
    sid = orc.siddon_op(g)
    rp, col, val = orc.siddon_csr(sid)
    val = np.maximum(np.round(val * 40.0), 1.0)          # integers 1 .. ~7
    m, n = len(rp) - 1, N ** 3
    kw = {"csr": (rp, col, val)} if given == "csr" else \
        {"csc": orc.csr_transpose(m, n, rp, col, val)}
    hyb = api.HybridOperator(ctx, g, "bresenham", "f32", **kw)
    dyn = api.DynamicOperator(ctx, g, "bresenham")
    assert (hyb.m, hyb.n) == (m, n)
    rng = np.random.default_rng(19)
    x = rng.integers(0, 5, size=n).astype(np.float32)
    r = rng.integers(-4, 5, size=m).astype(np.float32)
    assert (H(hyb.mat_vec_mul(T(x))) == H(dyn.mat_vec_mul(T(x)))).all()
    want = orc.matvec(orc.csr_op(m, n, rp, col, val), 1, r.astype(np.float64))
    got = H(hyb.mat_vec_mul(T(r), trans=True))
    assert (got == want).all()
    bres = orc.matvec(orc.bresenham_op(g), 1, r.astype(np.float64))
    assert not np.allclose(want, bres)                    # the two matrices differ
