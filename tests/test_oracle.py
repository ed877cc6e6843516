"""Pins the CPU oracle against closed forms, brute force and invariants.

The reference ships no tests, golden vectors or fixtures for this path, and
it cannot be built (CUDA texture references, missing headers), so every
oracle function is pinned here by mathematics that a plausible mistake (a
dropped term, a wrong sign or axis, a transposed operand) would violate.
"""
import numpy as np
import pytest

from tval3_amd import workload as wl


def rnd(shape, seed):
    return np.random.default_rng(seed).normal(size=shape)


# ---------------------------------------------------------------- ray list
def test_build_paths_matches_enumeration(orc):
    E, R, mps = 5, 7, 3
    rng = np.random.default_rng(0)
    mask = (rng.uniform(size=E * R) < 0.6).astype(np.uint8)
    pe, pr = orc.build_paths(E, R, mps, mask)
    exp = [(mp * E + e, mp * R + r) for mp in range(mps) for e in range(E) for r in range(R)
           if mask[e * R + r]]
    assert list(zip(pe.tolist(), pr.tolist())) == exp
    pe, pr = orc.build_paths(E, R, 1, None)
    assert len(pe) == E * R and pe[R] == 1 and pr[R] == 0


# ---------------------------------------------------------------- Bresenham
def _coords(vox, rv):
    rvx, rvy, rvz = rv
    z = vox // (rvx * rvy)
    x = (vox % (rvx * rvy)) // rvy
    y = vox % rvy
    return np.stack([x, y, z], 1)


def test_bresenham_axis_and_diagonal(orc):
    rv = (16, 16, 16)
    v = orc.bresenham_ray((2, 3, 4), (12, 3, 4), rv)
    c = _coords(v, rv)
    assert len(v) == 11 and (c[:, 0] == np.arange(2, 13)).all()
    assert (c[:, 1] == 3).all() and (c[:, 2] == 4).all()
    v = orc.bresenham_ray((1, 1, 1), (9, 9, 9), rv)
    c = _coords(v, rv)
    assert len(v) == 9 and (c == np.arange(1, 10)[:, None]).all()
    # backwards along z (z is the fast axis: dz > dx, dz > dy)
    c = _coords(orc.bresenham_ray((5, 5, 10), (5, 7, 2), rv), rv)
    assert len(c) == 9 and (c[:, 2] == np.arange(10, 1, -1)).all()


def test_bresenham_closed_form(orc):
    """Every step of the walk of 3d/tval3_gpu_3d_kernels.cu:1672-1759 puts the
    slow coordinates at ceil(j*m_k/m1 - 1/2) (round half down), and the fast
    axis follows the >= cascade (x, then y, then z) of :1679,:1687."""
    rv = (40, 40, 40)
    rng = np.random.default_rng(1)
    for _ in range(400):
        a = rng.integers(0, 40, 3)
        b = rng.integers(0, 40, 3)
        if rng.uniform() < 0.3:                # force ties in |d|
            b[1] = np.clip(a[1] + (b[0] - a[0]), 0, 39)
        c = _coords(orc.bresenham_ray(a, b, rv), rv)
        d = b - a
        m = np.abs(d)
        if m[0] >= m[1] and m[0] >= m[2]:
            order = (0, 1, 2)
        elif m[1] >= m[0] and m[1] >= m[2]:
            order = (1, 0, 2)
        else:
            order = (2, 0, 1)
        f, s2, s3 = order
        m1 = m[f]
        assert len(c) == m1 + 1
        j = np.arange(m1 + 1)
        assert (c[:, f] == a[f] + np.sign(d[f] if d[f] else 1) * j).all()
        for s in (s2, s3):
            if m1 == 0:
                continue
            inc = -1 if d[s] < 0 else 1
            # ceil((2 j m - m1) / (2 m1)) in integers, clipped at 0 for j = 0
            k = -((-(2 * j * m[s] - m1)) // (2 * m1))
            k[0] = 0
            assert (c[:, s] == a[s] + inc * k).all(), (a, b)


# ---------------------------------------------------------------- Siddon
def _chord(p0, p1, lo, hi):
    d = p1 - p0
    amin, amax = 0.0, 1.0
    for a in range(3):
        if d[a] == 0:
            if not (lo[a] < p0[a] < hi[a]):
                return 0.0
            continue
        t0, t1 = sorted(((lo[a] - p0[a]) / d[a], (hi[a] - p0[a]) / d[a]))
        amin, amax = max(amin, t0), min(amax, t1)
    return max(0.0, amax - amin) * np.linalg.norm(d)


def test_siddon_chord_length(orc):
    rv = (7, 9, 5)
    origin = np.array([-1.0, -0.5, -0.8])
    h = np.array([0.3, 0.2, 0.35])
    hi = origin + np.array(rv) * h
    rng = np.random.default_rng(2)
    for _ in range(300):
        p0 = rng.uniform(-2.5, 2.5, 3)
        p1 = rng.uniform(-2.5, 2.5, 3)
        vox, ln = orc.siddon_ray(p0, p1, origin, h, rv)
        assert np.isclose(ln.sum(), _chord(p0, p1, origin, hi), rtol=1e-12, atol=1e-12)
        assert (ln > 0).all()
        assert (vox >= 0).all() and (vox < np.prod(rv)).all()


def test_siddon_axis_ray_through_centres(orc):
    rv = (6, 5, 4)
    origin = np.zeros(3)
    h = np.array([0.5, 0.25, 1.0])
    p0 = np.array([-1.0, 2.5 * 0.25, 1.5])       # y centre of voxel 2, z centre of voxel 1
    p1 = np.array([10.0, 2.5 * 0.25, 1.5])
    vox, ln = orc.siddon_ray(p0, p1, origin, h, rv)
    assert len(vox) == 6 and np.allclose(ln, 0.5, rtol=0, atol=1e-14)
    c = _coords(vox, rv)
    assert (c[:, 0] == np.arange(6)).all() and (c[:, 1] == 2).all() and (c[:, 2] == 1).all()


def test_siddon_supersampled_line_integral(orc):
    """Per-voxel lengths against brute-force sampling of the segment."""
    rng = np.random.default_rng(3)
    for N in (4, 6, 8):
        rv = (N, N, N)
        origin = -np.ones(3)
        h = np.full(3, 2.0 / N)
        for _ in range(10):
            p0, p1 = wl.sphere_points(2, 1.1, seed=int(rng.integers(1 << 30)))
            vox, ln = orc.siddon_ray(p0, p1, origin, h, rv)
            K = 400000
            t = (np.arange(K) + 0.5) / K
            P = p0[None] + t[:, None] * (p1 - p0)[None]
            idx = np.floor((P - origin) / h).astype(int)
            ok = ((idx >= 0) & (idx < N)).all(1)
            lin = idx[ok, 2] * N * N + idx[ok, 0] * N + idx[ok, 1]
            L = np.linalg.norm(p1 - p0)
            bf = np.bincount(lin, minlength=N ** 3) * (L / K)
            got = np.zeros(N ** 3)
            np.add.at(got, vox, ln)
            assert np.abs(got - bf).max() < 3 * L / K + 1e-12


def test_siddon_csr_matches_rays(orc):
    g = wl.transducer_geometry(6, 10, N=8)
    op = orc.siddon_op(g)
    rowptr, col, val = orc.siddon_csr(op)
    assert rowptr[-1] == len(col)
    for k in (0, 13, 59):
        e, r = k // 10, k % 10
        vox, ln = orc.siddon_ray(g["e_pos"][e], g["r_pos"][r], g["origin"], g["voxel"], g["rv"])
        assert (col[rowptr[k]:rowptr[k + 1]] == vox).all()
        assert np.allclose(val[rowptr[k]:rowptr[k + 1]], ln)


# ---------------------------------------------------------------- products
def test_dense_and_csr_matvec_brute_force(orc):
    rng = np.random.default_rng(4)
    m, n = 23, 31
    A = rng.normal(size=(m, n)) * (rng.uniform(size=(m, n)) < 0.3)
    x, y = rng.normal(size=n), rng.normal(size=m)
    for op in (orc.dense_op(A), None):
        if op is None:
            rows, cols = np.nonzero(A)
            rowptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))])
            op = orc.csr_op(m, n, rowptr, cols, A[rows, cols])
        assert np.allclose(orc.matvec(op, 0, x), A @ x, rtol=1e-13, atol=1e-13)
        assert np.allclose(orc.matvec(op, 1, y), A.T @ y, rtol=1e-13, atol=1e-13)
        # integer-valued inputs: FP32 result is exact
        Ai = np.round(A * 4)
    rows, cols = np.nonzero(Ai)
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))])
    op = orc.csr_op(m, n, rowptr, cols, Ai[rows, cols])
    xi = rng.integers(-5, 6, n).astype(np.float32)
    assert (orc.matvec(op, 0, xi, np.float32) == (Ai @ xi).astype(np.float32)).all()


def test_csr_transpose_round_trip(orc):
    rng = np.random.default_rng(5)
    m, n = 40, 17
    A = rng.normal(size=(m, n)) * (rng.uniform(size=(m, n)) < 0.25)
    rows, cols = np.nonzero(A)
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))])
    cp, rw, cv = orc.csr_transpose(m, n, rowptr, cols, A[rows, cols])
    B = np.zeros((n, m))
    for j in range(n):
        seg = slice(cp[j], cp[j + 1])
        assert (np.diff(rw[seg]) > 0).all()            # rows ascending within a column
        B[j, rw[seg]] = cv[seg]
    assert (B == A.T).all()
    rp2, c2, v2 = orc.csr_transpose(n, m, cp, rw, cv)
    assert (rp2 == rowptr).all() and (c2 == cols).all() and (v2 == A[rows, cols]).all()


def test_bresenham_and_siddon_adjoint(orc):
    g = wl.transducer_geometry(5, 9, N=6)
    for op in (orc.bresenham_op(g), orc.siddon_op(g)):
        x = rnd(op.s.n, 6)
        y = rnd(op.s.m, 7)
        lhs = orc.matvec(op, 0, x) @ y
        rhs = x @ orc.matvec(op, 1, y)
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1)


# ---------------------------------------------------------------- stencils
def _vol(U, dims):
    p, q, r = dims
    return U.reshape(r, q, p)          # [z, x, y]


def test_D_closed_form_and_adjoint(orc):
    dims = (5, 4, 3)
    U = rnd(60, 8)
    Ux, Uy, Uz = orc.D(U, dims)
    V = _vol(U, dims)
    assert np.allclose(_vol(Uy, dims), np.roll(V, -1, 2) - V)
    assert np.allclose(_vol(Ux, dims), np.roll(V, -1, 1) - V)
    assert np.allclose(_vol(Uz, dims), np.roll(V, -1, 0) - V)
    X, Y, Z = rnd(60, 9), rnd(60, 10), rnd(60, 11)
    lhs = Ux @ X + Uy @ Y + Uz @ Z
    rhs = U @ orc.Dt(X, Y, Z, dims)
    assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + 1)
    c = np.full(60, 3.25)
    assert all((d == 0).all() for d in orc.D(c, dims))
    # each difference sums to zero along its own axis
    assert np.allclose(_vol(Ux, dims).sum(1), 0)


def test_DtD_is_periodic_laplacian(orc):
    dims = (6, 5, 4)
    p, q, r = dims
    U = rnd(p * q * r, 12)
    L = orc.Dt(*orc.D(U, dims), dims)
    lam = sum(np.meshgrid(2 - 2 * np.cos(2 * np.pi * np.arange(r) / r),
                          2 - 2 * np.cos(2 * np.pi * np.arange(q) / q),
                          2 - 2 * np.cos(2 * np.pi * np.arange(p) / p), indexing="ij"))
    ref = np.real(np.fft.ifftn(lam * np.fft.fftn(_vol(U, dims))))
    assert np.allclose(_vol(L, dims), ref, atol=1e-12)


# ---------------------------------------------------------------- shrinkage
def test_shrink_aniso_closed_form(orc):
    n = 500
    U = [rnd(n, 20 + k) for k in range(3)]
    S = [rnd(n, 30 + k) for k in range(3)]
    beta = 4.0
    Wx, Wy, Wz, lam1 = orc.shrink(*U, *S, beta, iso=False)
    exp = []
    for u, s in zip(U, S):
        ub = u - s / beta
        exp.append(np.sign(ub) * np.maximum(np.abs(ub) - 1 / beta, 0))
    for w, e in zip((Wx, Wy, Wz), exp):
        assert np.allclose(w, e, atol=1e-15)
    assert np.isclose(lam1, sum(np.abs(e).sum() for e in exp))
    # dead zone
    z = np.full(n, 0.2)
    W = orc.shrink(z, -z, z, 0 * z, 0 * z, 0 * z, 4.0)
    assert all((w == 0).all() for w in W[:3]) and W[3] == 0


def test_shrink_iso_closed_form(orc):
    n = 500
    U = [rnd(n, 40 + k) for k in range(3)]
    S = [rnd(n, 50 + k) for k in range(3)]
    beta = 2.0
    Wx, Wy, Wz, lam1 = orc.shrink(*U, *S, beta, iso=True)
    Ub = np.stack([u - s / beta for u, s in zip(U, S)])
    nrm = np.linalg.norm(Ub, axis=0)
    fac = np.maximum(nrm - 1 / beta, 0) / np.where(nrm > 0, nrm, 1)
    for k, w in enumerate((Wx, Wy, Wz)):
        assert np.allclose(w, Ub[k] * fac, atol=1e-15)
    assert np.isclose(lam1, np.maximum(nrm - 1 / beta, 0).sum())
    # one non-zero component: iso == aniso
    z = np.zeros(n)
    a = orc.shrink(U[0], z, z, S[0], z, z, beta, iso=True)
    b = orc.shrink(U[0], z, z, S[0], z, z, beta, iso=False)
    assert np.allclose(a[0], b[0]) and np.isclose(a[3], b[3])


def test_scaleb_closed_form(orc):
    b = np.array([0.1, -0.05, 0.2, 0.15])
    s, scl = orc.scaleb(b)
    # smallest |b| is -0.05, largest 0.2: range 0.25 -> scaled to 0.5
    assert np.isclose(scl, 2.0) and np.allclose(s, 2 * b)
    b = np.array([3.0, -1.0, 0.5])
    s, scl = orc.scaleb(b)                      # range 3.0 - 0.5 = 2.5 -> 1.5
    assert np.isclose(scl, 1.5 / 2.5)
    s, scl = orc.scaleb(np.array([1.0, 2.0]))   # range 1.0: unchanged
    assert scl == 1.0


# ---------------------------------------------------------------- whole loop
def test_solve_noise_free_well_conditioned(orc):
    """m = n Gaussian A, no noise: TVAL3 recovers a piecewise-constant image."""
    N = 16
    u = wl.phantom_2d(N, 4, seed=7)
    A = wl.dense_gaussian(N * N, N * N, seed=8)
    b = A @ u
    op = orc.dense_op(A)
    o = orc.default_opts(mu=2.0 ** 12, mu0=2.0 ** 12, beta=2.0 ** 7, beta0=2.0 ** 7,
                         tol=1e-6, tol_inn=1e-4, maxit=600)
    U, info, hist, st = orc.solve(op, (N, N, 1), b, o, Ut=u)
    assert st == 0
    assert info.rel_error < 2e-3, info.rel_error
    assert info.total_iters <= 600 and info.outer_iters <= o.maxcnt


def test_skip_heuristic_never_skipping_equals_baseline(orc):
    """A^T-skip rule (3d/tval3_gpu_3d.cpp:421-467) with maxRelativeChange = 0:
    Au always changes, so every product is taken and the iterates equal the
    run without the heuristic bit for bit, with the same product count."""
    N = 12
    u = wl.phantom_2d(N, 4, seed=21)
    A = wl.dense_gaussian(90, N * N, seed=22)
    b = wl.add_noise(A @ u, 1e-3, seed=23)
    base = dict(maxit=60, tol_inn=1e-3)
    Ub, ib, hb, _ = orc.solve(orc.dense_op(A), (N, N, 1), b, orc.default_opts(**base),
                              dtype=np.float32)
    Us, is_, hs, _ = orc.solve(orc.dense_op(A), (N, N, 1), b,
                               orc.default_opts(**base, skipMulRatio=0.5, numInitIterations=2,
                                                maxRelativeChange=0.0), dtype=np.float32)
    assert (Ub == Us).all() and (hb["f"] == hs["f"]).all()
    assert ib.n_matvec_T == is_.n_matvec_T == ib.total_iters + 1 + ib.fallbacks


def test_skip_heuristic_skips_and_still_recovers(orc):
    """With no change limit the rule skips on a fixed schedule: after the
    forced first product and numInitIterations products, skipMulRatio = 2
    gives one product then two skips.  A skipped product keeps g =
    A^T(last Au) - Atb (the reference would subtract Atb a second time), so
    the noise-free well-conditioned problem is still recovered."""
    N = 16
    u = wl.phantom_2d(N, 4, seed=7)
    A = wl.dense_gaussian(N * N, N * N, seed=8)
    b = A @ u
    o = orc.default_opts(mu=2.0 ** 12, mu0=2.0 ** 12, beta=2.0 ** 7, beta0=2.0 ** 7,
                         tol=1e-6, tol_inn=1e-4, maxit=600, skipMulRatio=2.0,
                         numInitIterations=3, maxRelativeChange=1e30)
    U, info, hist, st = orc.solve(orc.dense_op(A), (N, N, 1), b, o, Ut=u)
    assert st == 0
    calls = info.total_iters + 1 + info.fallbacks       # get_g calls
    assert info.n_matvec_T < 0.6 * calls
    assert info.rel_error < 5e-3, info.rel_error


@pytest.mark.parametrize("ratio,period", [(2.0, [1, 0, 0]), (1.0, [1, 0]), (0.5, [1, 1, 0])])
def test_skip_heuristic_schedule_closed_form(orc, ratio, period):
    """maxRelativeChange = 1e30: the decisions follow the counters alone
    (3d/tval3_gpu_3d.cpp:437-466).  The skip value starts at 0 after the
    forced first product; each of the first numInitIterations calls
    multiplies; then a call multiplies while the value is below 1 (adding the
    ratio) and otherwise skips (subtracting 1).  Ratio 2 gives 1,0,0 repeated,
    ratio 1 gives 1,0 and ratio 0.5 gives 1,1,0.  A fallback (cnt 6) is a
    forced product that resets the value, after a rejected trial that also
    counted: the closed form is checked on the rows before the first one,
    which cover at least a full period after the initial products; the row
    of a fallback multiplies."""
    N = 16
    u = wl.phantom_2d(N, 4, seed=7)
    A = wl.dense_gaussian(N * N, N * N, seed=8)
    b = wl.add_noise(A @ u, 1e-3, seed=9)
    ninit = 3
    o = orc.default_opts(maxit=40, tol_inn=1e-9, maxcnt=1, skipMulRatio=ratio,
                         numInitIterations=ninit, maxRelativeChange=1e30)
    U, info, hist, st = orc.solve(orc.dense_op(A), (N, N, 1), b, o, dtype=np.float32)
    assert st == 0
    fb = np.flatnonzero(hist["cnt"] == 6)
    k = int(fb[0]) if len(fb) else len(hist["cnt"])
    assert k >= ninit + len(period) + 1, k
    want = [1] * ninit + [period[j % len(period)] for j in range(k - ninit)]
    assert hist["mulT"][:k].tolist() == want
    assert (hist["mulT"][fb] == 1).all()
    # every decision is counted: the forced initial product, one per row, one
    # more per rejected trial that multiplied
    assert info.n_matvec_T >= 1 + int(hist["mulT"].sum())


def test_solve_nonmonotone_invariants(orc):
    N = 16
    u = wl.phantom_2d(N, 5, seed=9)
    A = wl.dense_gaussian(100, N * N, seed=10)
    b = wl.add_noise(A @ u, 1e-3, seed=11)
    o = orc.default_opts(maxit=150, tol_inn=1e-3, tol=1e-5)
    U, info, hist, st = orc.solve(orc.dense_op(A), (N, N, 1), b, o, Ut=u)
    assert st == 0 and info.total_iters == len(hist["f"]) <= 150
    assert (U >= 0).all()                                 # nonneg projection
    assert info.outer_iters <= o.maxcnt
    # mu, beta follow the continuation scheme and never exceed their caps
    assert (hist["mu"] <= o.mu).all() and (hist["beta"] <= o.beta).all()
    assert np.all(np.diff(hist["outer"]) >= 0)


def test_solve_maxit_cap_and_fallback_branch(orc):
    """With gamma = 1 the backtracking never shrinks alpha, so an iteration
    either passes the test |f| > |C - alpha c| at once or takes the cnt==5
    fallback (3d/tval3_gpu_3d.cpp:848,875-883).  The test uses absolute
    values, so a huge c makes it accept rather than fail."""
    N = 8
    u = wl.phantom_2d(N, 3, seed=12)
    A = wl.dense_gaussian(40, N * N, seed=13)
    b = wl.add_noise(A @ u, 0.05, seed=14)
    o = orc.default_opts(maxit=40, gamma=1.0, c=0.0, gam=0.0, tol_inn=1e-12)
    U, info, hist, st = orc.solve(orc.dense_op(A), (N, N, 1), b, o)
    assert info.total_iters == 40 == len(hist["cnt"])
    assert set(hist["cnt"].tolist()) <= {0, 6}
    assert info.fallbacks == int((hist["cnt"] == 6).sum()) > 0
    assert info.backtracks == 5 * info.fallbacks


def test_solve_rejects_nonfinite(orc):
    A = wl.dense_gaussian(10, 16, seed=1)
    b = np.ones(10)
    b[3] = np.nan
    U, info, hist, st = orc.solve(orc.dense_op(A), (4, 4, 1), b)
    assert st == 6


def test_solve_fp32_tracks_fp64_config0(orc):
    """Config 0 (2-D 64x64, dense A, m = 0.4 N, 0.1% noise, aniso TV,
    mu=2^8, beta=2^5, tol 1e-4, tol_inn 1e-3, maxit 300): FP32 vs FP64."""
    N = 64
    u = wl.phantom_2d(N, 8)
    A = wl.dense_gaussian(1638, N * N)
    b = wl.add_noise(A @ u, 1e-3)
    op = orc.dense_op(A)
    o = orc.default_opts(tol=1e-4, tol_inn=1e-3, maxit=300)
    U64, i64, h64, _ = orc.solve(op, (N, N, 1), b, o, Ut=u)
    U32, i32, h32, _ = orc.solve(op, (N, N, 1), b, o, Ut=u, dtype=np.float32)
    rel = np.linalg.norm(U32 - U64) / np.linalg.norm(U64)
    assert i64.rel_error < 0.2
    assert rel < 5e-3, rel


# ------------------------------------------------ whole-solve pins (orc_solve)
# Independent numpy restatements (periodic differences by np.roll, shrinkage
# as soft thresholding) used as the checker of the oracle's whole solve.
def _np_D(u3):
    """(Ux, Uy, Uz) of u3[z, x, y]: periodic forward differences"""
    return (np.roll(u3, -1, 1) - u3, np.roll(u3, -1, 2) - u3, np.roll(u3, -1, 0) - u3)


def _np_Dt(X, Y, Z):
    return (np.roll(X, 1, 1) - X) + (np.roll(Y, 1, 2) - Y) + (np.roll(Z, 1, 0) - Z)


def _np_shrink(X, Y, Z, t, iso):
    if iso:
        nr = np.sqrt(X * X + Y * Y + Z * Z)
        f = np.where(nr > t, (nr - t) / np.where(nr > 0, nr, 1.0), 0.0)
        return X * f, Y * f, Z * f
    s = lambda a: np.sign(a) * np.maximum(np.abs(a) - t, 0.0)
    return s(X), s(Y), s(Z)


def _huber_tv_grad(u3, b3, beta, mu, iso):
    """gradient of F(u) = sum_i min_w |w| + beta/2 |D_i u - w|^2 + mu/2 |u - b|^2,
    the objective of TVAL3's first inner loop for A = I (sigma = delta = 0)"""
    X, Y, Z = _np_D(u3)
    Wx, Wy, Wz = _np_shrink(X, Y, Z, 1.0 / beta, iso)
    return beta * _np_Dt(X - Wx, Y - Wy, Z - Wz) + mu * (u3 - b3)


def _fista_nonneg(b3, beta, mu, iso, iters=20000):
    """projected accelerated gradient (Beck-Teboulle 2009) on F over u >= 0;
    |D^T D| <= 12 bounds the Lipschitz constant"""
    L = 12.0 * beta + mu
    x = b3.copy()
    y = x.copy()
    t = 1.0
    for _ in range(iters):
        xn = np.maximum(y - _huber_tv_grad(y, b3, beta, mu, iso) / L, 0.0)
        tn = (1.0 + np.sqrt(1.0 + 4.0 * t * t)) / 2.0
        y = xn + (t - 1.0) / tn * (xn - x)
        x, t = xn, tn
    return x


def _kkt_residual(u3, b3, beta, mu, iso):
    """natural residual |u - P_{u>=0}(u - grad F/L)| L / (mu |b|)"""
    L = 12.0 * beta + mu
    G = _huber_tv_grad(u3, b3, beta, mu, iso)
    return float(np.linalg.norm(u3 - np.maximum(u3 - G / L, 0.0)) * L / (mu * np.linalg.norm(b3)))


def _identity_op(orc, n):
    ip = np.arange(n + 1, dtype=np.int32)
    return orc.csr_op(n, n, ip, np.arange(n, dtype=np.int32), np.ones(n))


@pytest.mark.parametrize("N,iso,offset", [(8, 1, 0.2), (8, 0, 0.2), (10, 1, -0.3), (10, 0, -0.3)])
def test_solve_tv_denoising_optimality(orc, N, iso, offset):
    """A = I, one outer iteration (mu0 = mu, beta0 = beta, so sigma = delta = 0):
    the inner loop of 3d/tval3_gpu_3d.cpp:1161-1202 minimises
    F(u) = sum_i |D_i u|_beta + mu/2 |u - b|^2 over u >= 0, the Huber-smoothed
    TV (iso or aniso) denoising problem.  Its solution satisfies the projected
    optimality condition and equals an independent accelerated projected
    gradient solution of F.  b is chosen in the range scaleb leaves
    unchanged; offset -0.3 makes the nonnegativity constraint active."""
    dims = (N, N, N)
    u = wl.phantom_3d(dims=dims, seed=3, contrast=0.5)
    u = (u - u.min()) / (u.max() - u.min()) * 0.8 + offset
    b = u + 0.05 * np.random.default_rng(5).standard_normal(u.size)
    assert orc.scaleb(b)[1] == 1.0
    mu, beta = 16.0, 32.0
    o = orc.default_opts(mu=mu, mu0=mu, beta=beta, beta0=beta, maxcnt=1, tol_inn=1e-10,
                         maxit=3000, tv_iso=iso)
    U, info, hist, st = orc.solve(_identity_op(orc, N ** 3), dims, b, o)
    assert st == 0
    u3, b3 = U.reshape(N, N, N), b.reshape(N, N, N)
    ref = _fista_nonneg(b3, beta, mu, iso)
    assert _kkt_residual(ref, b3, beta, mu, iso) < 1e-12
    kkt = _kkt_residual(u3, b3, beta, mu, iso)
    err = np.linalg.norm(u3 - ref) / np.linalg.norm(ref)
    print(f"N={N} iso={iso}: {info.total_iters} iterations, KKT {kkt:.1e}, |U - U_fista| {err:.1e}, "
          f"zeros {(U == 0).mean():.3f}")
    assert kkt < 1e-7 and err < 1e-7
    if offset < 0:
        assert (U == 0).any() and ((u3 == 0) == (ref == 0)).all()


def _cs_problem(N=32, frac=0.35, seed=31):
    u = wl.phantom_2d(N, 6, seed=seed)
    m = int(frac * N * N)
    A = wl.dense_gaussian(m, N * N, seed=seed + 1)
    return u, A, A @ u


_CS_OPTS = dict(mu=2.0 ** 12, mu0=2.0 ** 8, beta=2.0 ** 7, beta0=2.0 ** 4, tol=1e-7,
                tol_inn=1e-4, maxit=3000, maxcnt=30)


@pytest.mark.parametrize("iso", [0, 1])
def test_solve_underdetermined_recovery_with_continuation(orc, iso):
    """m = 0.35 n Gaussian measurements of a piecewise-constant image, no
    noise: the null space of A is large (the least-squares solution is off by
    ~0.8), TV minimisation subject to Au = b picks the image.  The run uses
    the continuation scheme (mu0 < mu, beta0 < beta, 3d/tval3_gpu_3d.cpp:
    1222-1249) and the multiplier updates, so it needs the outer loop right."""
    u, A, b = _cs_problem()
    N = 32
    U, info, hist, st = orc.solve(orc.dense_op(A), (N, N, 1), b,
                                  orc.default_opts(**_CS_OPTS, tv_iso=iso), Ut=u)
    lsq = np.linalg.lstsq(A, b, rcond=None)[0]
    e_lsq = np.linalg.norm(lsq - u) / np.linalg.norm(u)
    res = np.linalg.norm(A @ U - b) / np.linalg.norm(b)
    print(f"iso={iso}: rel error {info.rel_error:.2e} (least squares {e_lsq:.2f}), residual "
          f"{res:.1e}, {info.total_iters} iterations, {info.outer_iters} outer")
    assert st == 0 and e_lsq > 0.5
    assert info.rel_error < 3e-3 and res < 1e-4
    assert hist["beta"][0] == 2.0 ** 4 and hist["beta"][-1] == 2.0 ** 7


@pytest.mark.parametrize("iso", [0, 1])
def test_solve_multiplier_state_identity(orc, iso):
    """update_mlp (3d/tval3_gpu_3d.cpp:1034-1083) sets sigma -= beta V and
    delta -= mu (Au - b); the outer step then sets DtsAtd = beta0/beta d
    (:1240-1249).  With d = D^T V + mu/beta A^T(Au - b) + DtsAtd this keeps
    DtsAtd = -(D^T sigma + A^T delta)/beta at every outer step, so the final
    state of a run with continuation must satisfy it (sign, ratio and
    multiplier updates all enter)."""
    u, A, b = _cs_problem()
    N = 32
    o = orc.default_opts(**dict(_CS_OPTS, maxcnt=8), tv_iso=iso)
    U, info, h, st = orc.solve(orc.dense_op(A), (N, N, 1), b, o, state=True)
    assert st == 0 and info.outer_iters >= 6
    beta = h["beta"][-1]
    sig = h["sigma"].reshape(3, 1, N, N)
    want = -(_np_Dt(sig[0], sig[1], sig[2]).reshape(-1) + A.T @ h["delta"]) / beta
    err = np.linalg.norm(want - h["DtsAtd"]) / np.linalg.norm(h["DtsAtd"])
    print(f"iso={iso}: |DtsAtd + (D^T sigma + A^T delta)/beta| rel {err:.1e}")
    assert np.linalg.norm(h["DtsAtd"]) > 0 and err < 1e-9


@pytest.mark.parametrize("iso", [0, 1])
def test_solve_history_f_closed_form(orc, iso):
    """update_W (3d/tval3_gpu_3d.cpp:965-1029) and update_mlp (:1034-1083)
    keep f incrementally (subtract the old terms, add the new); every recorded
    f must equal lam1 + beta/2 lam2 + mu/2 lam3 - lam4 - lam5 of the same
    row."""
    u, A, b = _cs_problem()
    U, info, h, st = orc.solve(orc.dense_op(A), (32, 32, 1), b,
                               orc.default_opts(**dict(_CS_OPTS, maxcnt=6), tv_iso=iso))
    fc = h["lam1"] + h["beta"] / 2 * h["lam2"] + h["mu"] / 2 * h["lam3"] - h["lam4"] - h["lam5"]
    assert len(h["f"]) > 50 and len(set(h["beta"].tolist())) > 1
    assert np.max(np.abs(fc - h["f"]) / np.abs(h["f"])) < 1e-12


@pytest.mark.parametrize("case", ["cs", "noisy3d", "fallback"])
def test_solve_nonmonotone_acceptance_and_reference_recursion(orc, case):
    """Default gam (0.9995): C is the Zhang-Hager weighted average
    Q_k = gam Q_{k-1} + 1, C_k = (gam Q_{k-1} C_{k-1} + f_k)/Q_k
    (3d/tval3_gpu_3d.cpp:1176-1177), reset to Q = 1, C = f at each outer step
    (:1253), with gam shrunk by rate_gam on each cnt == 5 fallback (:877).
    Recomputed from the recorded f, cnt and outer index; and every accepted
    step satisfies the Armijo test of min_u (:848, :875):
    |f(alpha)| <= |C - alpha c| with alpha = gamma^cnt and
    c = |d|^2 tau c beta, while a backtracked step failed it at alpha = 1.
    Case "fallback" takes gamma = 1, so every failed test ends in the
    fallback and gam shrinks (and is reset by the next outer step)."""
    if case in ("cs", "fallback"):
        u, A, b = _cs_problem(seed=41)
        op, dims = orc.dense_op(A), (32, 32, 1)
        o = orc.default_opts(**dict(_CS_OPTS, maxcnt=8))
        if case == "fallback":
            o.gamma = 1.0
    else:
        dims = (12, 12, 12)
        u = wl.phantom_3d(dims=dims, seed=4, contrast=0.5)
        A = wl.dense_gaussian(700, u.size, seed=6)
        b = wl.add_noise(A @ u, 0.02, seed=7)
        op = orc.dense_op(A)
        o = orc.default_opts(maxit=200, tol_inn=1e-4, tv_iso=1, beta0=4.0, mu0=64.0)
    U, info, h, st = orc.solve(op, dims, b, o)
    assert st == 0
    cnt, outer = h["cnt"], h["outer"]
    if case == "fallback":                   # fallbacks, then a later outer loop
        fb = np.flatnonzero(cnt == 6)
        assert len(fb) >= 2 and outer[-1] > outer[fb[0]]
    else:
        assert (cnt[cnt < 6] > 0).any()      # the test exercises backtracking
    Q = C = gam = None
    for k in range(len(cnt)):
        if k == 0 or outer[k] != outer[k - 1]:
            Q, C, gam = 1.0, h["C_in"][k], o.gam
        else:
            assert h["C_in"][k] == C
        if cnt[k] == 6:
            gam *= o.rate_gam
        assert h["gam"][k] == gam
        Qn = gam * Q + 1
        C = (gam * Q * C + h["f"][k]) / Qn
        Q = Qn
        assert abs(h["C"][k] - C) <= 1e-13 * abs(C)
        c = h["c_armij"][k]
        if cnt[k] < 6:                       # a fallback row records the new tau
            assert c == h["dd"][k] * h["tau"][k] * o.c * h["beta"][k]
            al = o.gamma ** cnt[k]
            assert h["alpha"][k] == pytest.approx(al, rel=1e-15)
            assert abs(h["f_acc"][k]) <= abs(h["C_in"][k] - al * c)
            if cnt[k] == 0:
                assert h["f_trial"][k] == h["f_acc"][k]
            else:
                assert abs(h["f_trial"][k]) > abs(h["C_in"][k] - c)


# ------------------------------------------ tracked-plane Siddon traversal
def _geoms():
    yield wl.transducer_geometry(16, 24, N=12, seed=2)
    yield wl.transducer_geometry(64, 128, N=64, seed=2)
    yield wl.transducer_geometry(24, 40, dims=(20, 12, 8), seed=5)     # anisotropic voxels
    g = wl.transducer_geometry(8, 8, N=10, seed=3)                     # axis-parallel rays
    g["r_pos"] = g["e_pos"].copy()
    g["r_pos"][:, 0] *= -1.0
    g["e_pos"][:4, 1] = 0.3
    g["r_pos"][:4, 1] = 0.3
    yield g


def test_siddon_walk_fp64_equals_textbook_merge(orc):
    """orc_siddon_walk (tracked planes, the traversal of the device products)
    in FP64 against the textbook Siddon merge of sorted plane crossings
    (orc_siddon_ray): the same voxels in the same order and the same lengths
    to FP64 rounding, on every ray, including anisotropic voxels and rays
    parallel to two axes.  The two evaluate the plane parameters differently
    ((org + k h - p0) / d against fma(k, h/d, (org - p0)/d)), so lengths agree
    to the per-ray rounding bound 8 eps64 (|org| + N h + |p0|) max|1/d| L of
    the FP32 test below."""
    eps = 2.0 ** -53
    for g in _geoms():
        op = orc.siddon_op(g)
        ra, ca, va = orc.siddon_csr(op)
        rb, cb, vb = orc.siddon_walk_csr(op, np.float64)
        assert np.array_equal(ra, rb) and np.array_equal(ca, cb)
        pe, pr = orc.build_paths(g["E"], g["R"], g.get("mps", 1))
        ext = np.abs(np.asarray(g["origin"])) + np.asarray(g["rv"]) * np.asarray(g["voxel"])
        for k in range(len(ra) - 1):
            p0, p1 = g["e_pos"][pe[k]], g["r_pos"][pr[k]]
            d = p1 - p0
            inv = np.max(1.0 / np.maximum(np.abs(d), 1e-300))
            tol = 8 * eps * float(np.max(ext + np.abs(p0))) * inv * np.linalg.norm(d)
            s = slice(ra[k], ra[k + 1])
            assert np.all(np.abs(va[s] - vb[s]) <= tol), (k, tol)


def test_siddon_walk_fp32_tracks_fp64(orc):
    """The FP32 traversal (what the device computes) against the FP64 one:
    each ray's length sum (its chord through the volume) to FP32 accuracy; the
    voxels and lengths equal up to the FP32 rounding of the plane parameters
    t = fma(k, h/d, (org - p0)/d): per ray tol = 8 eps32 (|org| + N h + |p0|) max|1/d| L,
    which grows on grazing rays (small |d| along an axis); voxels whose
    length exceeds tol appear in both, with lengths within 2 tol."""
    eps = 2.0 ** -24
    for g in _geoms():
        op = orc.siddon_op(g)
        r64, c64, v64 = orc.siddon_walk_csr(op, np.float64)
        r32, c32, v32 = orc.siddon_walk_csr(op, np.float32)
        pe, pr = orc.build_paths(g["E"], g["R"], g.get("mps", 1))
        ext = np.abs(np.asarray(g["origin"])) + np.asarray(g["rv"]) * np.asarray(g["voxel"])
        for k in range(len(r64) - 1):
            p0, p1 = g["e_pos"][pe[k]], g["r_pos"][pr[k]]
            d = p1 - p0
            inv = np.max(1.0 / np.maximum(np.abs(d), 1e-30))
            tol = 8 * eps * float(np.max(ext + np.abs(p0))) * inv * np.linalg.norm(d)
            a, b = slice(r64[k], r64[k + 1]), slice(r32[k], r32[k + 1])
            m64 = dict(zip(c64[a].tolist(), v64[a].tolist()))
            m32 = dict(zip(c32[b].tolist(), v32[b].astype(np.float64).tolist()))
            for vox in set(m64) | set(m32):
                x, y = m64.get(vox, 0.0), m32.get(vox, 0.0)
                assert abs(x - y) <= 2 * tol + 1e-6 * x, (k, vox, x, y, tol)
            assert abs(v64[a].sum() - v32[b].sum()) <= 2 * tol * max(1, len(m64)) + 1e-6
