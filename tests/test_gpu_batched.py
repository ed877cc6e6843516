"""Batched TVAL3 (tv3_solve_batched): many problems of uneven size in one set
of launches per iteration.

- No cross-talk: a problem solved inside a batch gives bitwise the result,
  iteration count and history of the same problem solved alone (same launch
  shapes), and graph size does not change results.
- Each problem against the FP32 oracle with the envelope criterion of the
  fused solver (the batched path runs the same algorithm with FP32 products
  in its own summation order): rounding-level agreement at the start, then
  within a small multiple of the oracle's own sensitivity to 1e-7
  perturbations of b (three seeds), and the final error.
- Backtracking, fallbacks and the outer loop inside a batch; error paths.
- The A^T-skip heuristic decides per problem: each problem against the
  oracle running the same rule, batch against solo bitwise, and a rule that
  never skips against the run without it, bitwise.
- The context keeps the last batch (layouts, launch shapes, graph): a call
  with the same operators, shapes and captured options and new data reuses
  it and gives the bits a fresh build gives.
"""
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


def make(ctx, dims, E, R, seed, dtype="f32", noise=0.005):
    import torch
    u = wl.phantom_3d(dims=dims, seed=seed)
    g = wl.transducer_geometry(E, R, dims=dims, seed=seed + 1)
    dyn = api.DynamicOperator(ctx, g, "siddon")
    A = dyn.materialize(dtype)
    A32 = dyn.materialize("f32")
    rp, c, v = A32.download_csr()
    oop = orc.csr_op(A32.m, A32.n, rp, c, v)
    b = wl.add_noise(orc.matvec(oop, 0, u), noise, seed=seed + 2).astype(np.float32)
    return dict(dims=dims, u=u, A=A, oop=oop, b=b, bt=torch.from_numpy(b).cuda(),
                ut=torch.from_numpy(u.astype(np.float32)).cuda())


PROBS = [((16, 16, 16), 24, 64, 1), ((12, 10, 6), 16, 48, 4), ((8, 20, 12), 20, 40, 7),
         ((24, 8, 16), 24, 56, 10)]


@pytest.fixture(scope="module")
def batch(ctx):
    return [make(ctx, *pp) for pp in PROBS]


# every dimension a multiple of 8: update_W on 8^3 tiles for every problem
PROBS8 = [((16, 16, 16), 24, 64, 21), ((16, 8, 24), 16, 48, 24), ((8, 8, 8), 16, 32, 27),
          ((24, 16, 8), 20, 56, 30)]


@pytest.fixture(scope="module")
def batch8(ctx):
    return [make(ctx, *pp) for pp in PROBS8]


# even tile counts along every axis: A's panels may span 16^3 voxels
PROBS16 = [((16, 16, 16), 24, 64, 41), ((32, 16, 16), 20, 56, 44), ((16, 32, 16), 16, 48, 47)]


@pytest.fixture(scope="module")
def batch16(ctx):
    return [make(ctx, *pp) for pp in PROBS16]


def opts(**kw):
    base = dict(mu=2.0 ** 9, mu0=2.0 ** 9, tol_inn=1e-3, tv_mode="iso", maxit=60)
    base.update(kw)
    return api.default_opts(**base)


def solve(ctx, probs, o, lanes=(8, 8), history=True, panels=-1, panels_A=-1, box=None):
    return api.tval3_gpu_3d_batched(ctx, [p["A"] for p in probs], [p["bt"] for p in probs],
                                    [p["dims"] for p in probs], o, Uts=[p["ut"] for p in probs],
                                    history=history, lanes=lanes, panels=panels,
                                    panels_A=panels_A, panel_box_A=box)


@pytest.mark.parametrize("lanes,panels,panels_A", [((8, 8), -1, -1), ((-2048, -1024), -1, -1),
                                                   ((-2048, -1024), 2048, -1),
                                                   ((-2048, -1024), 2048, 1024)])
def test_batch_equals_solo_bitwise(ctx, batch, lanes, panels, panels_A):
    """row-group products (8 lanes per row), row-block products (blocks of
    2048 / 1024 entries, which must start at every problem's first row) and
    A^T and A in panel form; the batch mixes the B = 8, 4 and 1 numberings,
    so update_W runs on tiles for some problems and per voxel for the others"""
    o = opts()
    Ub, ib, bi, hb = solve(ctx, batch, o, lanes=lanes, panels=panels, panels_A=panels_A)
    assert (bi["panels_T"], bi["panels_A"]) == (max(panels, 0), max(panels_A, 0))
    print("batch info", bi)
    for k, p in enumerate(batch):
        Us, is_, _, hs = solve(ctx, [p], o, lanes=lanes, panels=panels, panels_A=panels_A)
        assert ib[k]["total_iters"] == is_[0]["total_iters"] > 0
        assert ib[k]["outer_iters"] == is_[0]["outer_iters"]
        assert (hb[k]["cnt"] == hs[0]["cnt"]).all()
        assert np.array_equal(hb[k]["f"], hs[0]["f"])
        assert np.array_equal(Ub[k].cpu().numpy(), Us[0].cpu().numpy()), k
    assert bi["total_iters"] == sum(i["total_iters"] for i in ib)


def test_batch_graph_size_and_lanes(ctx, batch):
    o1 = opts(graph_iters=1)
    o8 = opts(graph_iters=8)
    U1, i1, _, _ = solve(ctx, batch, o1, history=False)
    U8, i8, b8, _ = solve(ctx, batch, o8, history=False)
    for k in range(len(batch)):
        assert np.array_equal(U1[k].cpu().numpy(), U8[k].cpu().numpy())
        assert i1[k]["total_iters"] == i8[k]["total_iters"]
    assert b8["graph_launches"] > 0
    # other launch shapes: the same iterates up to FP32 summation order, which
    # TVAL3's BB step amplifies over the iterations (rounding level at the
    # start, final errors within the spread such small problems show)
    _, i8h, _, h8 = solve(ctx, batch, o8)
    Ut, it, bt, ht = solve(ctx, batch, o8, lanes=None)
    print("tuned lanes", bt["lanes_N"], bt["lanes_T"])
    shapes = (2, 4, 8, 16, 32, 64, -1024, -2048, -4096)
    assert bt["lanes_N"] in shapes and bt["lanes_T"] in shapes
    for k in range(len(batch)):
        f8, ft = h8[k]["f"][:4], ht[k]["f"][:4]
        assert (np.abs(ft - f8) <= 1e-5 * np.abs(f8)).all()
        assert abs(it[k]["rel_error"] - i8h[k]["rel_error"]) < 2e-2


def first_divergence(h1, h2):
    """first iteration whose decisions differ: backtracking count, outer
    index, A^T product taken or skipped"""
    n = min(len(h1["cnt"]), len(h2["cnt"]))
    for k in range(n):
        if (h1["cnt"][k] != h2["cnt"][k] or h1["outer"][k] != h2["outer"][k]
                or h1["mulT"][k] != h2["mulT"][k]):
            return k
    return n


@pytest.mark.parametrize("which,panels,box", [("batch", -1, None), ("batch8", 2048, None),
                                              ("batch16", 2048, 8)])
def test_batch_tracks_oracle(ctx, request, which, panels, box):
    """(batch8 with A^T and A in panel form; batch16 with A's panels over
    16^3 voxels)"""
    batch = request.getfixturevalue(which)
    o = opts()
    Ub, ib, bi, hb = solve(ctx, batch, o, panels=panels, panels_A=panels, box=box)
    assert bi["panels_T"] == bi["panels_A"] == max(panels, 0)
    assert bi["panel_box_A"] == (box or (1 if panels > 0 else 0))
    oo = orc.default_opts(mu=2.0 ** 9, mu0=2.0 ** 9, tol_inn=1e-3, maxit=60, tv_iso=1)
    for k, p in enumerate(batch):
        Uo, io, ho, _ = orc.solve(p["oop"], p["dims"], p["b"], oo, Ut=p["u"], dtype=np.float32)
        # the oracle's own sensitivity: three rounding-level perturbations of b
        # (trajectory envelope and spread of the final error: the largest of
        # the three; one perturbation can underestimate the drift tenfold)
        runs = []
        for seed in (0, 1, 2):
            xi = np.random.default_rng(seed).choice([-1.0, 1.0], size=p["b"].shape)
            bp = (p["b"].astype(np.float64) * (1 + 1e-7 * xi)).astype(np.float32)
            runs.append(orc.solve(p["oop"], p["dims"], bp, oo, Ut=p["u"], dtype=np.float32))
        Up, ip, hpp, _ = runs[0]
        spread = max(abs(r[1].rel_error - io.rel_error) for r in runs)
        kk = min([first_divergence(hb[k], ho)] + [first_divergence(r[2], ho) for r in runs])
        dev = np.abs(hb[k]["f"][:kk] - ho["f"][:kk]) / np.abs(ho["f"][:kk])
        env = np.max([np.maximum.accumulate(np.abs(r[2]["f"][:kk] - ho["f"][:kk])
                                            / np.abs(ho["f"][:kk])) for r in runs], axis=0)
        Ug = Ub[k].cpu().numpy()
        print(f"{p['dims']}: iters batch {ib[k]['total_iters']} oracle {io.total_iters}; "
              f"decisions agree for {kk}; err batch {ib[k]['rel_error']:.4e} oracle "
              f"{io.rel_error:.4e} perturbed {ip.rel_error:.4e}; rel(U batch, oracle) "
              f"{rel(Ug, Uo):.2e} rel(U perturbed, oracle) {rel(Up, Uo):.2e}")
        assert kk >= 10
        assert (dev[:4] < 1e-5).all()
        assert (dev <= 30 * env + 1e-6).all()
        # U within a small multiple of the oracle's own spread; the final
        # error then differs by at most |U - U_oracle| / |u_true| (triangle
        # inequality), which can exceed the error spread of the perturbed runs
        assert rel(Ug, Uo) <= max(10 * max(rel(r[0], Uo) for r in runs), 1e-4)
        bound = rel(Ug, Uo) * np.linalg.norm(Uo) / np.linalg.norm(p["u"])
        assert abs(ib[k]["rel_error"] - io.rel_error) <= max(3 * spread, 1e-3, 1.01 * bound)


def test_batch_outer_loop_backtracking_fallback(ctx, batch):
    # continuation over several outer iterations
    o = opts(maxit=120, tol_inn=5e-3, mu0=2.0 ** 6, beta0=4.0)
    Ub, ib, _, hb = solve(ctx, batch, o)
    for k in range(len(batch)):
        assert ib[k]["outer_iters"] >= 2
        assert (np.diff(hb[k]["outer"]) >= 0).all()
        assert hb[k]["mu"][-1] > hb[k]["mu"][0]
    # Armijo constant 0, gamma 1: every failing test ends in the fallback
    noisy = []
    for p in batch[:2]:
        import torch
        b = wl.add_noise(p["b"].astype(np.float64), 0.05, seed=14).astype(np.float32)
        q = dict(p)
        q["bt"] = torch.from_numpy(b).cuda()
        noisy.append(q)
    o = opts(maxit=40, gamma=1.0, c=0.0, gam=0.0, tol_inn=1e-12)
    Ub, ib, _, hb = solve(ctx, noisy, o)
    for k in range(len(noisy)):
        assert ib[k]["total_iters"] == 40
        assert set(hb[k]["cnt"].tolist()) <= {0, 6}
        assert ib[k]["fallbacks"] == int((hb[k]["cnt"] == 6).sum())
    assert sum(i["fallbacks"] for i in ib) > 0


def test_batch_errors(ctx, batch):
    import torch
    p16 = make(ctx, (8, 8, 8), 16, 32, 3, dtype="f16")
    with pytest.raises(TV3Error):        # mixed weight types
        solve(ctx, [batch[0], p16], opts())
    with pytest.raises(TV3Error):        # dims do not match the operator
        api.tval3_gpu_3d_batched(ctx, [batch[0]["A"]], [batch[0]["bt"]], [(16, 16, 15)], opts())
    with pytest.raises(TV3Error):        # negative skip ratio
        solve(ctx, batch[:1], opts(skipMulRatio=-0.5))
    with pytest.raises(TV3Error):        # no such row-block size
        solve(ctx, batch[:1], opts(), panels=512)
    with pytest.raises(ValueError):
        api.tval3_gpu_3d_batched(ctx, [batch[0]["A"]], [batch[0]["bt"][:-1]], [(16, 16, 16)])
    dense = api.DenseOperator(ctx, np.ones((4, 512)))
    with pytest.raises(TV3Error):        # dense operators are not batched this way
        api.tval3_gpu_3d_batched(ctx, [dense], [torch.zeros(4, device="cuda")], [(8, 8, 8)])


def _skip_kw(skip):
    ratio, ninit, maxrel = skip
    return dict(skipMulRatio=ratio, numInitIterations=ninit, maxRelativeChange=maxrel)


@pytest.mark.parametrize("skip", [(2.0, 3, 1e30), (1.0, 0, 0.05)])
def test_batch_skip_heuristic_tracks_oracle(ctx, batch, skip):
    """A^T-skip heuristic (3d/tval3_gpu_3d.cpp:398-468; a skipped product
    leaves g as it was), one decision per problem: each problem of the batch
    against the FP32 oracle running the same rule, with the envelope
    criterion of test_batch_tracks_oracle (decisions, the product/skip
    choice included, agree for at least 10 iterations), the same number of
    A^T products whenever both runs take the same iterations, products
    actually skipped, and the final error; and each problem of the batch
    equals its solo solve bit for bit (no cross-talk between decisions)."""
    kw = _skip_kw(skip)
    o = opts(**kw)
    Ub, ib, _, hb = solve(ctx, batch, o)
    oo = orc.default_opts(mu=2.0 ** 9, mu0=2.0 ** 9, tol_inn=1e-3, maxit=60, tv_iso=1)
    for key, v in kw.items():
        setattr(oo, key, v)
    for k, p in enumerate(batch):
        Uo, io, ho, _ = orc.solve(p["oop"], p["dims"], p["b"], oo, Ut=p["u"], dtype=np.float32)
        runs = []
        for seed in (0, 1, 2):
            xi = np.random.default_rng(seed).choice([-1.0, 1.0], size=p["b"].shape)
            bp = (p["b"].astype(np.float64) * (1 + 1e-7 * xi)).astype(np.float32)
            runs.append(orc.solve(p["oop"], p["dims"], bp, oo, Ut=p["u"], dtype=np.float32))
        spread = max(abs(r[1].rel_error - io.rel_error) for r in runs)
        kk = min([first_divergence(hb[k], ho)] + [first_divergence(r[2], ho) for r in runs])
        dev = np.abs(hb[k]["f"][:kk] - ho["f"][:kk]) / np.abs(ho["f"][:kk])
        env = np.max([np.maximum.accumulate(np.abs(r[2]["f"][:kk] - ho["f"][:kk])
                                            / np.abs(ho["f"][:kk])) for r in runs], axis=0)
        same = (ib[k]["total_iters"] == io.total_iters and ib[k]["fallbacks"] == io.fallbacks
                and np.array_equal(hb[k]["outer"], ho["outer"]))
        print(f"skip {skip} {p['dims']}: iters batch {ib[k]['total_iters']} oracle "
              f"{io.total_iters}; A^T products batch {ib[k]['n_matvec_T'] - 1} oracle "
              f"{io.n_matvec_T}; decisions agree for {kk}; err batch {ib[k]['rel_error']:.4e} "
              f"oracle {io.rel_error:.4e}; max f dev {dev.max():.1e} envelope {env.max():.1e}")
        assert io.n_matvec_T < io.total_iters             # the rule skipped products
        assert ib[k]["n_matvec_T"] - 1 < ib[k]["total_iters"]
        assert kk >= 10
        assert (dev[:4] < 1e-5).all()
        assert (dev <= 30 * env + 1e-6).all()
        if same:
            # the batched count holds A^T b besides the oracle's products
            assert ib[k]["n_matvec_T"] - 1 == io.n_matvec_T
        assert abs(ib[k]["rel_error"] - io.rel_error) <= max(3 * spread, 1e-3)
        Us, is_, _, hs = solve(ctx, [p], o)
        assert np.array_equal(Ub[k].cpu().numpy(), Us[0].cpu().numpy()), k
        assert np.array_equal(hb[k]["mulT"], hs[0]["mulT"])
        assert ib[k]["n_matvec_T"] == is_[0]["n_matvec_T"]


def test_batch_skip_rule_that_never_skips_changes_nothing(ctx, batch):
    """maxRelativeChange = 0: every get_g multiplies, so the batch with the
    heuristic on equals the batch without it bit for bit, product counts
    included; with fallbacks (Armijo constant 0, gamma 1) the forced product
    of the fallback step goes through the same gate."""
    for extra in ({}, dict(maxit=40, gamma=1.0, c=0.0, gam=0.0, tol_inn=1e-12)):
        o0 = opts(**extra)
        o1 = opts(**extra, **_skip_kw((0.5, 2, 0.0)))
        U0, i0, _, h0 = solve(ctx, batch, o0)
        U1, i1, _, h1 = solve(ctx, batch, o1)
        for k in range(len(batch)):
            assert np.array_equal(U0[k].cpu().numpy(), U1[k].cpu().numpy()), (extra, k)
            assert np.array_equal(h0[k]["f"], h1[k]["f"])
            assert i0[k]["total_iters"] == i1[k]["total_iters"]
            assert i0[k]["n_matvec_T"] == i1[k]["n_matvec_T"]
            assert (h1[k]["mulT"] == 1).all()


@pytest.mark.parametrize("blk", [1024, 2048])
def test_batch_panels_equal_solo_bitwise(ctx, batch8, blk):
    """A^T of the sparse problems in panel form (tiles of 512 voxels, each
    tile's rays staged in LDS) and A in two-pass panel form (segment sums
    per tile, then per ray): each problem of the batch equals its solo
    solve bit for bit -- the problems stop at different iterations, and the
    panels of a stopped problem are skipped whole while the others run --
    and agrees with the row-kernel batch to rounding level at the start."""
    o = opts(maxit=120, tol_inn=5e-3, mu0=2.0 ** 6, beta0=4.0)   # continuation steps
    Ub, ib, bi, hb = solve(ctx, batch8, o, panels=blk, panels_A=blk)
    print("iterations", [i["total_iters"] for i in ib])
    assert bi["panels_T"] == bi["panels_A"] == blk
    assert len({i["total_iters"] for i in ib}) > 1       # problems stop apart
    for k, p in enumerate(batch8):
        Us, is_, bs, hs = solve(ctx, [p], o, panels=blk, panels_A=blk)
        assert bs["panels_T"] == bs["panels_A"] == blk
        assert ib[k]["total_iters"] == is_[0]["total_iters"] > 0
        assert np.array_equal(hb[k]["f"], hs[0]["f"])
        assert np.array_equal(Ub[k].cpu().numpy(), Us[0].cpu().numpy()), k
    # the row kernels sum in another order: rounding-level agreement at the
    # start (accuracy over the whole run: test_batch_tracks_oracle)
    Ur, ir, br, hr = solve(ctx, batch8, o, panels=-1, panels_A=-1)
    assert br["panels_T"] == br["panels_A"] == 0
    for k in range(len(batch8)):
        f, fr = hb[k]["f"][:4], hr[k]["f"][:4]
        assert (np.abs(f - fr) <= 1e-5 * np.abs(fr)).all()
    # timed choice (0): one of the block sizes or the row kernels, reported
    _, _, bt, _ = solve(ctx, batch8, o, panels=0, panels_A=0, history=False)
    assert bt["panels_T"] in (0, 1024, 2048, 4096) and bt["panels_A"] in (0, 1024, 2048)


@pytest.mark.parametrize("blk", [2048, 4096])
def test_batch_box_panels_equal_solo_bitwise(ctx, batch16, blk):
    """A in two-pass panel form over 16^3 voxels (the tiles of each problem
    regrouped by 2 x 2 x 2, x staged through a tile map): each problem of the
    batch equals its solo solve bit for bit (A^T's panels in row blocks of
    the same size), and the batch agrees with the one-tile panels to
    rounding level at the start; panels of 8 tiles are refused for a batch
    with an odd tile count."""
    o = opts(maxit=80, tol_inn=5e-3, mu0=2.0 ** 6, beta0=4.0)
    Ub, ib, bi, hb = solve(ctx, batch16, o, panels=blk, panels_A=blk, box=8)
    assert (bi["panels_T"], bi["panels_A"], bi["panel_box_A"]) == (blk, blk, 8)
    for k, p in enumerate(batch16):
        Us, is_, bs, hs = solve(ctx, [p], o, panels=blk, panels_A=blk, box=8)
        assert (bs["panels_T"], bs["panels_A"], bs["panel_box_A"]) == (blk, blk, 8)
        assert ib[k]["total_iters"] == is_[0]["total_iters"] > 0
        assert np.array_equal(hb[k]["f"], hs[0]["f"])
        assert np.array_equal(Ub[k].cpu().numpy(), Us[0].cpu().numpy()), k
    U1, i1, b1, h1 = solve(ctx, batch16, o, panels=2048, panels_A=2048)
    assert (b1["panels_A"], b1["panel_box_A"]) == (2048, 1)
    for k in range(len(batch16)):
        f, f1 = hb[k]["f"][:4], h1[k]["f"][:4]
        assert (np.abs(f - f1) <= 1e-5 * np.abs(f1)).all()
    _, _, bt, _ = solve(ctx, batch16, o, panels=0, panels_A=0, history=False)
    assert bt["panel_box_A"] in (0, 1, 8)
    with pytest.raises(TV3Error, match="even tile count"):
        solve(ctx, [make(ctx, (24, 16, 16), 8, 16, 50)], o, panels_A=2048, box=8)


def test_batch_backtracking_blends(ctx, batch):
    """Backtracking steps (1 <= cnt <= 5, alpha < 1): update_W reads the
    trial blended with the current iterate at alpha and stores the accepted
    U and g (apply keeps only Au).  A larger Armijo constant makes the search
    backtrack; each problem of the batch equals its solo solve bit for bit
    and tracks the FP32 oracle (same decisions, the oracle's envelope)."""
    o = opts(c=0.3, maxit=40)
    Ub, ib, _, hb = solve(ctx, batch, o, lanes=(-2048, -1024))
    cnts = np.concatenate([h["cnt"] for h in hb])
    print("cnt histogram", np.bincount(cnts, minlength=7).tolist())
    assert ((cnts >= 1) & (cnts <= 5)).any()
    oo = orc.default_opts(mu=2.0 ** 9, mu0=2.0 ** 9, tol_inn=1e-3, maxit=40, tv_iso=1, c=0.3)
    for k, p in enumerate(batch):
        Us, is_, _, hs = solve(ctx, [p], o, lanes=(-2048, -1024))
        assert np.array_equal(Ub[k].cpu().numpy(), Us[0].cpu().numpy()), k
        assert np.array_equal(hb[k]["f"], hs[0]["f"])
        Uo, io, ho, _ = orc.solve(p["oop"], p["dims"], p["b"], oo, Ut=p["u"], dtype=np.float32)
        runs = []
        for seed in (0, 1, 2):
            xi = np.random.default_rng(seed).choice([-1.0, 1.0], size=p["b"].shape)
            bp = (p["b"].astype(np.float64) * (1 + 1e-7 * xi)).astype(np.float32)
            runs.append(orc.solve(p["oop"], p["dims"], bp, oo, Ut=p["u"], dtype=np.float32))
        kk = min([first_divergence(hb[k], ho)] + [first_divergence(r[2], ho) for r in runs])
        dev = np.abs(hb[k]["f"][:kk] - ho["f"][:kk]) / np.abs(ho["f"][:kk])
        env = np.max([np.maximum.accumulate(np.abs(r[2]["f"][:kk] - ho["f"][:kk])
                                            / np.abs(ho["f"][:kk])) for r in runs], axis=0)
        print(f"{p['dims']}: decisions agree for {kk}; backtracks {ib[k]['backtracks']}; "
              f"max f dev {dev.max():.1e} envelope {env.max():.1e}")
        assert kk >= 10
        assert (dev[:4] < 1e-5).all()
        assert (dev <= 30 * env + 1e-6).all()


def test_batch_continuation_tracks_oracle(ctx, batch8):
    """beta0 < beta: after each outer step the first trial of the new inner
    loop uses W of the last update_W, built from the old sigma at the old
    beta, in every problem of the batch (same criterion as the fused
    solver)."""
    from test_gpu_fused import _envelope_runs, check_continuation_tracks
    kw = dict(mu=2.0 ** 9, mu0=2.0 ** 6, beta=32.0, beta0=4.0, tol_inn=5e-3, maxit=100)
    o = opts(**kw)
    Ub, ib, bi, hb = solve(ctx, batch8, o, panels=2048, panels_A=2048)
    oo = orc.default_opts(**kw, tv_iso=1)
    for k, p in enumerate(batch8):
        Uo, io, ho, _ = orc.solve(p["oop"], p["dims"], p["b"], oo, Ut=p["u"], dtype=np.float32)
        runs = _envelope_runs(p["oop"], p["dims"], p["b"], oo, p["u"])
        kk = check_continuation_tracks(hb[k], ho, runs, f"batch {p['dims']}")
        assert kk >= 15


def test_batch_session_reuse_equals_fresh_build(ctx, batch):
    """Three calls with the same operators and shapes: b, then 0.7 b (the
    session is reused), then 0.7 b with a different graph size (a captured
    option: the session is rebuilt; graph size does not change results).
    The reused call equals the rebuilt one bit for bit, and the first call
    equals a repeat of itself."""
    import torch
    o8, o4 = opts(graph_iters=8), opts(graph_iters=4)
    U1, i1, _, h1 = solve(ctx, batch, o8, lanes=(-2048, -1024), panels=2048, panels_A=1024)
    scaled = [dict(p, bt=(0.7 * p["bt"]).contiguous()) for p in batch]
    U2, i2, b2, h2 = solve(ctx, scaled, o8, lanes=(-2048, -1024), panels=2048, panels_A=1024)
    U3, i3, b3, h3 = solve(ctx, scaled, o4, lanes=(-2048, -1024), panels=2048, panels_A=1024)
    U4, i4, _, h4 = solve(ctx, batch, o4, lanes=(-2048, -1024), panels=2048, panels_A=1024)
    print("host seconds: reused", b2["host_secs"], "rebuilt", b3["host_secs"])
    for k in range(len(batch)):
        assert torch.equal(U2[k], U3[k])
        assert i2[k]["total_iters"] == i3[k]["total_iters"] > 0
        assert (h2[k]["f"] == h3[k]["f"]).all()
        assert torch.equal(U1[k], U4[k])
        assert (h1[k]["f"] == h4[k]["f"]).all()
        assert not torch.equal(U1[k], U2[k])            # the data did change
