"""Dense multi-right-hand-side products on the matrix cores
(tv3_matvec_multi) and batched problems sharing one dense BF16 operator.

- Y = A X and Y = A^T X against FP64 products of the same BF16 matrix: the
  FP32 operand enters as three BF16 terms (about 24 significant bits) and
  the products accumulate in FP32, so the error is that of an FP32 product:
  |Y - Y64| <= 4 sqrt(K) 2^-24 |A| |X| elementwise; several shapes, column
  counts below, at and above 64, and leading dimensions above the length.
- Operators the matrix-core path does not take (sizes not multiples of 256,
  other weight types) run tv3_matvec per column: equal to it bitwise.
- A batch holding sparse problems and problems that share a dense operator:
  every problem bitwise equal to the same problem solved alone in a batch
  (same launch shapes); dense problems against the FP32 oracle with the
  fused solver's envelope criterion (rounding-level agreement at the start,
  then within a small multiple of the oracle's own drift under 1e-7
  perturbations of b, three seeds) and the error against u_true.
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


def bf16_round(a):
    v = np.ascontiguousarray(a, np.float32).view(np.uint32)
    r = ((v + 0x7FFF + ((v >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.view(np.float32).astype(np.float64)


def cols(torch, length, nrhs, ld=None, seed=0):
    ld = ld or length
    g = torch.Generator(device="cuda").manual_seed(seed)
    buf = torch.randn(nrhs, ld, generator=g, device="cuda", dtype=torch.float32)
    return buf[:, :length].t()           # (length, nrhs), columns contiguous, stride ld


@pytest.mark.parametrize("m,n", [(256, 512), (1024, 2048), (8192, 32768 // 8)])
def test_multi_rhs_products_against_fp64(ctx, m, n):
    import torch
    A = bf16_round(wl.dense_gaussian(m, n, seed=m + n))
    op = api.DenseOperator(ctx, A, dtype="bf16")
    At = torch.from_numpy(A).cuda()
    for trans in (False, True):
        xl, yl = (m, n) if trans else (n, m)
        M = At.t() if trans else At
        for nrhs, ld in ((1, None), (7, xl + 64), (64, None), (100, xl + 128)):
            X = cols(torch, xl, nrhs, ld, seed=nrhs)
            Y = op.mat_mat_mul(X, trans=trans)
            Y64 = M @ X.double()
            scale = M.abs() @ X.double().abs()
            bound = 4 * np.sqrt(xl) * 2.0 ** -24
            err = ((Y.double() - Y64).abs() / scale).max().item()
            rel = ((Y.double() - Y64).norm() / Y64.norm()).item()
            print(f"m={m} n={n} trans={trans} nrhs={nrhs}: max |err|/(|A||X|) = {err:.2e} "
                  f"(bound {bound:.1e}), rel L2 {rel:.2e}")
            assert err <= bound, err


def test_multi_rhs_fallback_paths_equal_matvec(ctx):
    import torch
    # 300 rows: not a multiple of 256 -> per-column products
    A = wl.dense_gaussian(300, 640, seed=3)
    for dt in ("bf16", "f32"):
        op = api.DenseOperator(ctx, A, dtype=dt)
        X = cols(torch, 640, 5, seed=1)
        Y = op.mat_mat_mul(X)
        for j in range(5):
            y = op.mat_vec_mul(X[:, j].contiguous())
            assert torch.equal(Y[:, j], y)
    # a sparse operator: per-column products
    g = wl.transducer_geometry(8, 16, N=8)
    S = api.DynamicOperator(ctx, g, "siddon").materialize("f32")
    X = cols(torch, S.n, 3, seed=2)
    Y = S.mat_mat_mul(X)
    for j in range(3):
        assert torch.equal(Y[:, j], S.mat_vec_mul(X[:, j].contiguous()))


def _dense_group(ctx, G, seed=40):
    """G problems of 16 x 8 x 8 voxels sharing one 256 x 1024 Gaussian BF16
    operator; problem j: phantom seed + j, its own noise."""
    import torch
    dims = (16, 8, 8)
    n = 1024
    A = bf16_round(wl.dense_gaussian(256, n, seed=seed))
    op = api.DenseOperator(ctx, A, dtype="bf16")
    out = []
    for j in range(G):
        u = wl.phantom_3d(dims=dims, seed=seed + j)
        b = wl.add_noise(A @ u, 0.005, seed=seed + 100 + j).astype(np.float32)
        out.append(dict(dims=dims, u=u, A=op, Ad=A, b=b, bt=torch.from_numpy(b).cuda(),
                        ut=torch.from_numpy(u.astype(np.float32)).cuda(), dense=True))
    return out


def _sparse(ctx, dims, E, R, seed):
    import torch
    u = wl.phantom_3d(dims=dims, seed=seed)
    g = wl.transducer_geometry(E, R, dims=dims, seed=seed + 1)
    A = api.DynamicOperator(ctx, g, "siddon").materialize("bf16")
    b = A.mat_vec_mul(torch.from_numpy(u.astype(np.float32)).cuda()).cpu().numpy()
    b = wl.add_noise(b.astype(np.float64), 0.005, seed=seed + 2).astype(np.float32)
    return dict(dims=dims, u=u, A=A, b=b, bt=torch.from_numpy(b).cuda(),
                ut=torch.from_numpy(u.astype(np.float32)).cuda())


def _solve(ctx, probs, o, history=True):
    return api.tval3_gpu_3d_batched(ctx, [p["A"] for p in probs], [p["bt"] for p in probs],
                                    [p["dims"] for p in probs], o, Uts=[p["ut"] for p in probs],
                                    history=history, lanes=(8, 8), splits=(4, 2))


def _first_divergence(h1, h2):
    n = min(len(h1["cnt"]), len(h2["cnt"]))
    for k in range(n):
        if h1["cnt"][k] != h2["cnt"][k] or h1["outer"][k] != h2["outer"][k]:
            return k
    return n


def test_batch_with_shared_dense_operator(ctx):
    group = _dense_group(ctx, 6)
    sparse = [_sparse(ctx, (16, 16, 16), 24, 64, 1), _sparse(ctx, (8, 16, 8), 16, 40, 5)]
    # interleaved caller order: the solver groups the dense problems itself
    probs = [group[0], sparse[0], group[1], group[2], sparse[1], group[3], group[4], group[5]]
    o = api.default_opts(mu=2.0 ** 9, mu0=2.0 ** 9, tv_mode="iso", maxit=50, tol_inn=1e-3)
    Ub, ib, bi, hb = _solve(ctx, probs, o)
    print("batch info", bi)
    assert bi["dense_groups"] == 1 and bi["splits_N"] == 4 and bi["splits_T"] == 2
    for k, p in enumerate(probs):
        Us, is_, _, hs = _solve(ctx, [p], o)
        assert ib[k]["total_iters"] == is_[0]["total_iters"] > 0
        assert np.array_equal(hb[k]["f"], hs[0]["f"])
        assert np.array_equal(Ub[k].cpu().numpy(), Us[0].cpu().numpy()), k
    oo = orc.default_opts(mu=2.0 ** 9, mu0=2.0 ** 9, tol_inn=1e-3, maxit=50, tv_iso=1)
    for k, p in enumerate(probs):
        if not p.get("dense"):
            continue
        oop = orc.dense_op(p["Ad"])
        Uo, io, ho, _ = orc.solve(oop, p["dims"], p["b"], oo, Ut=p["u"], dtype=np.float32)
        # the oracle's own sensitivity: three rounding-level perturbations of b
        runs = []
        for seed in (0, 1, 2):
            xi = np.random.default_rng(seed).choice([-1.0, 1.0], size=p["b"].shape)
            bp = (p["b"].astype(np.float64) * (1 + 1e-7 * xi)).astype(np.float32)
            runs.append(orc.solve(oop, p["dims"], bp, oo, Ut=p["u"], dtype=np.float32))
        ip = runs[0][1]
        n = min([_first_divergence(hb[k], ho)] + [_first_divergence(r[2], ho) for r in runs])
        dev = np.abs(hb[k]["f"][:n] - ho["f"][:n]) / np.abs(ho["f"][:n])
        env = np.max([np.maximum.accumulate(np.abs(r[2]["f"][:n] - ho["f"][:n])
                                            / np.abs(ho["f"][:n])) for r in runs], axis=0)
        spread = max(abs(r[1].rel_error - io.rel_error) for r in runs)
        print(f"dense problem {k}: iters {ib[k]['total_iters']} oracle {io.total_iters}; "
              f"decisions agree for {n}; err {ib[k]['rel_error']:.4e} oracle {io.rel_error:.4e} "
              f"perturbed {ip.rel_error:.4e}; f dev max {dev.max():.1e} envelope {env.max():.1e}")
        assert n >= 10
        assert (dev[:4] < 1e-5).all()
        # on dense Gaussian problems the fused schedule's iterates show early
        # transients near 100x the perturbation drift, through the sparse
        # (CSR) products of the same matrix as much as through the matrix-core
        # ones, hence 300 here against 30 for the sparse problems
        assert (dev <= 300 * env + 1e-6).all()
        assert abs(ib[k]["rel_error"] - io.rel_error) <= max(3 * spread, 1e-3)


def test_batch_dense_operator_errors(ctx):
    import torch
    A = wl.dense_gaussian(256, 1024, seed=1)
    opf = api.DenseOperator(ctx, A, dtype="f32")
    b = torch.zeros(256, device="cuda")
    o = api.default_opts(maxit=5)
    with pytest.raises(TV3Error):      # dense batches take BF16
        api.tval3_gpu_3d_batched(ctx, [opf], [b], [(16, 8, 8)], o)
    op = api.DenseOperator(ctx, wl.dense_gaussian(200, 1024, seed=1), dtype="bf16")
    with pytest.raises(TV3Error):      # rows not a multiple of 256
        api.tval3_gpu_3d_batched(ctx, [op], [torch.zeros(200, device="cuda")], [(16, 8, 8)], o)
