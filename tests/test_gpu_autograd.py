"""Differentiable building blocks and the unrolled TVAL3 on the GPU.

- Adjoint (dot-product) tests <Ax, y> = <x, A^T y> at 1e-6 relative for A,
  D and the shrinkage VJP's Jacobian pair.
- torch.autograd.gradcheck (FP64 central differences against the hand-written
  backward kernels) of each Function at small sizes.
- The shrinkage map against the plain solver's shrinkage entry point (bitwise)
  and the VJP against the autograd of a plain PyTorch FP64 restatement.
- The unrolled loop's forward against the FP64 oracle's solve with the same
  iteration counts (the oracle's stopping rules set so that it runs exactly
  those counts), decisions included.
- Gradients with respect to b and log(mu) against FP64 central finite
  differences (h = 1e-4) on seeded probes, evaluated on the recorded branch.
- Checkpointed and plain backward give the same gradient.
"""
import math

import numpy as np
import pytest

import tv3_oracle as orc
from tval3_amd import api, autograd as ad, workload as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available()
    return api.Context(0)


@pytest.fixture(scope="module")
def prob(ctx):
    """12^3 Siddon CSR problem (FP32 weights), its oracle operator, b, u_true."""
    N = 12
    u = wl.phantom_3d(N, 4, seed=5)
    g = wl.transducer_geometry(16, 40, N=N, seed=6)
    A = api.DynamicOperator(ctx, g, "siddon").materialize("f32")
    rp, c, v = A.download_csr()
    oop = orc.csr_op(A.m, A.n, rp, c, v)
    b = wl.add_noise(orc.matvec(oop, 0, u), 0.005, seed=7)
    return dict(dims=(N, N, N), u=u, A=A, oop=oop, b=b)


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def test_matvec_adjoint_and_gradcheck(ctx, prob):
    import torch
    A = prob["A"]
    gen = torch.Generator(device="cuda").manual_seed(11)
    for dt, tol in ((torch.float32, 1e-6), (torch.float64, 1e-12)):
        x = torch.randn(A.n, device="cuda", dtype=dt, generator=gen)
        y = torch.randn(A.m, device="cuda", dtype=dt, generator=gen)
        ax = ad.matvec(A, x)
        aty = ad.matvec(A, y, trans=True)
        lhs, rhs = float(torch.dot(ax.double(), y.double())), float(torch.dot(x.double(), aty.double()))
        assert _rel(lhs, rhs) < tol, (dt, lhs, rhs)
    x = torch.randn(A.n, device="cuda", dtype=torch.float64, generator=gen, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: ad.matvec(A, t), (x,), eps=1e-6, atol=1e-8,
                                    fast_mode=True)


def test_grad3_adjoint_and_gradcheck(ctx):
    import torch
    gen = torch.Generator(device="cuda").manual_seed(12)
    for dims in ((12, 12, 12), (7, 5, 3), (64, 1, 1), (1, 9, 2)):
        n = dims[0] * dims[1] * dims[2]
        for dt, tol in ((torch.float32, 1e-6), (torch.float64, 1e-13)):
            U = torch.randn(n, device="cuda", dtype=dt, generator=gen)
            V = torch.randn(3, n, device="cuda", dtype=dt, generator=gen)
            lhs = float(torch.dot(ad.grad3(ctx, U, dims).double().reshape(-1),
                                  V.double().reshape(-1)))
            rhs = float(torch.dot(U.double(), ad.grad3_t(ctx, V, dims).double()))
            assert _rel(lhs, rhs) < tol, (dims, dt, lhs, rhs)
    dims = (5, 4, 3)
    U = torch.randn(60, device="cuda", dtype=torch.float64, generator=gen, requires_grad=True)
    V = torch.randn(3, 60, device="cuda", dtype=torch.float64, generator=gen, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: ad.grad3(ctx, t, dims), (U,))
    assert torch.autograd.gradcheck(lambda t: ad.grad3_t(ctx, t, dims), (V,))


def _shrink_torch(V, t, iso):
    """Plain PyTorch FP64 restatement of the shrinkage map (the reference for
    the VJP kernel)."""
    import torch
    if iso:
        r = torch.sqrt((V * V).sum(0))
        sh = torch.clamp_min(r - t, 0)
        W = V * (sh / torch.where(r > 0, r, torch.ones_like(r)))
        return W, sh
    sh = torch.clamp_min(V.abs() - t, 0)
    return torch.sign(V) * sh, sh.sum(0)


@pytest.mark.parametrize("iso", [True, False])
def test_shrink_forward_vjp_and_gradcheck(ctx, iso):
    import torch
    gen = torch.Generator(device="cuda").manual_seed(13 + iso)
    n, beta = 5000, 8.0
    t = 1.0 / beta
    for dt in (torch.float32, torch.float64):
        DU = torch.randn(3, n, device="cuda", dtype=dt, generator=gen) * 0.2
        sig = torch.randn(3, n, device="cuda", dtype=dt, generator=gen) * 0.2
        if dt == torch.float32:     # the plain solver's shrinkage, bitwise
            *Wref, lam1 = ctx.shrink(DU[0], DU[1], DU[2], sig[0], sig[1], sig[2], beta, iso=iso)
            W, s, _ = ad.shrink(ctx, DU - sig / beta, ad._thresh(DU, beta), iso)
            for k in range(3):
                assert torch.equal(W[k], Wref[k])
            assert abs(float(s.double().sum()) - lam1) <= 1e-6 * abs(lam1)
    # VJP against torch autograd of the restatement (FP64), away from the kink
    V = torch.randn(3, n, device="cuda", dtype=torch.float64, generator=gen) * 0.2
    r = torch.sqrt((V * V).sum(0)) if iso else V.abs()
    keep = ((r - t).abs() > 1e-3).all(0) if not iso else (r - t).abs() > 1e-3
    V = V[:, keep].contiguous()
    gW = torch.randn_like(V)
    gs = torch.randn(V.shape[1], device="cuda", dtype=torch.float64, generator=gen)
    Vr = V.clone().requires_grad_(True)
    Wr, sr = _shrink_torch(Vr, t, iso)
    (Wr * gW).sum().add((sr * gs).sum()).backward()
    Vk = V.clone().requires_grad_(True)
    Wk, sk, _ = ad.shrink(ctx, Vk, t, iso)
    assert torch.allclose(Wk, Wr.detach(), rtol=1e-13, atol=1e-15)
    (Wk * gW).sum().add((sk * gs).sum()).backward()
    assert torch.allclose(Vk.grad, Vr.grad, rtol=1e-12, atol=1e-14)
    # gradcheck on a small sample
    Vs = V[:, :40].clone().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x: ad.shrink(ctx, x, t, iso)[:2], (Vs,), eps=1e-7)
    # a recorded active set selects the smooth branch: scaling V by 1.5 moves
    # voxels across the threshold, the replay keeps the recorded ones
    W0, s0, act = ad.shrink(ctx, V, t, iso)
    V2 = V * 1.5
    W2, s2, act2 = ad.shrink(ctx, V2, t, iso, mask_in=act)
    assert torch.equal(act2, act)
    if iso:
        r2 = torch.sqrt((V2 * V2).sum(0))
        on = act.bool()
        Wb = torch.where(on, V2 * ((r2 - t) / r2), torch.zeros_like(V2))
        sb = torch.where(on, r2 - t, torch.zeros_like(r2))
    else:
        on = torch.stack([(act >> k) & 1 for k in range(3)]).bool()
        tb = torch.where(on, V2.abs() - t, torch.zeros_like(V2))
        Wb, sb = torch.sign(V2) * tb, tb.sum(0)
    assert not torch.equal(ad.shrink(ctx, V2, t, iso)[2], act)
    assert torch.allclose(W2, Wb, rtol=1e-13, atol=1e-15)
    assert torch.allclose(s2, sb, rtol=1e-13, atol=1e-15)


def _opts(**kw):
    base = dict(mu=2.0 ** 9, mu0=2.0 ** 9, tv_mode="iso")
    base.update(kw)
    return api.default_opts(**base)


def _oracle_opts(o):
    return orc.default_opts(mu=o.mu, mu0=o.mu0, beta=o.beta, beta0=o.beta0, tol=o.tol,
                            tol_inn=o.tol_inn, maxit=o.maxit, maxcnt=o.maxcnt, c=o.c,
                            gamma=o.gamma, gam=o.gam, rate_gam=o.rate_gam,
                            rate_cnt=o.rate_cnt, nonneg=o.nonneg, tv_iso=int(o.tv_mode == 1))


@pytest.mark.parametrize("case", ["inner6", "outer3x2", "backtrack"])
def test_unrolled_forward_matches_oracle_fp64(ctx, prob, case):
    """outer x inner fixed counts; the oracle runs the same counts: one outer
    iteration with tol_inn = 0 runs maxit inner iterations; tol_inn = 1e30
    stops every outer iteration after its second inner iteration."""
    import torch
    if case == "inner6":
        outer, inner = 1, 6
        o = _opts(maxcnt=1, maxit=6, tol_inn=0.0, tol=0.0)
    elif case == "outer3x2":
        outer, inner = 3, 2
        o = _opts(maxcnt=3, maxit=100, tol_inn=1e30, tol=0.0, mu0=2.0 ** 7, beta0=8.0)
    else:
        outer, inner = 3, 2        # Armijo constant c = 1: backtracking and fallbacks
        o = _opts(maxcnt=3, maxit=100, tol_inn=1e30, tol=0.0, c=1.0)
    b = torch.tensor(prob["b"], dtype=torch.float64, device="cuda")
    with torch.no_grad():
        U, tape = ad.tval3_unrolled(ctx, prob["A"], b, prob["dims"], o, outer=outer,
                                    inner=inner, checkpoint=False)
    Uo, io, ho, _ = orc.solve(prob["oop"], prob["dims"], prob["b"], _oracle_opts(o),
                              dtype=np.float64)
    assert io.total_iters == outer * inner and io.outer_iters == outer
    cnt = [c for ks in tape.cnt for c in ks]
    assert cnt == [int(c) for c in ho["cnt"]], (cnt, ho["cnt"])
    if case == "backtrack":
        assert max(cnt) == 6, "the case is meant to take the fallback branch"
    Ug = U.cpu().numpy()
    err = np.linalg.norm(Ug - Uo) / np.linalg.norm(Uo)
    print(f"{case}: rel(U_unrolled, U_oracle_fp64) = {err:.2e}, cnt = {cnt}")
    assert err < 1e-10, err


def _loss_fn(ctx, prob, o, outer, inner, checkpoint=True):
    import torch
    ut = torch.tensor(prob["u"], dtype=torch.float64, device="cuda")

    def loss(b, log_mu, tape=None):
        U, tape = ad.tval3_unrolled(ctx, prob["A"], b, prob["dims"], o, mu=torch.exp(log_mu),
                                    outer=outer, inner=inner, checkpoint=checkpoint, tape=tape)
        return ((U - ut) ** 2).sum(), tape
    return loss


@pytest.mark.parametrize("c", [1e-5, 1.0])
def test_unrolled_gradient_finite_differences_fp64(ctx, prob, c):
    """c = 1e-5 (the default Armijo constant) and c = 1 (a branch with
    backtracking steps and fallbacks).  Probes in (b, log mu) with the b part
    scaled to |b|: h = 1e-4 on the recorded branch (decisions and active sets
    replayed: the function autograd differentiates), and h = 1e-6 on the
    map itself (nothing replayed).  On the branch, central differences carry
    an O(h^2) truncation error (it falls 4x per halving of h); Richardson's
    extrapolation (4 fd(h/2) - fd(h))/3 removes it to O(h^4)."""
    import torch
    o = _opts(c=c)
    outer, inner, probes, h, h_free = 4, 3, 8, 1e-4, 1e-6
    loss = _loss_fn(ctx, prob, o, outer, inner)
    b0 = torch.tensor(prob["b"], dtype=torch.float64, device="cuda")
    lm0 = torch.tensor(math.log(o.mu), dtype=torch.float64, device="cuda")
    b = b0.clone().requires_grad_(True)
    lm = lm0.clone().requires_grad_(True)
    L0, tape = loss(b, lm)
    print(f"c={c}: decisions {tape.cnt}")
    L0.backward()
    gb, glm = b.grad.clone(), float(lm.grad)
    gen = torch.Generator(device="cuda").manual_seed(21)
    worst = worst_r = worst_free = 0.0

    def fd_at(step, vb, vm, tp):
        Lp, _ = loss(b0 + step * vb, lm0 + step * vm, tp)
        Lm, _ = loss(b0 - step * vb, lm0 - step * vm, tp)
        return float(Lp - Lm) / (2 * step)

    with torch.no_grad():
        for i in range(probes):
            vb = torch.randn(b0.numel(), device="cuda", dtype=torch.float64, generator=gen)
            vb *= float(b0.norm() / vb.norm())
            vm = float(torch.randn((), device="cuda", dtype=torch.float64, generator=gen))
            an = float(torch.dot(gb, vb)) + glm * vm
            fd, fd2 = fd_at(h, vb, vm, tape), fd_at(h / 2, vb, vm, tape)
            fdr = (4 * fd2 - fd) / 3
            fd_free = fd_at(h_free, vb, vm, None)
            worst, worst_r = max(worst, _rel(fd, an)), max(worst_r, _rel(fdr, an))
            worst_free = max(worst_free, _rel(fd_free, an))
            print(f"probe {i}: analytic={an:.10e} fd(branch, h)={fd:.10e} rel={_rel(fd, an):.2e}"
                  f" richardson rel={_rel(fdr, an):.2e} fd(free, h=1e-6) rel="
                  f"{_rel(fd_free, an):.2e}")
    print(f"dL/dlog(mu) = {glm:.6e}; worst rel err: branch {worst:.2e}, richardson "
          f"{worst_r:.2e}, free {worst_free:.2e}")
    assert worst < 1e-4, worst
    assert worst_r < 1e-6, worst_r
    assert worst_free < 1e-6, worst_free


def test_checkpointed_backward_equals_plain(ctx, prob):
    import torch
    o = _opts()
    b0 = torch.tensor(prob["b"], dtype=torch.float64, device="cuda")
    lm0 = torch.tensor(math.log(o.mu), dtype=torch.float64, device="cuda")
    grads = []
    tape = None
    for ck in (True, False):
        b = b0.clone().requires_grad_(True)
        lm = lm0.clone().requires_grad_(True)
        L, tape = _loss_fn(ctx, prob, o, 3, 3, checkpoint=ck)(b, lm, tape)
        L.backward()
        grads.append((b.grad.clone(), lm.grad.clone(), float(L)))
    assert grads[0][2] == grads[1][2]
    assert torch.equal(grads[0][0], grads[1][0]) and torch.equal(grads[0][1], grads[1][1])


def test_unrolled_fp32_close_to_fp64(ctx, prob):
    import torch
    o = _opts()
    b64 = torch.tensor(prob["b"], dtype=torch.float64, device="cuda")
    with torch.no_grad():
        U64, tape = ad.tval3_unrolled(ctx, prob["A"], b64, prob["dims"], o, outer=3, inner=3)
        U32, _ = ad.tval3_unrolled(ctx, prob["A"], b64.float(), prob["dims"], o, outer=3,
                                   inner=3, tape=tape)
    err = float((U32.double() - U64).norm() / U64.norm())
    print(f"rel(U_fp32, U_fp64) = {err:.2e}")
    assert err < 1e-4, err


def test_unrolled_rejects_bad_arguments(ctx, prob):
    import torch
    b = torch.tensor(prob["b"], dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        ad.tval3_unrolled(ctx, prob["A"], b[:-1], prob["dims"], _opts())
    with pytest.raises(ValueError):
        ad.tval3_unrolled(ctx, prob["A"], b, (12, 12, 11), _opts())
    with pytest.raises(ValueError):
        ad.tval3_unrolled(ctx, prob["A"], b, prob["dims"], _opts(skipMulRatio=0.5))
    with pytest.raises(TypeError):
        ad.shrink(ctx, torch.zeros(3, 4, dtype=torch.float16, device="cuda"), 0.1, True)


def test_unrolled_frozen_steps_gradient(ctx, prob):
    """step_grad='frozen': same forward values; the gradient of the loop with
    the recorded step sizes, checked by central differences on the recorded
    branch (step sizes replayed), Richardson-extrapolated."""
    import torch
    o = _opts(c=1.0)
    b0 = torch.tensor(prob["b"], dtype=torch.float64, device="cuda")
    lm0 = torch.tensor(math.log(o.mu), dtype=torch.float64, device="cuda")
    ut = torch.tensor(prob["u"], dtype=torch.float64, device="cuda")

    def loss(b, lm, mode, tape=None):
        U, tape = ad.tval3_unrolled(ctx, prob["A"], b, prob["dims"], o, mu=torch.exp(lm),
                                    outer=4, inner=3, tape=tape, step_grad=mode)
        return ((U - ut) ** 2).sum(), tape

    with torch.no_grad():
        assert float(loss(b0, lm0, "full")[0]) == float(loss(b0, lm0, "frozen")[0])
    b = b0.clone().requires_grad_(True)
    lm = lm0.clone().requires_grad_(True)
    L, tape = loss(b, lm, "frozen")
    L.backward()
    assert sum(len(v) for v in tape.taus.values()) >= 12
    gen = torch.Generator(device="cuda").manual_seed(31)
    worst = 0.0
    with torch.no_grad():
        for _ in range(4):
            vb = torch.randn(b0.numel(), device="cuda", dtype=torch.float64, generator=gen)
            vb *= float(b0.norm() / vb.norm())
            vm = float(torch.randn((), device="cuda", dtype=torch.float64, generator=gen))
            an = float(torch.dot(b.grad, vb)) + float(lm.grad) * vm
            f = []
            for h in (1e-4, 5e-5):
                f.append(float(loss(b0 + h * vb, lm0 + h * vm, "frozen", tape)[0] -
                               loss(b0 - h * vb, lm0 - h * vm, "frozen", tape)[0]) / (2 * h))
            worst = max(worst, _rel((4 * f[1] - f[0]) / 3, an))
    print(f"frozen steps: worst Richardson rel err {worst:.2e}")
    assert worst < 1e-6, worst
