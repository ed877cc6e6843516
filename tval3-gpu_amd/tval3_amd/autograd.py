"""Differentiable TVAL3: A, D, D^T and shrinkage as torch.autograd.Functions
whose forward and backward both run hand-written HIP kernels of the library,
and an unrolled TVAL3 built from them.

Building blocks (float32 or float64 tensors on the device, torch's current
stream):

    matvec(op, x, trans)    y = A x / A^T x        backward: the other product
    grad3(ctx, U, dims)     D U  -> (3, n)          backward: D^T
    grad3_t(ctx, V, dims)   D^T V -> (n)            backward: D
    shrink(ctx, V, t, iso)  (W (3, n), s (n))       backward: tv3_shrink_adjoint

Unrolled solver: ``tval3_unrolled`` restates the reference's TVAL3 loop
(3d/tval3_gpu_3d.cpp:1088-1345 -- scaleb :225-266, get_g :366-572, get_tau
:577-666, descend :671-701, update_g :706-818, min_u :823-891, get_gradient
:896-919, shrinkage :924-960, update_W :965-1029, update_mlp :1034-1083)
with a fixed number of outer iterations and of inner iterations per outer
iteration, differentiable with respect to b and mu.  Its discrete decisions
-- the entries scaleb picks, and per inner iteration how many backtracking
steps the nonmonotone Armijo test took (6: the fallback after five) -- are
kept on a DecisionTape: the first run records them, a run with the same tape
replays them.  The gradient is that of the branch the recorded decisions
select; checkpoint recomputation and finite-difference probes replay the tape
so they evaluate the same branch.  beta, the continuation rates and the
line-search constants are plain numbers, not differentiated.

Each outer iteration can be wrapped in torch.utils.checkpoint, so memory for
the backward pass is one outer iteration's intermediates plus the state
vectors at each outer boundary.
"""
import ctypes as C

import torch

from . import _lib as L
from .api import _torch_stream, _vp


def _dtype_code(t):
    if t.dtype == torch.float32:
        return L.F32
    if t.dtype == torch.float64:
        return L.F64
    raise TypeError(f"float32 or float64 tensors only, got {t.dtype}")


def _c(t):
    return t if t.is_contiguous() else t.contiguous()


# ------------------------------------------------------------------ products
class MatVec(torch.autograd.Function):
    """y = A x (trans False) or A^T x through the operator's kernels
    (mat_vec_mul, 3d/mat_vec_mul_sparse.h:1703-1780 and its dense/dynamic
    overloads); the backward applies the other product."""

    @staticmethod
    def forward(ctx, x, op, trans):
        ctx.op, ctx.trans = op, trans
        return op.mat_vec_mul(_c(x), trans=trans)

    @staticmethod
    def backward(ctx, g):
        return ctx.op.mat_vec_mul(_c(g), trans=not ctx.trans), None, None


def matvec(op, x, trans=False):
    return MatVec.apply(x, op, trans)


def _D(tctx, U, dims):
    out = torch.empty((3, U.numel()), dtype=U.dtype, device=U.device)
    d = (C.c_int32 * 3)(*dims)
    f = L.load().tv3_D_f64 if U.dtype == torch.float64 else L.load().tv3_D
    _dtype_code(U)
    tctx.check(f(tctx.h, d, _vp(_c(U)), _vp(out[0]), _vp(out[1]), _vp(out[2]), _torch_stream()))
    return out


def _Dt(tctx, V, dims):
    V = _c(V)
    out = torch.empty(V.shape[1], dtype=V.dtype, device=V.device)
    d = (C.c_int32 * 3)(*dims)
    f = L.load().tv3_Dt_f64 if V.dtype == torch.float64 else L.load().tv3_Dt
    _dtype_code(V)
    tctx.check(f(tctx.h, d, _vp(V[0]), _vp(V[1]), _vp(V[2]), _vp(out), _torch_stream()))
    return out


class Grad3(torch.autograd.Function):
    """Periodic forward differences D (3d/tval3_gpu_3d.cpp:271-314): rows
    (Ux, Uy, Uz) = differences along dim_x, dim_y, dim_z; backward D^T."""

    @staticmethod
    def forward(ctx, U, tctx, dims):
        ctx.tctx, ctx.dims = tctx, tuple(dims)
        return _D(tctx, U, dims)

    @staticmethod
    def backward(ctx, g):
        return _Dt(ctx.tctx, g, ctx.dims), None, None


class Grad3T(torch.autograd.Function):
    """D^T (3d/tval3_gpu_3d.cpp:319-361) of a (3, n) tensor; backward D."""

    @staticmethod
    def forward(ctx, V, tctx, dims):
        ctx.tctx, ctx.dims = tctx, tuple(dims)
        return _Dt(tctx, V, dims)

    @staticmethod
    def backward(ctx, g):
        return _D(ctx.tctx, g, ctx.dims), None, None


def grad3(tctx, U, dims):
    return Grad3.apply(U, tctx, dims)


def grad3_t(tctx, V, dims):
    return Grad3T.apply(V, tctx, dims)


class Shrink(torch.autograd.Function):
    """W, s, active = shrink(V, t) (shrinkage_gpu, 3d/tval3_gpu_3d.cpp:924-960,
    as a map of V = DU - sigma/beta at threshold t = 1/beta); s per voxel,
    lam1 = s.sum(); active: the per-voxel active set (uint8).  With
    mask_in, the recorded active set selects the branch instead of the
    comparison with t.  t is a constant (no gradient)."""

    @staticmethod
    def forward(ctx, V, t, iso, tctx, mask_in=None):
        V = _c(V)
        n = V.shape[1]
        W = torch.empty_like(V)
        s = torch.empty(n, dtype=V.dtype, device=V.device)
        active = torch.empty(n, dtype=torch.uint8, device=V.device)
        if mask_in is not None:
            assert mask_in.dtype == torch.uint8 and mask_in.numel() == n
            mask_in = _c(mask_in)
        tctx.check(L.load().tv3_shrink_apply(tctx.h, n, L.ISO if iso else L.ANISO,
                                             _dtype_code(V), _vp(V[0]), _vp(V[1]), _vp(V[2]),
                                             float(t), _vp(W[0]), _vp(W[1]), _vp(W[2]), _vp(s),
                                             _vp(mask_in), _vp(active), _torch_stream()))
        ctx.save_for_backward(V)
        ctx.t, ctx.iso, ctx.tctx = float(t), iso, tctx
        ctx.mark_non_differentiable(active)
        return W, s, active

    @staticmethod
    def backward(ctx, gW, gs, _ga):
        (V,) = ctx.saved_tensors
        n = V.shape[1]
        if gW is None:
            gW = torch.zeros_like(V)
        gW = _c(gW)
        gs = None if gs is None else _c(gs)
        gV = torch.empty_like(V)
        ctx.tctx.check(L.load().tv3_shrink_adjoint(
            ctx.tctx.h, n, L.ISO if ctx.iso else L.ANISO, _dtype_code(V), _vp(V[0]), _vp(V[1]),
            _vp(V[2]), ctx.t, _vp(gW[0]), _vp(gW[1]), _vp(gW[2]), _vp(gs), _vp(gV[0]),
            _vp(gV[1]), _vp(gV[2]), _torch_stream()))
        return gV, None, None, None, None


def shrink(tctx, V, t, iso=True, mask_in=None):
    """(W, s, active) -- see Shrink."""
    return Shrink.apply(V, t, iso, tctx, mask_in)


# ------------------------------------------------------------------ unrolled TVAL3
class DecisionTape:
    """The discrete decisions of one unrolled solve.  ``scaleb``: (i_min,
    i_max, branch) of b's scaling (branch 0: none, 1: to 0.5, 2: to 1.5);
    ``cnt[k][j]``: backtracking steps of inner iteration j of outer iteration
    k (0..5; 6 = the fallback after five); ``masks[k]``: in order of use, the
    active sets of the nonnegativity clamp (bool) and of each shrinkage
    (uint8) -- the kinks of the map (k = -1: before the loop); ``taus[k]``:
    the step sizes, kept only when they are held constant.  A tape is
    recorded by the first run that uses it and replayed by every later one,
    which then evaluates the smooth branch the record selected."""

    def __init__(self):
        self.scaleb = None
        self.cnt = []
        self.masks = {}
        self.taus = {}

    def complete(self, outer):
        return self.scaleb is not None and len(self.cnt) >= outer

    def nbytes(self):
        return sum(m.numel() * m.element_size() for ms in self.masks.values() for m in ms)


class _StepLog:
    """Record/replay of the step sizes tau of one outer iteration (used when
    the step sizes are held constant, step_grad='frozen')."""

    def __init__(self, store):
        self.store, self.i, self.replay = store, 0, len(store) > 0

    def take(self, tau):
        if self.replay:
            t = self.store[self.i]
            self.i += 1
            return t
        t = tau.detach()
        self.store.append(t)
        return t


class _MaskLog:
    """Sequential record/replay of the active sets of one outer iteration."""

    def __init__(self, store):
        self.store, self.i, self.replay = store, 0, len(store) > 0

    def next(self):
        if not self.replay:
            return None
        m = self.store[self.i]
        self.i += 1
        return m

    def put(self, m):
        if not self.replay:
            self.store.append(m)


def _dot(a, b):
    """Inner product accumulated in FP64 (the solvers' reductions are FP64)."""
    return torch.dot(a.reshape(-1).double(), b.reshape(-1).double())


def _sq(a):
    return _dot(a, a)


_STATE = ("U", "DU", "W", "sig", "dlt", "Au", "g", "g2", "d", "DtsAtd", "f", "lam1", "lam2",
          "lam3", "lam4", "lam5", "mu")


class _Loop:
    """The solver's operators and constants; methods restate the reference's
    routines on a state dict S of tensors (vectors in the working precision,
    scalars as 0-d FP64 tensors)."""

    def __init__(self, tctx, op, dims, opts, inner, bs, Atb, mu_max, freeze_steps=False):
        self.tctx, self.op, self.dims, self.o = tctx, op, tuple(dims), opts
        self.freeze_steps = freeze_steps
        self.iso = opts.tv_mode == L.ISO
        self.nonneg = bool(opts.nonneg)
        self.inner = inner
        self.bs, self.Atb, self.mu_max = bs, Atb, mu_max

    # operators
    def D(self, U):
        return grad3(self.tctx, U, self.dims)

    def Dt(self, V):
        return grad3_t(self.tctx, V, self.dims)

    def A(self, x, trans=False):
        return matvec(self.op, x, trans)

    @staticmethod
    def fval(S, beta):
        return S["lam1"] + beta / 2 * S["lam2"] + S["mu"] / 2 * S["lam3"] - S["lam4"] - S["lam5"]

    def lam24(self, S):
        V = S["DU"] - S["W"]
        S["lam2"], S["lam4"] = _sq(V), _dot(V, S["sig"])
        return V

    def lam35(self, S):
        r = S["Au"] - self.bs
        S["lam3"], S["lam5"] = _sq(r), _dot(S["dlt"], r)

    def get_g(self, S, beta):
        """get_g (3d/tval3_gpu_3d.cpp:366-572), every product taken."""
        S["Au"] = self.A(S["U"])
        g = self.A(S["Au"], True)
        V = self.lam24(S)
        self.lam35(S)
        S["g2"] = self.Dt(V)
        S["g"] = g - self.Atb
        S["f"] = self.fval(S, beta)

    def descend(self, S, Ubase, tau, beta, masks):
        """descend (:671-701): U = Ubase - tau d, clamp at 0, DU, get_g."""
        U = Ubase - tau.to(Ubase.dtype) * S["d"]
        if self.nonneg:
            keep = masks.next()
            if keep is None:
                keep = U.detach() > 0
                masks.put(keep)
            U = torch.where(keep, U, torch.zeros((), dtype=U.dtype, device=U.device))
        S["U"] = U
        S["DU"] = self.D(S["U"])
        self.get_g(S, beta)

    def tau_first(self, S, mdb):
        """get_tau (:577-666) on the first inner iteration: exact step along d
        for the quadratic part."""
        d = S["d"]
        return (_sq(d) / (_sq(self.D(d)) + mdb * _sq(self.A(d)))).abs()

    def shrinkage(self, S, beta, masks):
        """shrinkage (:924-960): W = shrink(DU - sigma/beta, 1/beta), lam1."""
        W, s, active = shrink(self.tctx, S["DU"] - S["sig"] / beta, _thresh(S["U"], beta),
                              self.iso, masks.next())
        masks.put(active)
        S["W"], S["lam1"] = W, s.double().sum()

    def get_gradient(self, S, mdb):
        """get_gradient (:896-919): d = muDbeta g + g2 + DtsAtd."""
        S["d"] = (mdb.to(S["g"].dtype) * S["g"] + S["g2"]) + S["DtsAtd"]

    def outer_step(self, S, beta, cnts, record, masks, steps_log):
        """One outer iteration: `inner` inner iterations (min_u :823-891,
        update_W :965-1029, :1190-1202) then update_mlp (:1034-1083) and the
        continuation (:1235-1249).  Returns the new state and beta."""
        o = self.o
        S = dict(S)
        mdb = S["mu"] / beta
        gam, Q, C = o.gam, 1.0, S["f"]
        uup = gp = g2p = None
        for j in range(self.inner):
            # ---- get_tau
            if j == 0:
                tau = self.tau_first(S, mdb)
            else:
                dg, dg2 = S["g"] - gp, S["g2"] - g2p
                tau = (_sq(uup) / _dot(uup, dg2 + mdb.to(dg.dtype) * dg)).abs()
            if self.freeze_steps:
                tau = steps_log.take(tau)
            Up, gp, g2p, Aup, DUp = S["U"], S["g"], S["g2"], S["Au"], S["DU"]
            self.descend(S, Up, tau, beta, masks)
            alpha = 1.0
            dU = S["U"] - Up
            c_armij = _sq(S["d"]) * tau * o.c * beta

            def fails():
                # a decision, not part of the differentiated function
                with torch.no_grad():
                    return bool(S["f"].abs() > (C - alpha * c_armij).abs())

            if record:
                cnt = 0
                if fails():
                    cnt = -1        # decided below
            else:
                cnt = cnts[j]
            if cnt != 0:
                dg, dg2 = S["g"] - gp, S["g2"] - g2p
                dAu, dDU = S["Au"] - Aup, S["DU"] - DUp
                steps = 0
                while True:
                    if record:
                        if not fails():
                            cnt = steps
                            break
                        if steps == 5:
                            cnt = 6
                    elif steps == cnt:
                        break
                    if cnt == 6 and steps == 5:
                        # fallback (:876-882): fresh exact step from Up
                        gam *= o.rate_gam
                        tau = self.tau_first(S, mdb)
                        if self.freeze_steps:
                            tau = steps_log.take(tau)
                        self.descend(S, Up, tau, beta, masks)
                        break
                    alpha *= o.gamma
                    # update_g (:706-818)
                    S["g"] = alpha * dg + gp
                    S["g2"] = alpha * dg2 + g2p
                    S["U"] = alpha * dU + Up
                    S["Au"] = alpha * dAu + Aup
                    S["DU"] = alpha * dDU + DUp
                    self.lam24(S)
                    self.lam35(S)
                    S["f"] = self.fval(S, beta)
                    steps += 1
            if record:
                cnts.append(cnt)
            # ---- update_W
            f = S["f"] - (S["lam1"] + beta / 2 * S["lam2"] - S["lam4"])
            self.shrinkage(S, beta, masks)
            V = self.lam24(S)
            S["g2"] = self.Dt(V)
            S["f"] = f + (S["lam1"] + beta / 2 * S["lam2"] - S["lam4"])
            Qp = Q
            Q = gam * Qp + 1
            C = (gam * Qp * C + S["f"]) / Q
            uup = S["U"] - Up
            self.get_gradient(S, mdb)
        # ---- update_mlp
        f = S["f"] + (S["lam4"] + S["lam5"])
        V = S["DU"] - S["W"]
        S["sig"] = S["sig"] - beta * V
        r = S["Au"] - self.bs
        S["dlt"] = S["dlt"] - S["mu"].to(r.dtype) * r
        S["lam4"], S["lam5"] = _dot(S["sig"], V), _dot(S["dlt"], r)
        S["f"] = f - (S["lam4"] + S["lam5"])
        # ---- continuation
        beta0 = beta
        beta = min(beta * o.rate_cnt, o.beta)
        S["mu"] = torch.minimum(S["mu"] * o.rate_cnt, self.mu_max)
        mdb = S["mu"] / beta
        S["f"] = self.fval(S, beta)
        S["DtsAtd"] = _ratio(beta0, beta, S["d"]) * S["d"]
        self.get_gradient(S, mdb)
        return S, beta


def _thresh(U, beta):
    """1/beta as the solvers round it: (REAL)1 / (REAL)beta."""
    if U.dtype == torch.float32:
        one = torch.tensor(1.0, dtype=torch.float32)
        return float(one / torch.tensor(beta, dtype=torch.float32))
    return 1.0 / beta


def _ratio(beta0, beta, like):
    """(REAL)(beta0/beta), the DtsAtd factor of :1247."""
    r = beta0 / beta
    return float(torch.tensor(r, dtype=like.dtype)) if like.dtype == torch.float32 else r


def _scaleb(b, tape):
    """scaleb (3d/tval3_gpu_3d.cpp:225-266): the entries of smallest and
    largest magnitude (first occurrences) set the range, which is scaled to
    0.5 if below 0.5 or 1.5 if above 1.5.  Returns (scaled b, scale)."""
    if tape.scaleb is None:
        a = b.detach().abs()
        i_min, i_max = int(torch.argmin(a)), int(torch.argmax(a))
        bd = b.detach()
        dif = float((bd[i_max] - bd[i_min]).abs())
        branch = 1 if dif < 0.5 else (2 if dif > 1.5 else 0)
        tape.scaleb = (i_min, i_max, branch)
    i_min, i_max, branch = tape.scaleb
    if branch == 0:
        return b, None
    s = (0.5 if branch == 1 else 1.5) / (b[i_max] - b[i_min]).abs()
    return s * b, s


def tval3_unrolled(tctx, op, b, dims, opts, mu=None, outer=20, inner=4, checkpoint=True,
                   tape=None, step_grad="full"):
    """Unrolled TVAL3: `outer` outer iterations of `inner` inner iterations
    each, from U0 = A^T b (3d/tval3_gpu_3d.cpp:1122-1160), differentiable with
    respect to b (a float32 or float64 tensor, which sets the working
    precision) and mu (a 0-d tensor; default opts.mu; mu0 follows as
    mu * opts.mu0/opts.mu, the continuation as in the solver).

    opts: a tv3_opts (tval3_amd.api.default_opts); tol, tol_inn, maxit and
    maxcnt are not used (the counts are fixed).  The skip heuristic is not
    part of the unrolled loop (skipMulRatio must be 0).

    step_grad: 'full' differentiates through the step sizes too (the
    Barzilai-Borwein step tau = <s,s>/<s,y> and the exact first step), the
    derivative of the map itself.  The BB recursion makes long unrolls very
    sensitive (the derivative grows geometrically with the number of outer
    iterations).  'frozen' holds every tau at its recorded value (kept on
    the tape): the derivative of the loop with those step sizes, a
    well-conditioned gradient for learning b- or mu-dependent quantities.

    Returns (U, tape): U in b's precision, unscaled like the solver's
    result."""
    if step_grad not in ("full", "frozen"):
        raise ValueError("step_grad must be 'full' or 'frozen'")
    if opts.skipMulRatio != 0:
        raise ValueError("the unrolled loop takes every product (skipMulRatio must be 0)")
    if outer < 1 or inner < 1:
        raise ValueError("outer and inner must be >= 1")
    if b.dim() != 1 or b.numel() != op.m:
        raise ValueError(f"b must be a vector of {op.m} entries")
    n = 1
    for v in dims:
        n *= int(v)
    if n != op.n:
        raise ValueError("prod(dims) must equal the columns of A")
    dev = b.device
    if mu is None:
        mu = torch.tensor(float(opts.mu), dtype=torch.float64, device=dev)
    mu = mu.to(torch.float64)
    tape = DecisionTape() if tape is None else tape
    replay = tape.complete(outer)

    bs, scl = _scaleb(b, tape)
    Atb = matvec(op, bs, True)
    mu0 = mu * (opts.mu0 / opts.mu) if opts.mu0 != opts.mu else mu
    loop = _Loop(tctx, op, dims, opts, inner, bs, Atb, mu, step_grad == "frozen")
    beta = min(opts.beta0, opts.beta)
    zn = torch.zeros(n, dtype=b.dtype, device=dev)
    z0 = torch.zeros((), dtype=torch.float64, device=dev)
    S = dict(U=Atb, sig=torch.zeros((3, n), dtype=b.dtype, device=dev),
             dlt=torch.zeros_like(b), DtsAtd=zn, mu=torch.minimum(mu0, mu),
             lam1=z0, lam2=z0, lam3=z0, lam4=z0, lam5=z0)
    S["DU"] = loop.D(S["U"])
    loop.shrinkage(S, beta, _MaskLog(tape.masks.setdefault(-1, [])))
    loop.get_g(S, beta)
    loop.get_gradient(S, S["mu"] / beta)

    for k in range(outer):
        rec = not replay and len(tape.cnt) <= k
        if rec:
            tape.cnt.append([])
        cnts = tape.cnt[k]
        mstore = tape.masks.setdefault(k, [])
        tstore = tape.taus.setdefault(k, [])

        def step(*flat, _beta=beta, _cnts=cnts, _ms=mstore, _ts=tstore):
            # a recorded iteration is replayed when checkpointing recomputes it
            record = len(_cnts) == 0
            St, _ = loop.outer_step(dict(zip(_STATE, flat)), _beta, _cnts, record,
                                    _MaskLog(_ms), _StepLog(_ts))
            return tuple(St[key] for key in _STATE)

        flat = tuple(S[key] for key in _STATE)
        if checkpoint and torch.is_grad_enabled():
            out = torch.utils.checkpoint.checkpoint(step, *flat, use_reentrant=False)
        else:
            out = step(*flat)
        S = dict(zip(_STATE, out))
        beta = min(beta * opts.rate_cnt, opts.beta)

    U = S["U"]
    if scl is not None:
        U = U / scl.to(U.dtype)
    return U, tape
