"""Gradient check of the differentiable unrolled TVAL3 at full size.

    python -m tval3_amd.gradcheck [--N 96] [--E 192] [--R 768] [--outer 20]
                                  [--inner 4] [--probes 64] [--h 1e-4]

Workload: an N^3 slowness phantom, E x R rays on the sphere of radius 1.1,
Siddon weights as an FP32 CSR built on the device, b = A u_true + 0.5% noise,
isotropic TV, mu = 2^9.  Loss: |u_K - u_true|^2 after `outer` outer
iterations of `inner` inner iterations, each outer iteration checkpointed.
Gradients with respect to b and log(mu).  Checks, all on the GPU through the
library's kernels:

  * adjoint (dot-product) tests <Ax, y> = <x, A^T y> for A (FP32 and FP64
    products) and <D u, V> = <u, D^T V>;
  * for each gradient mode (step_grad 'frozen': step sizes held at their
    recorded values; 'full': differentiated through the step-size rule) the
    FP64 gradient against FP64 central finite differences on `probes` seeded
    directions in (b, log mu) (the b part scaled to |b|): at step h on the
    branch the first run recorded (line-search decisions, active sets of the
    clamp and the shrinkage, and in 'frozen' mode the step sizes replayed),
    with Richardson's extrapolation from h and h/2; in 'full' mode also at
    steps 1e-6 and 1e-8 on the map itself, in 'frozen' mode also at step
    1e-6 on the branch;
  * the growth of |dL/dlog(mu)| with the number of outer iterations
    ('full' mode);
  * the FP32 gradient against the FP64 one.

Prints one JSON object (timings in seconds, errors relative).
"""
import argparse
import json
import math
import time

import numpy as np


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def run(N=96, E=192, R=768, outer=20, inner=4, probes=64, h=1e-4, seed=1, log=print):
    import torch

    from . import api, autograd as ad, workload as wl

    ctx = api.Context(0)
    dims = (N, N, N)
    t0 = time.time()
    u = wl.phantom_3d(N, 6, seed=seed)
    geom = wl.transducer_geometry(E, R, N=N, seed=seed + 1)
    A = api.DynamicOperator(ctx, geom, "siddon").materialize("f32")
    ut = torch.tensor(u, dtype=torch.float64, device="cuda")
    Au = A.mat_vec_mul(ut)
    b_np = wl.add_noise(Au.cpu().numpy(), 0.005, seed=seed + 2)
    b64 = torch.tensor(b_np, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    out = dict(workload=dict(N=N, rays=A.m, nnz=A.nnz, weights="f32 CSR (Siddon)",
                             outer=outer, inner=inner, probes=probes, h=h),
               setup_s=time.time() - t0)
    log(f"problem: {N}^3 voxels, {A.m} rays, nnz {A.nnz} ({out['setup_s']:.1f}s)")

    # ---- adjoint tests
    gen = torch.Generator(device="cuda").manual_seed(seed + 100)
    adj = {}
    for name, dt in (("A_f32", torch.float32), ("A_f64", torch.float64)):
        x = torch.randn(A.n, device="cuda", dtype=dt, generator=gen)
        y = torch.randn(A.m, device="cuda", dtype=dt, generator=gen)
        lhs = float(torch.dot(ad.matvec(A, x).double(), y.double()))
        rhs = float(torch.dot(x.double(), ad.matvec(A, y, True).double()))
        adj[name] = _rel(lhs, rhs)
    x = torch.randn(A.n, device="cuda", dtype=torch.float64, generator=gen)
    V = torch.randn(3, A.n, device="cuda", dtype=torch.float64, generator=gen)
    adj["D_f64"] = _rel(float((ad.grad3(ctx, x, dims) * V).sum()),
                        float(torch.dot(x, ad.grad3_t(ctx, V, dims))))
    out["adjoint_rel"] = adj
    log(f"adjoint tests: {adj}")

    o = api.default_opts(mu=2.0 ** 9, mu0=2.0 ** 9, tv_mode="iso")
    lm0 = torch.tensor(math.log(o.mu), dtype=torch.float64, device="cuda")

    def loss(b, lm, mode, tape=None, checkpoint=True, n_outer=outer):
        U, tape = ad.tval3_unrolled(ctx, A, b, dims, o, mu=torch.exp(lm), outer=n_outer,
                                    inner=inner, checkpoint=checkpoint, tape=tape,
                                    step_grad=mode)
        return ((U.double() - ut) ** 2).sum(), tape, U

    def grads(b0, mode, n_outer=outer):
        b = b0.clone().requires_grad_(True)
        lm = lm0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        t = time.time()
        L, tape, U = loss(b, lm, mode, n_outer=n_outer)
        torch.cuda.synchronize()
        tf = time.time() - t
        t = time.time()
        L.backward()
        torch.cuda.synchronize()
        tb = time.time() - t
        err = float((U.detach().double() - ut).norm() / ut.norm())
        return dict(L=float(L.detach()), gb=b.grad.detach().clone(), glm=float(lm.grad),
                    tape=tape, fwd_s=tf, bwd_s=tb,
                    peak_gib=torch.cuda.max_memory_allocated() / 2 ** 30, rel_err=err)

    keys = ("L", "glm", "fwd_s", "bwd_s", "peak_gib", "rel_err")
    out["sensitivity_full"] = {}
    for k in sorted({1, 5, 10, outer}):
        if k <= outer:
            out["sensitivity_full"][k] = abs(grads(b64, "full", k)["glm"])
    log(f"|dL/dlog(mu)| (full) by outer iterations: {out['sensitivity_full']}")

    for mode in ("frozen", "full"):
        g64 = grads(b64, mode)
        g32 = grads(b64.float(), mode)
        tape = g64["tape"]
        ncnt = [c for ks in tape.cnt for c in ks]
        rec = dict(fp64={k: g64[k] for k in keys}, fp32={k: g32[k] for k in keys},
                   backtracks=int(sum(min(c, 5) for c in ncnt)),
                   fallbacks=int(sum(c == 6 for c in ncnt)),
                   tape_mask_mib=tape.nbytes() / 2 ** 20,
                   grad_b_fp32_vs_fp64_rel=float((g32["gb"].double() - g64["gb"]).norm() /
                                                 g64["gb"].norm()),
                   grad_logmu_fp32_vs_fp64_rel=_rel(g32["glm"], g64["glm"]))
        log(f"[{mode}] FP64: L={g64['L']:.6e} dL/dlog(mu)={g64['glm']:.6e} fwd "
            f"{g64['fwd_s']:.2f}s bwd {g64['bwd_s']:.2f}s peak {g64['peak_gib']:.2f} GiB; "
            f"FP32 fwd {g32['fwd_s']:.2f}s bwd {g32['bwd_s']:.2f}s; grad_b fp32 vs fp64 "
            f"{rec['grad_b_fp32_vs_fp64_rel']:.2e}")
        gen = torch.Generator(device="cuda").manual_seed(seed + 200)
        res = []
        t = time.time()
        with torch.no_grad():
            def fd(step, vb, vm, tp):
                Lp = loss(b64 + step * vb, lm0 + step * vm, mode, tp, False)[0]
                Lm = loss(b64 - step * vb, lm0 - step * vm, mode, tp, False)[0]
                return float(Lp - Lm) / (2 * step)

            for i in range(probes):
                vb = torch.randn(b64.numel(), device="cuda", dtype=torch.float64, generator=gen)
                vb *= float(b64.norm() / vb.norm())
                vm = float(torch.randn((), device="cuda", dtype=torch.float64, generator=gen))
                an = float(torch.dot(g64["gb"], vb)) + g64["glm"] * vm
                f1, f2 = fd(h, vb, vm, tape), fd(h / 2, vb, vm, tape)
                row = [_rel(f1, an), _rel((4 * f2 - f1) / 3, an)]
                if mode == "frozen":
                    row += [_rel(fd(1e-6, vb, vm, tape), an)]
                else:
                    row += [_rel(fd(1e-6, vb, vm, None), an), _rel(fd(1e-8, vb, vm, None), an)]
                res.append(row)
                log(f"[{mode}] probe {i}: analytic {an:.10e} rel err " +
                    " ".join(f"{v:.2e}" for v in row))
        r = np.array(res)
        names = ["branch_h", "richardson"] + (["free_1e-6", "free_1e-8"] if mode == "full"
                                              else ["branch_1e-6"])
        rec["fd"] = {"seconds": time.time() - t}
        for c, nm in enumerate(names):
            rec["fd"][nm + "_max"] = float(r[:, c].max())
            rec["fd"][nm + "_median"] = float(np.median(r[:, c]))
        out[mode] = rec
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--N", type=int, default=96)
    ap.add_argument("--E", type=int, default=192)
    ap.add_argument("--R", type=int, default=768)
    ap.add_argument("--outer", type=int, default=20)
    ap.add_argument("--inner", type=int, default=4)
    ap.add_argument("--probes", type=int, default=64)
    ap.add_argument("--h", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    out = run(a.N, a.E, a.R, a.outer, a.inner, a.probes, a.h, a.seed,
              log=lambda s: print(s, flush=True))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
