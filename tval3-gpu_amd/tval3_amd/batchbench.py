"""Batched TVAL3 on a batch of uneven problems (one persistent set of launches
per iteration, graph-captured iterations).

    python -m tval3_amd.batchbench [--problems 40] [--dense 8] [--rhs 8]
                                   [--maxit 100] [--seed 1000] [--no-fp64]
                                   [--no-sequential]

Workload: `problems` independent problems, problem i seeded with
seed + 1000 i: every volume edge drawn uniformly from {32, 48, 64, 80, 96}
(per axis), a slowness phantom, E x R rays on the sphere of radius 1.1 with
E x R nearest to 0.25 x voxels at E:R = 1:4, Siddon weights in a BF16 CSR
built on the device, b = A u_true + 0.5% noise; isotropic TV, mu = 2^9.
Plus `dense` problems of 32^3 voxels sharing one dense A (8192 x 32768,
i.i.d. N(0, 1/8192), stored BF16), each with `rhs` right-hand sides (its
phantom, `rhs` noise draws): dense x rhs TVAL3 instances whose products run
as one matrix-core product with dense x rhs columns (at most 64).

Reports the products' launch shapes picked by the timing autotuner,
aggregate inner iterations per second of the batch, the same problems solved
one after another (the batched solver on one problem at a time, and the
fused solver), and per problem the error against u_true and against the FP64
plain solve of the same problem (tv3_solve_f64, same weights), beside the
same figures for the reference-shaped FP32 solve (tv3_solve, plain): TVAL3's
step-size recursion makes FP32 and FP64 runs of 100 iterations drift apart,
and that drift is the scale to read the batch's against.  Prints one JSON
object.
"""
import argparse
import json
import time

import numpy as np

EDGES = (32, 48, 64, 80, 96)


def problem_shapes(count, seed):
    """(dims, E, R) of each problem: edges uniform over EDGES per axis; E x R
    nearest to n/4 with R = 4E."""
    out = []
    for i in range(count):
        rng = np.random.default_rng(seed + 1000 * i)
        dims = tuple(int(v) for v in rng.choice(EDGES, size=3))
        n = dims[0] * dims[1] * dims[2]
        E = max(1, int(round(np.sqrt(n / 16.0))))
        out.append((dims, E, 4 * E, seed + 1000 * i))
    return out


def build(ctx, shapes, dtype="bf16"):
    import torch

    from . import api, workload as wl
    probs = []
    for dims, E, R, sd in shapes:
        u = wl.phantom_3d(dims=dims, seed=sd)
        g = wl.transducer_geometry(E, R, dims=dims, seed=sd + 1)
        A = api.DynamicOperator(ctx, g, "siddon").materialize(dtype)
        ut = torch.tensor(u, dtype=torch.float32, device="cuda")
        Au = A.mat_vec_mul(ut)
        b = wl.add_noise(Au.double().cpu().numpy(), 0.005, seed=sd + 2).astype(np.float32)
        probs.append(dict(dims=dims, E=E, R=R, A=A, u=u, ut=ut, b=b,
                          bt=torch.from_numpy(b).cuda()))
    return probs


def build_dense(ctx, count, rhs, seed):
    """count problems of 32^3 voxels sharing one 8192 x 32768 BF16 Gaussian
    A; problem g: phantom seed + 7 g, right-hand side r: noise seed + 7 g + r."""
    import torch

    from . import api, workload as wl
    dims, m = (32, 32, 32), 8192
    n = 32768
    g = torch.Generator(device="cuda").manual_seed(seed)
    Ad = (torch.randn(m, n, generator=g, device="cuda") / np.sqrt(m)).to(torch.bfloat16)
    Ah = Ad.float().cpu().numpy()
    A = api.DenseOperator(ctx, Ah, dtype="bf16")
    A32 = api.DenseOperator(ctx, Ah, dtype="f32")     # the same values, for FP64 solves
    del Ah
    probs = []
    for k in range(count):
        u = wl.phantom_3d(dims=dims, seed=seed + 7 * k)
        ut = torch.tensor(u, dtype=torch.float32, device="cuda")
        Au = A.mat_vec_mul(ut).double().cpu().numpy()
        for r in range(rhs):
            b = wl.add_noise(Au, 0.005, seed=seed + 7 * k + r + 1).astype(np.float32)
            probs.append(dict(dims=dims, A=A, A32=A32, u=u, ut=ut, b=b,
                              bt=torch.from_numpy(b).cuda(), dense=True))
    del Ad
    return probs


def run(problems=40, maxit=100, seed=1000, fp64=True, sequential=True, dense=8, rhs=8,
        log=print):
    import torch

    from . import api
    ctx = api.Context(0)
    t0 = time.time()
    shapes = problem_shapes(problems, seed)
    probs = build(ctx, shapes)
    if dense > 0:
        probs += build_dense(ctx, dense, rhs, seed + 500)
    torch.cuda.synchronize()
    nvox = sum(int(np.prod(p["dims"])) for p in probs)
    nnz = sum(p["A"].nnz for p in probs if not p.get("dense"))
    rays = sum(p["A"].m for p in probs)
    log(f"{problems} sparse problems + {dense} x {rhs} dense right-hand sides: {nvox} voxels, "
        f"{rays} rays, {nnz} sparse nnz (BF16), built in {time.time() - t0:.1f}s")
    o = api.default_opts(mu=2.0 ** 9, mu0=2.0 ** 9, tv_mode="iso", maxit=maxit)
    ops = [p["A"] for p in probs]
    bs = [p["bt"] for p in probs]
    dims = [p["dims"] for p in probs]
    uts = [p["ut"] for p in probs]
    # warm-up (tunes the launch shapes), then the timed batch
    _, _, bi0, _ = api.tval3_gpu_3d_batched(ctx, ops, bs, dims, o, Uts=uts)
    lanes = (bi0["lanes_N"], bi0["lanes_T"])
    splits = (bi0["splits_N"], bi0["splits_T"]) if dense > 0 else None
    panels, panels_A = bi0["panels_T"] or -1, bi0["panels_A"] or -1
    Ub, ib, bi, _ = api.tval3_gpu_3d_batched(ctx, ops, bs, dims, o, Uts=uts, lanes=lanes,
                                             splits=splits, panels=panels, panels_A=panels_A,
                                             panel_box_A=bi0["panel_box_A"])
    out = dict(workload=dict(problems=problems, dense_problems=dense, rhs=rhs, voxels=nvox,
                             rays=rays, nnz=nnz, weights="bf16", maxit=maxit, seed=seed,
                             shapes=[list(p["dims"]) for p in probs if not p.get("dense")]),
               tuned=dict(lanes_N=bi0["lanes_N"], lanes_T=bi0["lanes_T"],
                          splits_N=bi0["splits_N"], splits_T=bi0["splits_T"],
                          panels_T=bi0["panels_T"], panels_A=bi0["panels_A"],
                          panel_box_A=bi0["panel_box_A"]),
               batch=dict(loop_s=bi["loop_secs"], total_iters=bi["total_iters"],
                          iter_per_s=bi["total_iters"] / bi["loop_secs"],
                          graph_launches=bi["graph_launches"], secs=bi["secs"],
                          host_secs=bi["host_secs"]))
    log(f"batch: {bi['total_iters']} inner iterations in {bi['loop_secs']:.3f}s -> "
        f"{out['batch']['iter_per_s']:.0f} iter/s (lanes {lanes}, splits {splits}, "
        f"A^T panels {bi['panels_T'] or 'off'}, A panels {bi['panels_A'] or 'off'} over "
        f"{bi['panel_box_A']} tiles)")
    sparse = [p for p in probs if not p.get("dense")]
    if dense > 0:
        dps = [p for p in probs if p.get("dense")]
        _, idn, bdn, _ = api.tval3_gpu_3d_batched(ctx, [p["A"] for p in dps],
                                                  [p["bt"] for p in dps],
                                                  [p["dims"] for p in dps], o,
                                                  Uts=[p["ut"] for p in dps], splits=splits)
        out["dense_group_alone"] = dict(loop_s=bdn["loop_secs"], total_iters=bdn["total_iters"],
                                        iter_per_s=bdn["total_iters"] / bdn["loop_secs"])
        log(f"dense group alone ({len(dps)} right-hand sides): "
            f"{out['dense_group_alone']['iter_per_s']:.0f} iter/s")
    if sequential:
        t_solo, it_solo = 0.0, 0
        for p in sparse:
            _, i1, b1, _ = api.tval3_gpu_3d_batched(ctx, [p["A"]], [p["bt"]], [p["dims"]], o,
                                                    lanes=lanes)
            t_solo += b1["loop_secs"]
            it_solo += b1["total_iters"]
        out["one_at_a_time"] = dict(loop_s=t_solo, total_iters=it_solo,
                                    iter_per_s=it_solo / t_solo)
        of = api.default_opts(mu=2.0 ** 9, mu0=2.0 ** 9, tv_mode="iso", maxit=maxit,
                              solver="fused")
        t_f, it_f = 0.0, 0
        for p in sparse:
            p["A"].prepare(p["dims"])
            _, inf, _ = api.tval3_gpu_3d(ctx, p["A"], p["bt"], p["dims"], of, history=False)
            t_f += inf["loop_secs"]
            it_f += inf["total_iters"]
        out["fused_one_at_a_time"] = dict(loop_s=t_f, total_iters=it_f, iter_per_s=it_f / t_f)
        log(f"sparse problems one at a time: batched solver "
            f"{out['one_at_a_time']['iter_per_s']:.0f} iter/s, fused solver "
            f"{out['fused_one_at_a_time']['iter_per_s']:.0f} iter/s")
    per = []
    for k, p in enumerate(probs):
        rec = dict(dims=list(p["dims"]), iters=ib[k]["total_iters"], outer=ib[k]["outer_iters"],
                   err_utrue=ib[k]["rel_error"])
        rec["dense"] = bool(p.get("dense"))
        if fp64 and p.get("dense") and k % rhs != 0:
            per.append(rec)          # FP64 reference for the first right-hand side of each
            continue
        if fp64:
            # the BF16 weights exactly, as F32 (FP64 products take F32 or F64 weights)
            if p.get("dense"):
                A32 = p["A32"]
            else:
                rp, col, val = p["A"].download_csr()
                A32 = api.SparseOperator(ctx, p["A"].m, p["A"].n, rp, col, val, dtype="f32")
            o64 = api.default_opts(mu=2.0 ** 9, mu0=2.0 ** 9, tv_mode="iso", maxit=maxit)
            U64, i64, _ = api.tval3_gpu_3d(ctx, A32, p["b"].astype(np.float64), p["dims"],
                                           o64, Ut=p["u"], history=False)
            Ug = Ub[k].double().cpu().numpy()
            rec["err_fp64"] = float(np.linalg.norm(Ug - U64) / np.linalg.norm(U64))
            rec["err_utrue_fp64"] = i64["rel_error"]
            rec["iters_fp64"] = i64["total_iters"]
            # the reference-shaped FP32 solve of the same problem: how far FP32
            # itself drifts from FP64 over these iterations
            U32, i32, _ = api.tval3_gpu_3d(ctx, A32, p["b"], p["dims"], o, Ut=p["u"],
                                           history=False)
            rec["plain32_err_fp64"] = float(np.linalg.norm(U32 - U64) / np.linalg.norm(U64))
            rec["plain32_err_utrue"] = i32["rel_error"]
            if not p.get("dense"):
                A32.close()
        per.append(rec)
    out["per_problem"] = per
    e = np.array([r["err_utrue"] for r in per])
    out["err_utrue_median"] = float(np.median(e))
    if fp64:
        e64 = np.array([r["err_fp64"] for r in per if "err_fp64" in r])
        d = np.array([r["err_utrue"] - r["err_utrue_fp64"] for r in per if "err_fp64" in r])
        p64 = np.array([r["plain32_err_fp64"] for r in per if "err_fp64" in r])
        dp = np.array([r["plain32_err_utrue"] - r["err_utrue_fp64"] for r in per
                       if "err_fp64" in r])
        out["err_fp64_median"] = float(np.median(e64))
        out["err_fp64_max"] = float(e64.max())
        out["err_utrue_minus_fp64_max"] = float(np.abs(d).max())
        out["err_utrue_minus_fp64_median"] = float(np.median(np.abs(d)))
        out["plain32_err_fp64_median"] = float(np.median(p64))
        out["plain32_err_fp64_max"] = float(p64.max())
        out["plain32_err_utrue_minus_fp64_median"] = float(np.median(np.abs(dp)))
        out["plain32_err_utrue_minus_fp64_max"] = float(np.abs(dp).max())
        log(f"|U - U_fp64| / |U_fp64|: batch median {out['err_fp64_median']:.2e} max "
            f"{out['err_fp64_max']:.2e}; plain FP32 solver median "
            f"{out['plain32_err_fp64_median']:.2e} max {out['plain32_err_fp64_max']:.2e}")
        log(f"error vs u_true: batch median {out['err_utrue_median']:.4f}; |err - err_fp64|: "
            f"batch median {out['err_utrue_minus_fp64_median']:.2e} max "
            f"{out['err_utrue_minus_fp64_max']:.2e}, plain FP32 median "
            f"{out['plain32_err_utrue_minus_fp64_median']:.2e} max "
            f"{out['plain32_err_utrue_minus_fp64_max']:.2e}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--problems", type=int, default=40)
    ap.add_argument("--dense", type=int, default=8)
    ap.add_argument("--rhs", type=int, default=8)
    ap.add_argument("--maxit", type=int, default=100)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-sequential", action="store_true")
    a = ap.parse_args(argv)
    out = run(a.problems, a.maxit, a.seed, not a.no_fp64, not a.no_sequential, a.dense, a.rhs,
              log=lambda s: print(s, flush=True))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
