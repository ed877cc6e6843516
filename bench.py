"""TVAL3 benchmark on one MI355X: 3-D reconstruction at 128^3 voxels.

Workload (synthetic, seeded): 128^3 slowness phantom (background 1/1500 s/m,
6 ellipsoids at +/-3 %), 256 emitters x 1024 receivers on a sphere of radius
1.1 around the [-1,1]^3 volume (262,144 rays), Siddon intersection-length
weights, b = A u_true + 0.5 % noise, isotropic TV, mu = 2^9.

A "step" is one TVAL3 inner iteration (1 A-product, 1 A^T-product, the
stencil/shrinkage/line-search work).  The timed region is one solve of
exactly --steps inner iterations (outer-loop continuation included, stopping
tests disabled so the count is fixed) with all inputs resident in HBM.

Prints ONE JSON line.  Extra keys: per-path iter/s, A / A^T kernel GB/s,
relative errors, the roofline object of the dominant kernel and the CPU
baseline (the oracle on the host cores, bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tval3-gpu_amd"))

import numpy as np  # noqa: E402

PEAK_HBM = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(ctx, N, E, R):
    import torch
    from tval3_amd import api, workload as wl
    t0 = time.time()
    u = wl.phantom_3d(N, 6).astype(np.float32)
    geom = wl.transducer_geometry(E, R, N=N)
    dyn = api.DynamicOperator(ctx, geom, "siddon")
    A32 = dyn.materialize("f32")
    ut = torch.from_numpy(u).cuda()
    Au = A32.mat_vec_mul(ut)
    torch.cuda.synchronize()
    b = wl.add_noise(Au.cpu().numpy().astype(np.float64), 0.005).astype(np.float32)
    log(f"workload: N={N} rays={A32.m} nnz={A32.nnz} ({time.time() - t0:.1f}s)")
    return u, geom, dyn, A32, b


def bytes_A(nnz, m, n, sv):
    """CSR product y = A x: entries (weight + int32 column), row pointers,
    x read once, y written."""
    return nnz * (sv + 4) + 4 * (m + 1) + 4 * n + 4 * m


def bytes_AT(nnz, m, n, sv):
    """CSC product y = A^T r: entries (weight + int32 row index), column
    pointers, r read once, y written."""
    return nnz * (sv + 4) + 4 * (n + 1) + 4 * m + 4 * n


def measured_traffic(key, N, path):
    """HBM bytes per launch of the dominant product kernel(s) from the
    committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, each scaled
    by what the counter reads for a copy of known size), or None when no
    summary for this workload exists."""
    f = os.path.join(ROOT, "profiles", f"traffic_{N}_{path}.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        d = json.load(fh)
    k = d.get("kernels", {}).get(key)
    if not k:
        return None
    return (k["fetch_kib"] * d["fetch_factor"] + k["write_kib"] * d["write_factor"]) * 1024.0


def time_spmv(op, trans, reps=50):
    """Average duration of one product, HIP events on the launch stream."""
    import torch
    x = torch.rand(op.m if trans else op.n, device="cuda")
    y = torch.empty(op.n if trans else op.m, device="cuda")
    for _ in range(3):
        op.mat_vec_mul(x, y, trans)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        op.mat_vec_mul(x, y, trans)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps / 1e3


def timed_solve(ctx, op, b_dev, dims, opts, ut_dev):
    import torch
    from tval3_amd import api
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    U, info, _ = api.tval3_gpu_3d(ctx, op, b_dev, dims, opts, Ut=ut_dev, history=False)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, U, info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--emitters", type=int, default=256)
    ap.add_argument("--receivers", type=int, default=1024)
    ap.add_argument("--path", default="fused-f16", choices=["plain-f32", "fused-f32", "fused-f16"])
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run 5 plain device copies of 256 MiB (reference for counters)")
    ap.add_argument("--shapes", default="auto",
                    help="fused layout launch shapes: 'auto' (timed on this GPU) or "
                         "'code_N,code_T,pan_nnz,panels' as Operator.layout_force takes them "
                         "(under a counter profiler the timings that pick shapes are skewed)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the solver has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch
    from tval3_amd import api
    if args.calibrate_copy:
        src = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    ctx = api.Context(0)
    N = args.size
    dims = (N, N, N)
    u, geom, dyn, A32, b = build_workload(ctx, N, args.emitters, args.receivers)
    m, n, nnz = A32.m, A32.n, A32.nnz
    b_dev = torch.from_numpy(b).cuda()
    ut_dev = torch.from_numpy(u).cuda()

    wdt = "f16" if args.path.endswith("f16") else "f32"
    op = A32 if wdt == "f32" else dyn.materialize(wdt)
    solver = "plain" if args.path.startswith("plain") else "fused"
    layout = None
    if solver == "fused":
        t0 = time.time()
        layout = op.prepare(dims)      # renumbering, panels, tuned shapes: setup, not timed
        if args.shapes != "auto":
            cn, ct, pn, pm = (int(v) for v in args.shapes.split(","))
            layout = op.layout_force(dims, cn, ct, pn, pm)
        lanes = (layout["code_N"], layout["code_T"])
        log(f"layout: {layout} ({time.time() - t0:.1f}s)")
    else:
        lanes = op.autotune()
    base = dict(mu=2.0 ** 9, mu0=2.0 ** 9, tv_mode="iso", solver=solver)
    # fixed iteration count: no outer cap, no stopping test
    fixed = dict(base, tol=0.0, tol_inn=0.0, maxcnt=1 << 30)

    log(f"path={args.path} lanes(N,T)={lanes} warmup={args.warmup} steps={args.steps}")
    timed_solve(ctx, op, b_dev, dims, api.default_opts(**dict(fixed, maxit=args.warmup)), ut_dev)
    t, U, info = timed_solve(ctx, op, b_dev, dims, api.default_opts(**dict(fixed, maxit=args.steps)),
                             ut_dev)
    assert info["total_iters"] == args.steps, info
    value = args.steps / t
    log(f"timed: {args.steps} iters in {t:.3f}s -> {value:.1f} iter/s "
        f"(loop-only {info['total_iters'] / info['loop_secs']:.1f}/s) rel_err={info['rel_error']:.4e}")

    # dominant kernels: the A and A^T products the solver runs, timed alone
    # with HIP events on the launch stream (layout_time) for the fused path
    sv = {"f32": 4, "f16": 2, "bf16": 2}[wdt]
    panels = layout["panels"] if layout else 0
    if layout:
        tA = op.layout_time(dims, False, 20)
        tAT = op.layout_time(dims, True, 20)
    else:
        tA = time_spmv(op, False)
        tAT = time_spmv(op, True)
    bA, bAT = bytes_A(nnz, m, n, sv), bytes_AT(nnz, m, n, sv)
    gA, gAT = bA / tA, bAT / tAT
    log(f"SpMV A: {tA * 1e6:.1f} us {gA / 1e9:.0f} GB/s; A^T: {tAT * 1e6:.1f} us {gAT / 1e9:.0f} GB/s")
    if tAT >= tA:
        dom_t, dom_b, dom_key = tAT, bAT, "AT"
        dom_name = ("A^T product, panel form (k_panel, rays staged in LDS)" if panels & 2
                    else "A^T product (CSC row blocks)")
    else:
        dom_t, dom_b, dom_key = tA, bA, "A"
        dom_name = ("A product, panel form (k_panel segment sums + k_seg_sum)" if panels & 1
                    else "A product (CSR row blocks)")
    traffic = measured_traffic(dom_key, N, args.path)

    extra = {}
    if not args.no_extra:
        # equal rel-err: FP16 fused against the reference-shaped FP32 loop, both
        # run to the default stopping rule (tol 1e-6, tol_inn 1e-3, maxcnt 10,
        # maxit 1025)
        conv = dict(mu=2.0 ** 9, mu0=2.0 ** 9, tv_mode="iso")
        tf, _, inf16 = timed_solve(ctx, op, b_dev, dims,
                                   api.default_opts(**dict(conv, solver=solver)), ut_dev)
        tp, _, inp = timed_solve(ctx, A32, b_dev, dims,
                                 api.default_opts(**dict(conv, solver="plain")), ut_dev)
        extra["equal_error"] = {
            "stopping": "tol 1e-6, tol_inn 1e-3, maxcnt 10, maxit 1025",
            "fast": {"path": args.path, "rel_err": inf16["rel_error"],
                     "inner": inf16["total_iters"], "outer": inf16["outer_iters"], "secs": tf,
                     "inner_per_s": inf16["total_iters"] / tf,
                     "outer_per_s": inf16["outer_iters"] / tf},
            "plain_f32": {"rel_err": inp["rel_error"], "inner": inp["total_iters"],
                          "outer": inp["outer_iters"], "secs": tp,
                          "inner_per_s": inp["total_iters"] / tp,
                          "outer_per_s": inp["outer_iters"] / tp},
        }
        ee = extra["equal_error"]
        ee["err_diff"] = ee["fast"]["rel_err"] - ee["plain_f32"]["rel_err"]
        ee["speedup_outer"] = ee["fast"]["outer_per_s"] / ee["plain_f32"]["outer_per_s"]
        ee["pass"] = bool(ee["err_diff"] <= 1e-3 and ee["speedup_outer"] >= 1.8)
        log(f"equal error: fast err {ee['fast']['rel_err']:.4e} ({ee['fast']['inner']} inner, "
            f"{ee['fast']['outer']} outer, {tf:.3f}s); plain FP32 err {ee['plain_f32']['rel_err']:.4e} "
            f"({ee['plain_f32']['inner']} inner, {ee['plain_f32']['outer']} outer, {tp:.3f}s); "
            f"outer-iter speedup {ee['speedup_outer']:.2f}")
        # the reference-shaped FP32 loop and the dynamic Siddon operator,
        # fixed iteration counts
        kp = min(args.steps, 100)
        tp, _, ip = timed_solve(ctx, A32, b_dev, dims,
                                api.default_opts(**dict(fixed, solver="plain", maxit=kp)), ut_dev)
        extra["plain_f32_iter_per_s"] = kp / tp
        kd = min(args.steps, 20)
        td, _, idn = timed_solve(ctx, dyn, b_dev, dims,
                                 api.default_opts(**dict(fixed, solver="plain", maxit=kd)), ut_dev)
        extra["dynamic_siddon_f32_iter_per_s"] = kd / td
        tdn = time_spmv(dyn, False, 10)
        tdt = time_spmv(dyn, True, 10)
        extra["dynamic_A_us"] = tdn * 1e6
        extra["dynamic_AT_us"] = tdt * 1e6
        extra["dynamic_csr_equiv_GBps"] = [bytes_A(nnz, m, n, 4) / tdn / 1e9,
                                           bytes_AT(nnz, m, n, 4) / tdt / 1e9]
        log(f"plain-f32 {extra['plain_f32_iter_per_s']:.1f} it/s; dynamic {extra['dynamic_siddon_f32_iter_per_s']:.1f} it/s "
            f"(A {tdn * 1e6:.0f} us, A^T {tdt * 1e6:.0f} us)")

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import tv3_oracle as orc
        rowptr, col, val = A32.download_csr()
        oop = orc.csr_op(m, n, rowptr, col, val)
        oo = orc.default_opts(mu=2.0 ** 9, mu0=2.0 ** 9, tv_iso=1, tol=0.0, tol_inn=0.0,
                              maxcnt=1 << 30, maxit=args.cpu_iters)
        t0 = time.perf_counter()
        orc.solve(oop, dims, b, oo, dtype=np.float32)
        tc = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": args.cpu_iters / tc, "unit": "inner iter/s", "cores": cores,
               "kind": "port",
               "sample": f"oracle FP32 TVAL3 (C, OpenMP) on the same {N}^3 Siddon CSR "
                         f"({nnz} nnz, FP32 weights), {args.cpu_iters} inner iterations "
                         f"including its setup (scaleb, A^T b), {tc:.1f}s"}
        log(f"cpu: {cpu['value']:.3f} iter/s on {cores} threads")

    out = {
        "metric": "TVAL3 iter/s + A,A^T HBM GB/s at 128^3 vox; fp16 vs fp32 at equal rel-err",
        "value": value,
        "unit": "inner iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded phantom + spherical transducer geometry, Siddon weights)",
        "config": {"workload": f"TVAL3 3-D {N}^3 voxels, {m} rays, Siddon CSR {nnz} nnz, "
                               f"{wdt} weights, {args.path} solver, iso TV mu=2^9",
                   "voxels": n, "rays": m, "nnz": nnz, "weights": wdt, "path": args.path,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": dom_b / dom_t / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": dom_b / dom_t / PEAK_HBM, "traffic": traffic,
                     "kernel": dom_name},
        "cpu_baseline": cpu,
        "A_GBps": gA / 1e9,
        "AT_GBps": gAT / 1e9,
        "A_us": tA * 1e6,
        "AT_us": tAT * 1e6,
        "rel_err_vs_utrue": info["rel_error"],
        "loop_iter_per_s": info["total_iters"] / info["loop_secs"],
        "lanes": list(lanes),
        "layout": layout,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
