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
    dyn = api.DynamicOperator(ctx, geom, "siddon")   # builds its tile layout
    log(f"matrix-free operator: {time.time() - t0:.1f}s")
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


def rel_l2(a, b, mask=None):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if mask is not None:
        a, b = a[mask], b[mask]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_leg(A32, b, dims, runs_all=5, iters_all=20, runs_one=3, iters_one=10):
    """The CPU port (oracle/, C + OpenMP) of the same TVAL3 loop on the same
    CSR on the host cores, loop-only time as the reference's loop timer
    (3d/tval3_gpu_3d.cpp:1159,1254-1257): one warm-up, then the median of
    `runs` runs of a fixed iteration count, FP32 and FP64, all threads and
    one thread.  Its FP64 solve to the default stopping rule is the FP64
    reference the GPU results are compared with."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tv3_oracle as orc
    rowptr, col, val = A32.download_csr()
    oop = orc.csr_op(A32.m, A32.n, rowptr, col, val)
    fixed = dict(mu=2.0 ** 9, mu0=2.0 ** 9, tv_iso=1, tol=0.0, tol_inn=0.0, maxcnt=1 << 30)
    threads = orc.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))

    def rate(dtype, nthr, runs, iters):
        orc.set_threads(nthr)
        o = orc.default_opts(**fixed, maxit=iters)
        orc.solve(oop, dims, b, o, dtype=dtype)                 # warm-up
        r = []
        for _ in range(runs):
            _, info, _, _ = orc.solve(oop, dims, b, o, dtype=dtype)
            r.append(iters / info.loop_secs)
        return float(np.median(r))

    t0 = time.perf_counter()
    res = {"f32_all": rate(np.float32, threads, runs_all, iters_all),
           "f64_all": rate(np.float64, threads, runs_all, iters_all),
           "f32_one": rate(np.float32, 1, runs_one, iters_one),
           "f64_one": rate(np.float64, 1, runs_one, iters_one)}
    tcpu = time.perf_counter() - t0
    orc.set_threads(threads)
    conv = orc.default_opts(mu=2.0 ** 9, mu0=2.0 ** 9, tv_iso=1)
    t0 = time.perf_counter()
    U64, i64, _, _ = orc.solve(oop, dims, b.astype(np.float64), conv)
    tref = time.perf_counter() - t0
    log(f"cpu: FP32 {res['f32_all']:.2f} it/s ({threads} threads), FP64 {res['f64_all']:.2f}; "
        f"one thread FP32 {res['f32_one']:.2f}, FP64 {res['f64_one']:.2f} ({tcpu:.0f}s); "
        f"FP64 reference solve {i64.total_iters} inner / {i64.outer_iters} outer ({tref:.0f}s)")
    cpu = {"value": res["f32_all"], "unit": "inner iter/s", "cores": threads, "kind": "port",
           "sample": f"oracle TVAL3 loop (C, OpenMP) on the same {dims[0]}^3 Siddon CSR "
                     f"({A32.nnz} nnz, FP32 weights), loop-only, 1 warm-up then median of "
                     f"{runs_all} runs x {iters_all} inner iterations ({threads} threads) and "
                     f"of {runs_one} x {iters_one} (1 thread)",
           "f64": res["f64_all"], "f32_1core": res["f32_one"], "f64_1core": res["f64_one"],
           "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
           "cpu_model": cpu_model()}
    ref = {"inner": i64.total_iters, "outer": i64.outer_iters, "secs": tref}
    return cpu, U64, ref


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--emitters", type=int, default=256)
    ap.add_argument("--receivers", type=int, default=1024)
    ap.add_argument("--path", default="fused-f16", choices=["plain-f32", "fused-f32", "fused-f16"])
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
    copy_bw, read_bw = ctx.bandwidth_probe(1 << 30, 10)
    log(f"HBM probe (1 GiB): copy {copy_bw:.0f} GB/s, read {read_bw:.0f} GB/s")
    N = args.size
    dims = (N, N, N)
    u, geom, dyn, A32, b = build_workload(ctx, N, args.emitters, args.receivers)
    m, n, nnz = A32.m, A32.n, A32.nnz
    dlay = dyn.prepare(dims)           # tile layout of the matrix-free operator (built at creation)
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
    log(f"timed: {args.steps} iters in {t:.4f}s -> {value:.1f} iter/s "
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
    # the panel kernels: k_panel12 with packed entries (layout bits 8 and 9)
    kA = "k_panel12" if panels & 256 else "k_panel"
    kT = "k_panel12" if panels & 512 else "k_panel"
    if tAT >= tA:
        dom_t, dom_b, dom_key = tAT, bAT, "AT"
        dom_name = (f"A^T product, panel form ({kT}, rays staged in LDS)" if panels & 2
                    else "A^T product (CSC row blocks)")
    else:
        dom_t, dom_b, dom_key = tA, bA, "A"
        dom_name = (f"A product, panel form ({kA} segment sums + k_seg_sum)" if panels & 1
                    else "A product (CSR row blocks)")
    traffic = measured_traffic(dom_key, N, args.path)

    extra = {}
    runs = {}
    if not args.no_extra:
        # equal rel-err: the fast path against the reference-shaped FP32 loop,
        # both to the default stopping rule (tol 1e-6, tol_inn 1e-3, maxcnt 10,
        # maxit 1025), on b and three rounding-level (1e-7) perturbations of it
        conv = dict(mu=2.0 ** 9, mu0=2.0 ** 9, tv_mode="iso")
        for seed in (None, 0, 1, 2):
            if seed is None:
                bd = b_dev
            else:
                xi = np.random.default_rng(seed).choice([-1.0, 1.0], size=b.shape)
                bd = torch.from_numpy((b.astype(np.float64) * (1 + 1e-7 * xi)).astype(np.float32)).cuda()
            tf, Uf, inf16 = timed_solve(ctx, op, bd, dims,
                                        api.default_opts(**dict(conv, solver=solver)), ut_dev)
            tp, Up, inp = timed_solve(ctx, A32, bd, dims,
                                      api.default_opts(**dict(conv, solver="plain")), ut_dev)
            runs[seed] = {"fast": (tf, Uf.cpu().numpy(), inf16), "plain": (tp, Up.cpu().numpy(), inp)}
        # the reference-shaped FP32 loop and the dynamic Siddon operator
        kp = min(args.steps, 100)
        tp, _, ip = timed_solve(ctx, A32, b_dev, dims,
                                api.default_opts(**dict(fixed, solver="plain", maxit=kp)), ut_dev)
        extra["plain_f32_iter_per_s"] = kp / tp
        # the matrix-free Siddon operator (weights recomputed in every product,
        # tile products without atomics): fused and plain solves
        kd = min(args.steps, 100)
        dfixed = dict(fixed, solver="fused", maxit=kd)
        timed_solve(ctx, dyn, b_dev, dims, api.default_opts(**dict(dfixed, maxit=5)), ut_dev)
        td, _, idn = timed_solve(ctx, dyn, b_dev, dims, api.default_opts(**dfixed), ut_dev)
        extra["dynamic_fused_f32_iter_per_s"] = kd / td
        extra["dynamic_fused_f32_loop_iter_per_s"] = idn["total_iters"] / idn["loop_secs"]
        kp2 = min(args.steps, 20)
        tdp, _, _ = timed_solve(ctx, dyn, b_dev, dims,
                                api.default_opts(**dict(fixed, solver="plain", maxit=kp2)), ut_dev)
        extra["dynamic_plain_f32_iter_per_s"] = kp2 / tdp
        tdc, Udyn, idc = timed_solve(ctx, dyn, b_dev, dims,
                                     api.default_opts(**dict(conv, solver="fused")), ut_dev)
        runs["dynamic"] = (tdc, Udyn.cpu().numpy(), idc)
        tdn = dyn.layout_time(dims, False, 20)
        tdt = dyn.layout_time(dims, True, 20)
        extra["dynamic_A_us"] = tdn * 1e6
        extra["dynamic_AT_us"] = tdt * 1e6
        extra["dynamic_csr_equiv_GBps"] = [bytes_A(nnz, m, n, 4) / tdn / 1e9,
                                           bytes_AT(nnz, m, n, 4) / tdt / 1e9]
        extra["dynamic_layout"] = dlay
        log(f"plain-f32 {extra['plain_f32_iter_per_s']:.1f} it/s; matrix-free fused "
            f"{extra['dynamic_fused_f32_iter_per_s']:.1f} it/s (loop "
            f"{extra['dynamic_fused_f32_loop_iter_per_s']:.1f}), plain "
            f"{extra['dynamic_plain_f32_iter_per_s']:.1f} it/s (A {tdn * 1e6:.0f} us, "
            f"A^T {tdt * 1e6:.0f} us)")

    cpu, U64, ref64 = None, None, None
    if not args.no_cpu:
        cpu, U64, ref64 = cpu_baseline_leg(A32, b, dims)

    if runs:
        # voxels some ray crosses (non-zero column of A); the rest are fixed
        # by the TV term alone
        _, col, _ = A32.download_csr()
        covered = np.bincount(col, minlength=n) > 0
        ut = u.astype(np.float64)

        def errs(Ux):
            e = {"vs_utrue": rel_l2(Ux, ut), "vs_utrue_covered": rel_l2(Ux, ut, covered)}
            if U64 is not None:
                e["vs_fp64"] = rel_l2(Ux, U64)
                e["vs_fp64_covered"] = rel_l2(Ux, U64, covered)
            return e

        per = {"fast": [], "plain_f32": []}
        for seed in (None, 0, 1, 2):
            for key, name in (("fast", "fast"), ("plain", "plain_f32")):
                tt, Ux, inf_ = runs[seed][key]
                per[name].append(dict(errs(Ux), inner=inf_["total_iters"], outer=inf_["outer_iters"],
                                      secs=tt, outer_per_s=inf_["outer_iters"] / tt,
                                      inner_per_s=inf_["total_iters"] / tt, seed=seed))
        med = lambda rows, k: float(np.median([r[k] for r in rows]))
        ee = {"stopping": "tol 1e-6, tol_inn 1e-3, maxcnt 10, maxit 1025",
              "perturbations": "b and b(1 + 1e-7 xi), xi = +-1 seeds 0, 1, 2; medians over the four",
              "fast_path": args.path, "runs": per}
        for name in ("fast", "plain_f32"):
            ee[name] = {k: med(per[name], k) for k in per[name][0] if k != "seed"}
        ee["err_diff"] = ee["fast"]["vs_utrue"] - ee["plain_f32"]["vs_utrue"]
        ee["err_diff_covered"] = ee["fast"]["vs_utrue_covered"] - ee["plain_f32"]["vs_utrue_covered"]
        ee["speedup_outer"] = ee["fast"]["outer_per_s"] / ee["plain_f32"]["outer_per_s"]
        ee["pass"] = bool(ee["err_diff"] <= 1e-3 and ee["speedup_outer"] >= 1.8)
        extra["equal_error"] = ee
        tdc, Udyn, idc = runs["dynamic"]
        extra["dynamic_fused_f32_errors"] = dict(errs(Udyn), inner=idc["total_iters"],
                                                 outer=idc["outer_iters"], secs=tdc)
        if ref64 is not None:
            extra["fp64_reference"] = dict(ref64, rel_err_vs_utrue=rel_l2(U64, ut),
                                           rel_err_vs_utrue_covered=rel_l2(U64, ut, covered),
                                           covered_voxels=int(covered.sum()))
        log(f"equal error (medians): fast err {ee['fast']['vs_utrue']:.4e} (covered "
            f"{ee['fast']['vs_utrue_covered']:.4e}, vs FP64 {ee['fast'].get('vs_fp64', float('nan')):.2e}); "
            f"plain FP32 err {ee['plain_f32']['vs_utrue']:.4e} (covered {ee['plain_f32']['vs_utrue_covered']:.4e}, "
            f"vs FP64 {ee['plain_f32'].get('vs_fp64', float('nan')):.2e}); outer-iter speedup "
            f"{ee['speedup_outer']:.2f}; pass {ee['pass']}")

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
                     "kernel": dom_name, "achievable_read_GBps": read_bw,
                     "achievable_copy_GBps": copy_bw,
                     "frac_of_read_probe": dom_b / dom_t / 1e9 / read_bw},
        "cpu_baseline": cpu,
        "A_GBps": gA / 1e9,
        "AT_GBps": gAT / 1e9,
        "A_us": tA * 1e6,
        "AT_us": tAT * 1e6,
        "rel_err_vs_utrue": info["rel_error"],
        "loop_iter_per_s": info["total_iters"] / info["loop_secs"],
        "lanes": list(lanes),
        "layout": layout,
        "hbm_probe_GBps": {"copy": copy_bw, "read": read_bw, "bytes": 1 << 30},
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
