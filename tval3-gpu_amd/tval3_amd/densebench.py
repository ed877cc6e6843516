"""Dense multi-right-hand-side products on the matrix cores (tv3_matvec_multi).

    python -m tval3_amd.densebench [--rows 8192] [--cols 32768] [--rhs 64] [--reps 20]

A dense BF16 operator of rows x cols (i.i.d. N(0, 1/rows), the shared dense A
of the batched workload at 32^3 voxels) times `rhs` FP32 columns, in both
directions: Y = A X (a row-major BF16 copy of A) and Z = A^T R (A's own
column-major storage).  Each product is timed with HIP events on the current
stream (median of `reps`), and reported as HBM GB/s of A (the operator is
read once per product) and as the matrix-core rate of the BF16 MFMAs issued
(each FP32 value of X enters as three BF16 terms: 3 x 2 rows cols rhs
flops).  Prints one JSON object.
"""
import argparse
import json

import numpy as np


def main():
    import torch
    from . import api
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", type=int, default=8192)
    ap.add_argument("--cols", type=int, default=32768)
    ap.add_argument("--rhs", type=int, default=64)
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args()
    ctx = api.Context(0)
    rng = np.random.default_rng(4)
    A = np.asfortranarray(rng.normal(0.0, 1.0 / np.sqrt(args.rows),
                                     size=(args.rows, args.cols)).astype(np.float32))
    op = api.DenseOperator(ctx, A, "bf16")
    del A
    out = {"rows": args.rows, "cols": args.cols, "rhs": args.rhs, "weights": "bf16"}
    for trans in (False, True):
        xl, yl = (args.rows, args.cols) if trans else (args.cols, args.rows)
        X = torch.randn(args.rhs, xl, device="cuda").t()          # columns contiguous
        Y = torch.empty(args.rhs, yl, device="cuda").t()
        for _ in range(3):
            op.mat_mat_mul(X, Y, trans)
        times = []
        for _ in range(args.reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            op.mat_mat_mul(X, Y, trans)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) * 1e-3)
        t = float(np.median(times))
        a_bytes = 2.0 * args.rows * args.cols
        flops = 3 * 2.0 * args.rows * args.cols * args.rhs
        out["AT" if trans else "A"] = {"us": t * 1e6, "A_GBps": a_bytes / t / 1e9,
                                       "mfma_TFLOPs": flops / t / 1e12}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
