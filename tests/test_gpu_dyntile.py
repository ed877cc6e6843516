"""Matrix-free Siddon operator: the traversal shared by every product and
the deterministic tile products.

- The device CSR build (the FP32 traversal) equals the oracle's FP32
  restatement of the same traversal exactly: row pointers, columns and every
  weight bit for bit (orc_siddon_walk_csr_f32, pinned against the textbook
  Siddon merge in FP64 by tests/test_oracle.py).
- The tile products (no atomics) are the same matrix: A x and A^T r within
  the FP32 summation bound of the exact products of that CSR, bitwise the
  same on every run, with whole 8^3 tiles and with tiles split into 4^3
  bricks (B = 8) and with bricks as boxes (B = 4).
- A fused solve on the matrix-free operator (no matrix stored) tracks the
  oracle solving the traversal's CSR within its rounding envelope."""
import numpy as np
import pytest

import tv3_oracle as orc
from tval3_amd import api, workload as wl

from test_gpu_fused import first_divergence, _envelope_runs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available()
    return api.Context(0)


def T(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def H(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


# A x boxes: 2 x 2 x 2 tiles where every tile count is even ("tiles", "bricks",
# "box16"), tile pairs along y ("pairs"), single tiles ("singles"), bricks (B4)
GEOMS = {"tiles": dict(E=24, R=40, dims=(16, 16, 16)),       # few rays: whole 8^3 tiles
         "bricks": dict(E=96, R=256, dims=(16, 16, 16)),     # dense: tiles split into bricks
         "B4": dict(E=24, R=64, dims=(20, 12, 8)),           # bricks only (dims multiples of 4)
         "box16": dict(E=32, R=96, dims=(32, 16, 32)),
         "pairs": dict(E=32, R=96, dims=(16, 24, 8)),
         "singles": dict(E=32, R=96, dims=(24, 16, 8))}


def geom(name, seed=2):
    k = GEOMS[name]
    return wl.transducer_geometry(k["E"], k["R"], dims=k["dims"], seed=seed)


@pytest.mark.parametrize("name", ["tiles", "B4"])
def test_device_csr_equals_oracle_fp32_walk(ctx, name):
    g = geom(name)
    dyn = api.DynamicOperator(ctx, g, "siddon")
    rp, c, v = dyn.materialize("f32").download_csr()
    orp, oc, ov = orc.siddon_walk_csr(orc.siddon_op(g), np.float32)
    print(f"{name}: nnz device {len(c)} oracle {len(oc)}")
    assert np.array_equal(rp, orp) and np.array_equal(c, oc)
    assert np.array_equal(v.astype(np.float32), ov)


@pytest.mark.parametrize("name", ["tiles", "bricks", "B4", "box16", "pairs", "singles"])
def test_tile_products_same_matrix_and_repeatable(ctx, name):
    g = geom(name)
    dims = GEOMS[name]["dims"]
    dyn = api.DynamicOperator(ctx, g, "siddon")
    lay = dyn.prepare(dims)
    ntiles = int(np.prod(dims)) // 512
    print(f"{name}: layout {lay}")
    if name == "tiles":
        assert lay["npan"] == ntiles                     # no tile split
    elif name == "bricks":
        assert lay["npan"] > ntiles                      # tiles split into bricks
    elif name == "B4":
        assert lay["npan"] == int(np.prod(dims)) // 64   # B = 4: bricks
    else:
        assert lay["npan"] >= ntiles                     # A^T: tiles or their bricks
    orp, oc, ov = orc.siddon_walk_csr(orc.siddon_op(g), np.float32)
    oop = orc.csr_op(dyn.m, dyn.n, orp, oc, ov.astype(np.float64))
    aop = orc.csr_op(dyn.m, dyn.n, orp, oc, np.abs(ov.astype(np.float64)))
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, size=dyn.n).astype(np.float32)
    r = rng.normal(size=dyn.m).astype(np.float32)
    for trans, vec in ((False, x), (True, r)):
        y1 = H(dyn.mat_vec_mul(T(vec), trans=trans))
        y2 = H(dyn.mat_vec_mul(T(vec), trans=trans))
        assert np.array_equal(y1, y2)                    # the same bits every run
        want = orc.matvec(oop, int(trans), vec.astype(np.float64))
        bound = orc.matvec(aop, int(trans), np.abs(vec).astype(np.float64))
        k = 2 * int(max(np.max(np.diff(orp)), np.max(np.bincount(oc, minlength=dyn.n))))
        err = np.abs(y1 - want)
        print(f"{name} trans={trans}: max err {err.max():.2e}, bound "
              f"{(k * 2.0 ** -24 * bound).max():.2e}")
        assert (err <= k * 2.0 ** -24 * bound + 1e-30).all()
    # FP64 products (per ray, the FP32 traversal's weights): adjoint to 1e-12
    y64 = H(dyn.mat_vec_mul(T(x.astype(np.float64))))
    z64 = H(dyn.mat_vec_mul(T(r.astype(np.float64)), trans=True))
    lhs, rhs = y64 @ r, x.astype(np.float64) @ z64
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    assert np.abs(y64 - orc.matvec(oop, 0, x.astype(np.float64))).max() <= 1e-12 * np.abs(y64).max()


def test_fused_solve_on_matrix_free_operator_tracks_oracle(ctx):
    dims = (16, 16, 16)
    g = wl.transducer_geometry(48, 128, dims=dims, seed=3)
    u = wl.phantom_3d(dims=dims, seed=4)
    dyn = api.DynamicOperator(ctx, g, "siddon")
    orp, oc, ov = orc.siddon_walk_csr(orc.siddon_op(g), np.float32)
    oop = orc.csr_op(dyn.m, dyn.n, orp, oc, ov.astype(np.float64))
    b = wl.add_noise(orc.matvec(oop, 0, u), 0.005, seed=5).astype(np.float32)
    kw = dict(mu=2.0 ** 9, mu0=2.0 ** 9, tol_inn=1e-3, maxit=60)
    U, info, h = api.tval3_gpu_3d(ctx, dyn, b, dims,
                                  api.default_opts(**kw, tv_mode="iso", solver="fused"), Ut=u)
    U2, info2, _ = api.tval3_gpu_3d(ctx, dyn, b, dims,
                                    api.default_opts(**kw, tv_mode="iso", solver="fused"), Ut=u)
    assert np.array_equal(U, U2)                        # deterministic
    oo = orc.default_opts(**kw, tv_iso=1)
    Uo, io, ho, _ = orc.solve(oop, dims, b, oo, Ut=u, dtype=np.float32)
    runs = _envelope_runs(oop, dims, b, oo, u)
    k = min([first_divergence(h, ho)] + [first_divergence(rr[2], ho) for rr in runs])
    dev = np.abs(h["f"][:k] - ho["f"][:k]) / np.abs(ho["f"][:k])
    env = np.max([np.maximum.accumulate(np.abs(rr[2]["f"][:k] - ho["f"][:k]) / np.abs(ho["f"][:k]))
                  for rr in runs], axis=0)
    print(f"matrix-free fused: iters {info['total_iters']} oracle {io.total_iters}; decisions "
          f"agree for {k}; err {info['rel_error']:.4e} oracle {io.rel_error:.4e}; max f dev "
          f"{dev.max():.1e} envelope {env.max():.1e}; rel(U, oracle) {rel(U, Uo):.2e}")
    assert k >= 10
    assert (dev[:4] < 1e-5).all()
    assert (dev <= 30 * env + 1e-6).all()
    assert rel(U, Uo) <= max(10 * max(rel(rr[0], Uo) for rr in runs), 1e-4)
