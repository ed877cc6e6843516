"""The fused solver's product kernels in its internal layout (8x8x8 tiles of
4x4x4 bricks, bricks, or natural voxel order; rays renumbered; panel forms:
A^T r with each tile's rays staged in LDS, A u as per-tile ray-segment sums
over the tile's voxels staged in LDS, then per-ray sums) against the oracle's
products.

Integer-valued weights and vectors make every product exact in FP32 whatever
the summation order, so those cases are compared bit-exactly and catch any
dropped, doubled or misrouted entry.  Real Siddon weights are compared within
the FP32 rounding bound gamma_k * (|A| |x|)_i, k = row length + 1."""
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


def T(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def H(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def siddon_structure(ctx, dims, E, R, seed=1):
    g = wl.transducer_geometry(E, R, dims=dims, seed=seed)
    dyn = api.DynamicOperator(ctx, g, "siddon")
    A = dyn.materialize("f32")
    rp, c, v = A.download_csr()
    return A.m, A.n, rp, c, v


def integer_case(ctx, dims, E, R, dtype):
    """Siddon sparsity pattern with weights 1..7 (exact in FP16 and BF16)."""
    m, n, rp, c, _ = siddon_structure(ctx, dims, E, R)
    w = (1 + (np.arange(len(c)) * 2654435761 % 7)).astype(np.float64)
    op = api.SparseOperator(ctx, m, n, rp, c, w, dtype=dtype)
    oop = orc.csr_op(m, n, rp, c, w)
    rng = np.random.default_rng(7)
    x = rng.integers(0, 1024, n).astype(np.float32)
    r = rng.integers(0, 1024, m).astype(np.float32)
    return op, oop, x, r


SHAPES = [
    # (code_N, code_T, pan_nnz, panels): both products in panel form with
    # three block sizes, panels split into brick runs (small pan_nnz), one
    # direction in panel form, and the plain kernels
    (-1024, -1024, 16384, 3), (-2048, -2048, 16384, 3), (-4096, -4096, 16384, 3),
    (-1024, -2048, 600, 3), (-2048, -1024, 100, 3),
    (32, -1024, 16384, 2), (-1024, 64, 16384, 1),
    (8, -2048, 0, 0), (-4096, 16, 0, 0),
    # panels over two tiles (A: bit 2, A^T: bit 3), whole and split into
    # brick runs
    (-2048, -2048, 16384, 7), (-1024, -1024, 300, 7), (-4096, 32, 16384, 5),
    (-2048, -2048, 16384, 15), (16, -1024, 300, 10), (-1024, -4096, 16384, 11),
    # A^T's entries packed (bit 9: 12-bit local ray indices, FP16 weights,
    # rows in groups of 4), over one and two tiles, whole and in brick runs
    (-2048, -2048, 16384, 3 | 512), (32, -4096, 300, 2 | 512), (-1024, -2048, 16384, 11 | 512),
    (-2048, -4096, 300, 15 | 512),
]


@pytest.mark.parametrize("dims,brick", [((16, 16, 16), 8), ((24, 16, 8), 8), ((12, 12, 12), 4),
                                        ((12, 10, 6), 1)])
def test_layout_products_exact(ctx, dims, brick):
    op, oop, x, r = integer_case(ctx, dims, 24, 64, "f16")
    y_ref = orc.matvec(oop, 0, x.astype(np.float64)).astype(np.float32)
    z_ref = orc.matvec(oop, 1, r.astype(np.float64)).astype(np.float32)
    assert np.abs(z_ref).max() < 2 ** 24 and np.abs(y_ref).max() < 2 ** 24
    info = op.prepare(dims)
    assert info["brick"] == brick
    for code_N, code_T, pan, panels in SHAPES:
        info = op.layout_force(dims, code_N, code_T, pan, panels)
        assert info["panels"] == panels
        y = H(op.layout_mat_vec_mul(dims, T(x)))
        z = H(op.layout_mat_vec_mul(dims, T(r), trans=True))
        assert (y == y_ref).all(), (dims, code_N, pan, info)
        assert (z == z_ref).all(), (dims, code_T, pan, info)


# A's panels over 2 x 2 x 1 tiles (bit 4) and 2 x 2 x 2 tiles (bit 5), whole
# and split into brick runs, with A^T in either form; bit 8: their entries
# packed (12-bit indices, FP16 weights); bit 9: A^T's entries packed
SHAPES_SUPER = [
    (-2048, -2048, 16384, 19), (-1024, -1024, 300, 19), (-4096, 16, 16384, 17),
    (-2048, -2048, 16384, 43), (-1024, -2048, 100, 33), (-2048, -1024, 16384, 35),
    (-4096, -2048, 16384, 299), (-2048, -1024, 300, 275), (-4096, 16, 100, 289),
    (-2048, -2048, 16384, 291), (-4096, -2048, 16384, 299 | 512),
    (-2048, -4096, 300, 291 | 512),
]


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 16, 32), (32, 32, 16), (48, 32, 64)])
def test_layout_supertile_panels_exact(ctx, dims):
    """A in panel form over boxes of 4 and 8 tiles: the tile-ordered voxel
    rows are regrouped by 2 x 2 x 2 blocks of tiles and the panels stage x
    through a tile map; integer inputs give the oracle's bits."""
    op, oop, x, r = integer_case(ctx, dims, 24, 64, "f16")
    y_ref = orc.matvec(oop, 0, x.astype(np.float64)).astype(np.float32)
    z_ref = orc.matvec(oop, 1, r.astype(np.float64)).astype(np.float32)
    assert np.abs(z_ref).max() < 2 ** 24 and np.abs(y_ref).max() < 2 ** 24
    op.prepare(dims)
    for code_N, code_T, pan, panels in SHAPES_SUPER:
        info = op.layout_force(dims, code_N, code_T, pan, panels)
        assert info["panels"] == panels
        y = H(op.layout_mat_vec_mul(dims, T(x)))
        z = H(op.layout_mat_vec_mul(dims, T(r), trans=True))
        assert (y == y_ref).all(), (dims, code_N, pan, info)
        assert (z == z_ref).all(), (dims, code_T, pan, info)


def test_layout_supertile_panels_need_even_tiles(ctx):
    """24 x 16 x 8 voxels hold 3 x 2 x 1 tiles: no 2 x 2 x 2 blocks, so the
    larger panels are refused, and two bits for A's panel size at once are
    an argument error."""
    dims = (24, 16, 8)
    op, _, _, _ = integer_case(ctx, dims, 8, 16, "f16")
    op.prepare(dims)
    with pytest.raises(TV3Error, match="even tile count"):
        op.layout_force(dims, -2048, -2048, 16384, 19)
    with pytest.raises(TV3Error, match="at most one"):
        op.layout_force(dims, -2048, -2048, 16384, 1 | 4 | 16)
    with pytest.raises(TV3Error, match="bit 4 or 5"):
        op.layout_force(dims, -2048, -2048, 16384, 1 | 4 | 256)


def test_layout_packed_entries_need_fp16(ctx):
    """packed entries hold FP16 or BF16 weights: an FP32 operator is refused
    in either direction; packed A^T entries need A^T's panel form and blocks
    of 2048 or 4096"""
    dims = (16, 16, 16)
    op, _, _, _ = integer_case(ctx, dims, 8, 16, "f32")
    with pytest.raises(TV3Error, match="FP16"):
        op.layout_force(dims, -2048, -2048, 16384, 1 | 32 | 256)
    with pytest.raises(TV3Error, match="FP16"):
        op.layout_force(dims, -2048, -2048, 16384, 2 | 512)
    op16, _, _, _ = integer_case(ctx, dims, 8, 16, "f16")
    with pytest.raises(TV3Error, match="bit 1"):
        op16.layout_force(dims, -2048, -2048, 16384, 1 | 512)
    with pytest.raises(TV3Error, match="2048 or 4096"):
        op16.layout_force(dims, -2048, -1024, 16384, 2 | 512)


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 16, 32)])
def test_layout_packed_bf16_exact(ctx, dims):
    """packed entries with BF16 weights (the upper half of an FP32): A over
    16^3 and 16 x 16 x 8 boxes and A^T over one and two tiles, integer
    inputs give the oracle's bits"""
    op, oop, x, r = integer_case(ctx, dims, 24, 64, "bf16")
    y_ref = orc.matvec(oop, 0, x.astype(np.float64)).astype(np.float32)
    z_ref = orc.matvec(oop, 1, r.astype(np.float64)).astype(np.float32)
    op.prepare(dims)
    for code_N, code_T, pan, panels in [(-4096, -4096, 16384, 1 | 2 | 32 | 256 | 512),
                                        (-2048, -2048, 300, 1 | 2 | 8 | 16 | 256 | 512),
                                        (-2048, -4096, 16384, 1 | 2 | 8 | 32 | 256 | 512)]:
        info = op.layout_force(dims, code_N, code_T, pan, panels)
        assert info["panels"] == panels
        y = H(op.layout_mat_vec_mul(dims, T(x)))
        z = H(op.layout_mat_vec_mul(dims, T(r), trans=True))
        assert (y == y_ref).all(), (dims, code_N, pan, info)
        assert (z == z_ref).all(), (dims, code_T, pan, info)


@pytest.mark.parametrize("unit_bits", [0, 8])
def test_layout_packed_transposed_equals_unpacked(ctx, unit_bits):
    """A^T with packed entries forms the same group sums and row sums as the
    4-byte form with the same block size, so real-valued products agree bit
    for bit (panels over one tile and over two; the packed plan caps a
    panel's rays at 4096)"""
    dims = (32, 32, 32)
    g = wl.transducer_geometry(48, 96, dims=dims)
    op = api.DynamicOperator(ctx, g, "siddon").materialize("f16")
    rng = np.random.default_rng(11)
    r = rng.uniform(-1, 1, op.m).astype(np.float32)
    op.prepare(dims)
    for code_T in (-2048, -4096):
        base = 2 | unit_bits
        i0 = op.layout_force(dims, -2048, code_T, 16384, base)
        assert i0["panels"] == base
        # no panel exceeds 4096 rays, so both forms share one plan and blocks
        assert i0["max_seg"] <= 4096
        z0 = H(op.layout_mat_vec_mul(dims, T(r), trans=True))
        i1 = op.layout_force(dims, -2048, code_T, 16384, base | 512)
        assert i1["panels"] == base | 512 and i1["npan"] == i0["npan"]
        z1 = H(op.layout_mat_vec_mul(dims, T(r), trans=True))
        assert np.array_equal(z0, z1), (code_T, unit_bits)


def test_layout_panels_with_empty_tiles(ctx):
    """Two emitters and four receivers leave most tiles without rays: empty
    panels still produce zeros for their voxels."""
    dims = (32, 16, 16)
    op, oop, x, r = integer_case(ctx, dims, 2, 4, "f32")
    y_ref = orc.matvec(oop, 0, x.astype(np.float64)).astype(np.float32)
    z_ref = orc.matvec(oop, 1, r.astype(np.float64)).astype(np.float32)
    assert (z_ref == 0).sum() > z_ref.size // 2
    for code, pan, panels in ((-1024, 16384, 3), (-2048, 64, 3), (-4096, 16384, 35),
                              (-2048, 64, 19)):
        info = op.layout_force(dims, code, code, pan, panels)
        assert info["panels"] == panels and info["npan"] >= 16
        y = H(op.layout_mat_vec_mul(dims, T(x)))
        z = H(op.layout_mat_vec_mul(dims, T(r), trans=True))
        assert (y == y_ref).all()
        assert (z == z_ref).all()


def test_layout_panels_first_tile_empty(ctx):
    """At 64^3 the corner tiles lie outside the transducer sphere, so no ray
    crosses them: the segment form of A has panels without rows (the first
    one at row 0), which must get no row blocks."""
    dims = (64, 64, 64)
    op, oop, x, r = integer_case(ctx, dims, 16, 32, "f16")
    y_ref = orc.matvec(oop, 0, x.astype(np.float64)).astype(np.float32)
    z_ref = orc.matvec(oop, 1, r.astype(np.float64)).astype(np.float32)
    for panels, code in ((3, -1024), (35, -4096)):   # A over tiles, over 16^3 boxes
        info = op.layout_force(dims, code, -1024, 16384, panels)
        assert info["panels"] == panels and info["brick"] == 8
        y = H(op.layout_mat_vec_mul(dims, T(x)))
        z = H(op.layout_mat_vec_mul(dims, T(r), trans=True))
        assert (y == y_ref).all()
        assert (z == z_ref).all()
    assert (z_ref[:8] == 0).all()          # voxels (0..7, 0, 0) of the corner tile


@pytest.mark.parametrize("dtype", ["f32", "f16", "bf16"])
def test_layout_siddon_rounding_bound(ctx, dtype):
    dims = (32, 32, 32)
    g = wl.transducer_geometry(48, 96, dims=dims)
    dyn = api.DynamicOperator(ctx, g, "siddon")
    op = dyn.materialize(dtype)
    rp, c, v = op.download_csr()          # the stored (rounded) weights
    oop = orc.csr_op(op.m, op.n, rp, c, v)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, op.n).astype(np.float32)
    r = rng.uniform(-1, 1, op.m).astype(np.float32)
    info = op.prepare(dims)
    u = 2.0 ** -24
    aop = orc.csr_op(op.m, op.n, rp, c, np.abs(v))
    lens_N = np.diff(rp)
    lens_T = np.bincount(c, minlength=op.n)
    for code_N, code_T, pan, panels in [(0, 0, 0, None)] + SHAPES[:3]:
        if panels is not None:
            op.layout_force(dims, code_N, code_T, pan, panels)
        y = H(op.layout_mat_vec_mul(dims, T(x))).astype(np.float64)
        z = H(op.layout_mat_vec_mul(dims, T(r), trans=True)).astype(np.float64)
        for got, trans, vec, lens in ((y, 0, x, lens_N), (z, 1, r, lens_T)):
            ref = orc.matvec(oop, trans, vec.astype(np.float64))
            mag = orc.matvec(aop, trans, np.abs(vec).astype(np.float64))
            k = lens + 1
            bound = k * u / (1 - k * u) * mag + 2 * u * np.abs(ref)   # + final FP32 store
            assert (np.abs(got - ref) <= bound + 1e-30).all(), (dtype, trans, code_T, info)
