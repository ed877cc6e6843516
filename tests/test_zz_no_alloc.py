"""tv3_matvec_noalloc: the products of tv3_matvec with no device allocation
(the dense A x keeps its split-K partial sums in the library's fixed per-GPU
scratch array).  Its outputs equal tv3_matvec's bit for bit over two calls
in a row, the second with a smaller operator (the array keeps the first
call's contents); it allocates no device memory; a dense A x whose partials
exceed the array is rejected before any launch."""
import numpy as np
import pytest

from tval3_amd import api
from tval3_amd._lib import TV3Error

pytestmark = pytest.mark.gpu


def _dense(ctx, m, n, dtype, seed):
    A = np.random.default_rng(seed).normal(size=(m, n)).astype(np.float32)
    return api.DenseOperator(ctx, np.asfortranarray(A), dtype)


def _csr(ctx, m, n, seed):
    rng = np.random.default_rng(seed)
    cnt = rng.integers(0, 40, size=m)
    rowptr = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    col = np.concatenate([np.sort(rng.choice(n, size=c, replace=False)) for c in cnt]).astype(np.int32)
    val = rng.normal(size=len(col))
    return api.SparseOperator(ctx, m, n, rowptr, col, val, "f32")


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_noalloc_equals_matvec_two_calls_second_smaller(dtype):
    import torch
    ctx = api.Context(0)
    big, small = _dense(ctx, 2048, 4096, dtype, 1), _dense(ctx, 300, 1000, dtype, 2)
    rng = np.random.default_rng(3)
    outs = []
    for op in (big, small):          # two calls in a row, the second smaller
        x = torch.from_numpy(rng.normal(size=op.n).astype(np.float32)).cuda()
        y = torch.full((op.m,), float("nan"), device="cuda")
        op.mat_vec_mul_noalloc(x, y)
        outs.append((op, x, y))
    torch.cuda.synchronize()
    for op, x, y in outs:
        want = op.mat_vec_mul(x)
        torch.cuda.synchronize()
        assert torch.equal(y, want), (op.m, op.n)
    # the transposed products and a sparse operator take the same kernels
    for op in (big, small, _csr(ctx, 500, 700, 4)):
        for trans in (False, True):
            x = torch.from_numpy(rng.normal(size=op.m if trans else op.n).astype(np.float32)).cuda()
            y = torch.empty(op.n if trans else op.m, device="cuda")
            op.mat_vec_mul_noalloc(x, y, trans)
            assert torch.equal(y, op.mat_vec_mul(x, trans=trans))


def test_noalloc_allocates_nothing_and_rejects_oversize():
    import torch
    ctx = api.Context(0)                 # a fresh context: empty workspace
    op = _dense(ctx, 8192, 8192, "bf16", 5)   # partials: 128 slabs x 8192 rows x 4 B = 4 MiB
    x = torch.ones(op.n, device="cuda")
    y1, y2 = torch.empty(op.m, device="cuda"), torch.empty(op.m, device="cuda")
    op.mat_vec_mul_noalloc(x, y1)        # first call: code objects loaded
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        op.mat_vec_mul_noalloc(x, y1)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    op.mat_vec_mul(x, y2)               # tv3_matvec grows the context workspace
    torch.cuda.synchronize()
    free2 = torch.cuda.mem_get_info()[0]
    print(f"free memory: before {free0}, after noalloc calls {free1}, after tv3_matvec {free2}")
    assert free1 == free0
    assert torch.equal(y1, y2)
    # partials of 34000 rows x 127 slabs x 4 B exceed the 16 MiB array
    big = api.DenseOperator(ctx, np.zeros((34000, 8193), np.float32, order="F"), "bf16")
    xb, yb = torch.ones(big.n, device="cuda"), torch.empty(big.m, device="cuda")
    with pytest.raises(TV3Error, match="allocation-free"):
        big.mat_vec_mul_noalloc(xb, yb)
    big.mat_vec_mul_noalloc(torch.ones(big.m, device="cuda"), torch.empty(big.n, device="cuda"),
                            trans=True)        # A^T x needs no scratch
