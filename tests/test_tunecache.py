"""The JSON tuning cache (tval3_amd.tunecache).

On the CPU: keys, the file round trip, replay instead of timing, a refused
entry timed again, unreadable files and other format versions ignored
(stand-in operators record what the cache asks of them).  On the GPU: a
fresh operator of the same matrix replays the stored shapes through
tv3_op_layout_force, which builds the layout without timing, and its products
equal the timed layout's bit for bit; a batch replays its shapes likewise.
"""
import json

import numpy as np
import pytest

from tval3_amd import tunecache
from tval3_amd._lib import TV3Error


class _Ctx:
    device = 0


class _Op:
    """stand-in operator: prepare() 'times' (counted), layout_force() replays"""

    def __init__(self, refuse=False):
        self.ctx, self.kind, self.m, self.n, self.nnz, self.dtype = _Ctx(), "csr", 100, 64, 999, "f16"
        self.timed, self.forced, self.refuse = 0, [], refuse

    def prepare(self, dims):
        self.timed += 1
        return dict(brick=8, panels=35, code_N=-4096, code_T=-2048, npan=1, nseg=2, max_seg=3)

    def layout_force(self, dims, code_N, code_T, pan_nnz, panels):
        if self.refuse:
            raise TV3Error(-1, "refused")
        self.forced.append((tuple(dims), code_N, code_T, pan_nnz, panels))
        return dict(brick=8, panels=panels, code_N=code_N, code_T=code_T, npan=1, nseg=2,
                    max_seg=3)


def test_cache_times_once_then_replays(tmp_path):
    path = tmp_path / "tune.json"
    c = tunecache.TuneCache(str(path), device="gpu0")
    op = _Op()
    info = c.prepare(op, (4, 4, 4))
    assert op.timed == 1 and not op.forced and info["panels"] == 35
    d = json.loads(path.read_text())
    assert d["version"] == tunecache.VERSION
    (key, e), = d["entries"].items()
    assert key == "layout|gpu0|csr:100x64:999:f16:4x4x4"
    assert e == dict(code_N=-4096, code_T=-2048, pan_nnz=16384, panels=35)
    # a new process: the stored shapes are replayed, nothing is timed
    op2 = _Op()
    info2 = tunecache.TuneCache(str(path), device="gpu0").prepare(op2, (4, 4, 4))
    assert op2.timed == 0 and op2.forced == [((4, 4, 4), -4096, -2048, 16384, 35)]
    assert info2["panels"] == 35
    # another GPU or another volume shape is another entry
    op3 = _Op()
    tunecache.TuneCache(str(path), device="gpu1").prepare(op3, (4, 4, 4))
    tunecache.TuneCache(str(path), device="gpu0").prepare(op3, (8, 4, 2))
    assert op3.timed == 2
    assert len(json.loads(path.read_text())["entries"]) == 3


def test_cache_refused_entry_is_timed_again(tmp_path):
    path = tmp_path / "tune.json"
    tunecache.TuneCache(str(path), device="g").prepare(_Op(), (4, 4, 4))
    op = _Op(refuse=True)
    info = tunecache.TuneCache(str(path), device="g").prepare(op, (4, 4, 4))
    assert op.timed == 1 and info["panels"] == 35


@pytest.mark.parametrize("text", ["not json", json.dumps({"version": 0, "entries": {"x": 1}}),
                                  json.dumps([1, 2])])
def test_cache_ignores_unreadable_or_other_versions(tmp_path, text):
    path = tmp_path / "tune.json"
    path.write_text(text)
    c = tunecache.TuneCache(str(path), device="g")
    assert c.entries == {}
    op = _Op()
    c.prepare(op, (4, 4, 4))
    assert op.timed == 1
    assert json.loads(path.read_text())["version"] == tunecache.VERSION


def test_cache_leaves_matrix_free_layouts_alone(tmp_path):
    path = tmp_path / "tune.json"
    op = _Op()
    op.kind = "dynamic"
    tunecache.TuneCache(str(path), device="g").prepare(op, (4, 4, 4))
    assert op.timed == 1 and not path.exists()


def test_batch_key_names_shared_operators():
    c = tunecache.TuneCache("unused.json", device="g")
    a, b = _Op(), _Op()
    k = c.batch_key(_Ctx(), [a, b, a], [(4, 4, 4), (4, 4, 4), (8, 8, 8)])
    assert k == ("batch|g|csr:100x64:999:f16:4x4x4;csr:100x64:999:f16:4x4x4;"
                 "csr:100x64:999:f16:8x8x8@0")


@pytest.mark.gpu
def test_cache_replay_on_gpu(tmp_path):
    import torch
    from tval3_amd import api, workload as wl
    ctx = api.Context(0)
    dims = (32, 32, 32)
    g = wl.transducer_geometry(48, 96, dims=dims)
    dyn = api.DynamicOperator(ctx, g, "siddon")
    path = str(tmp_path / "tune.json")
    A1 = dyn.materialize("f16")
    info1 = tunecache.TuneCache(path).prepare(A1, dims)
    A2 = dyn.materialize("f16")                 # a fresh operator: no layout yet
    info2 = tunecache.TuneCache(path).prepare(A2, dims)
    for k in ("brick", "panels", "code_N", "code_T"):
        assert info1[k] == info2[k], (k, info1, info2)
    rng = np.random.default_rng(1)
    x = torch.from_numpy(rng.uniform(-1, 1, A1.n).astype(np.float32)).cuda()
    r = torch.from_numpy(rng.uniform(-1, 1, A1.m).astype(np.float32)).cuda()
    for trans, v in ((False, x), (True, r)):
        y1 = A1.layout_mat_vec_mul(dims, v, trans=trans)
        y2 = A2.layout_mat_vec_mul(dims, v, trans=trans)
        torch.cuda.synchronize()
        assert torch.equal(y1, y2), trans
    # a batch of two problems: timed once, replayed with the same shapes
    probs = [(16, 16, 16), (32, 16, 16)]
    ops, bs = [], []
    for k, d in enumerate(probs):
        gg = wl.transducer_geometry(24, 64, dims=d, seed=k)
        op = api.DynamicOperator(ctx, gg, "siddon").materialize("f32")
        ops.append(op)
        bs.append(torch.rand(op.m, device="cuda"))
    o = api.default_opts(maxit=20)
    U1, i1, b1, _ = tunecache.TuneCache(path).batched(ctx, ops, bs, probs, o)
    U2, i2, b2, _ = tunecache.TuneCache(path).batched(ctx, ops, bs, probs, o)
    for k in ("lanes_N", "lanes_T", "panels_T", "panels_A", "panel_box_A"):
        assert b1[k] == b2[k], k
    for u1, u2 in zip(U1, U2):
        assert torch.equal(u1, u2)
