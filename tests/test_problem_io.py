"""The on-disk problem format (JSON manifest + raw little-endian arrays):
round trips for every weight type, geometry, and rejection of corrupted or
inconsistent files."""
import json
import os

import numpy as np
import pytest

import tv3_oracle as orc
from tval3_amd import problem_io as pio, workload as wl


@pytest.fixture(scope="module")
def small():
    g = wl.transducer_geometry(8, 16, N=8)
    op = orc.siddon_op(g)
    rp, col, val = orc.siddon_csr(op)
    u = wl.phantom_3d(8, 3)
    b = orc.matvec(orc.csr_op(len(rp) - 1, 512, rp, col, val), 0, u).astype(np.float32)
    return g, rp, col, val, u, b


@pytest.mark.parametrize("vt", ["f32", "f64", "f16", "bf16"])
def test_round_trip(tmp_path, small, vt):
    g, rp, col, val, u, b = small
    pio.save_problem(tmp_path, "p", (8, 8, 8), rp, col, val, b, val_type=vt, u_true=u,
                     geometry=g, meta={"seed": 2})
    q = pio.load_problem(tmp_path, "p")
    a = q["arrays"]
    assert q["m"] == len(rp) - 1 and q["n"] == 512 and q["nnz"] == len(col)
    assert (a["rowptr"] == rp).all() and (a["col"] == col).all()
    assert (a["b"] == b).all() and (a["u_true"] == u.astype(np.float32)).all()
    if vt == "bf16":
        back = (np.asarray(a["val"]).astype(np.uint32) << 16).view(np.float32)
        assert np.allclose(back, val, rtol=2 ** -8)
    else:
        assert (a["val"] == val.astype(np.dtype(pio._DTYPES[vt]))).all()
    assert (a["pos_emitters"] == g["e_pos"].astype(np.float32)).all()
    assert q["geometry"]["rv"] == [int(v) for v in g["rv"]] and q["meta"] == {"seed": 2}


def test_rejects_corruption(tmp_path, small):
    g, rp, col, val, u, b = small
    pio.save_problem(tmp_path, "p", (8, 8, 8), rp, col, val, b)
    path = os.path.join(tmp_path, "p.col.bin")
    raw = bytearray(open(path, "rb").read())
    raw[0] ^= 1
    open(path, "wb").write(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        pio.load_problem(tmp_path, "p")
    # a consistent checksum over an out-of-range column is still rejected
    man = json.load(open(os.path.join(tmp_path, "p.json")))
    bad = np.array(col, np.int32)
    bad[5] = 10 ** 6
    bad.tofile(path)
    man["arrays"]["col"]["sha256"] = pio._sha(path)
    json.dump(man, open(os.path.join(tmp_path, "p.json"), "w"))
    with pytest.raises(ValueError, match="column"):
        pio.load_problem(tmp_path, "p")
    with pytest.raises(ValueError):
        pio.save_problem(tmp_path, "q", (8, 8, 8), rp, col[:-1], val, b)


@pytest.mark.gpu
def test_operator_from_file_matches_in_memory(tmp_path, small):
    import torch
    from tval3_amd import api
    g, rp, col, val, u, b = small
    ctx = api.Context(0)
    pio.save_problem(tmp_path, "p", (8, 8, 8), rp, col, val, b, val_type="f16")
    A_file = pio.operator_from_problem(ctx, pio.load_problem(tmp_path, "p"))
    A_mem = api.SparseOperator(ctx, len(rp) - 1, 512, rp, col, val, dtype="f16")
    x = torch.from_numpy(u.astype(np.float32)).cuda()
    y1 = A_file.mat_vec_mul(x).cpu().numpy()
    y2 = A_mem.mat_vec_mul(x).cpu().numpy()
    assert (y1 == y2).all()
