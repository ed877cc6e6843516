"""On-disk form of a TVAL3 problem: one JSON manifest plus raw little-endian
binary arrays (the reference defines no file format; callers assemble host
arrays in memory, 3d/tval3_gpu_3d.cpp:1430-1470).

Layout of a problem named P in directory D:

    D/P.json          manifest (format "tval3-problem", version 1)
    D/P.<array>.bin   one file per array, C order, dtype and shape in the manifest

Arrays: the sparse matrix as CSR (``rowptr`` int32 m+1, ``col`` int32 nnz,
``val`` float32/float16/bfloat16-as-uint16/float64 nnz), ``b`` (m), optional
``u_true`` (n), and optional geometry arrays (``pos_emitters``,
``pos_receivers`` float32 (num_mps*E, 3), ``use_path`` uint8 E*R).  The
manifest records dims (dim_y, dim_x, dim_z), m, n, nnz, the weight type, the
geometry scalars and a SHA-256 of every array file, checked on load.
Loading reads files with numpy only; nothing in a file is executed.
"""
import hashlib
import json
import os

import numpy as np

FORMAT = "tval3-problem"
VERSION = 1
_DTYPES = {"f32": "<f4", "f64": "<f8", "f16": "<f2", "bf16": "<u2", "i32": "<i4", "u8": "|u1"}


def _sha(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def _write(directory, name, key, arr, kind):
    arr = np.ascontiguousarray(arr, dtype=np.dtype(_DTYPES[kind]))
    fn = f"{name}.{key}.bin"
    path = os.path.join(directory, fn)
    arr.tofile(path)
    return {"file": fn, "dtype": kind, "shape": list(arr.shape), "sha256": _sha(path)}


def save_problem(directory, name, dims, rowptr, col, val, b, val_type="f32", u_true=None,
                 geometry=None, meta=None):
    """Write a problem.  val_type is the stored weight type: f32, f64, f16
    or bf16 (bf16 values are passed as their uint16 bit patterns or as
    floats, which are rounded to nearest even).  geometry: a dict as
    tval3_amd.workload.transducer_geometry returns (optional)."""
    os.makedirs(directory, exist_ok=True)
    rowptr = np.asarray(rowptr)
    m = len(rowptr) - 1
    n = int(np.prod(dims))
    nnz = int(rowptr[-1])
    if len(col) != nnz or len(val) != nnz or len(b) != m:
        raise ValueError("rowptr, col, val and b do not agree in size")
    if val_type == "bf16" and np.asarray(val).dtype != np.uint16:
        v32 = np.ascontiguousarray(val, np.float32).view(np.uint32)
        val = ((v32 + 0x7FFF + ((v32 >> 16) & 1)) >> 16).astype(np.uint16)
    arrays = {
        "rowptr": _write(directory, name, "rowptr", rowptr, "i32"),
        "col": _write(directory, name, "col", col, "i32"),
        "val": _write(directory, name, "val", val, val_type),
        "b": _write(directory, name, "b", b, "f32"),
    }
    if u_true is not None:   # flat, dim_y fastest
        arrays["u_true"] = _write(directory, name, "u_true", np.ravel(u_true), "f32")
    geo = None
    if geometry is not None:
        g = geometry
        geo = {k: (list(map(float, g[k])) if isinstance(g[k], (tuple, list, np.ndarray))
                   else g[k]) for k in ("E", "R", "mps", "origin", "voxel") if k in g}
        geo["rv"] = [int(v) for v in g["rv"]]
        arrays["pos_emitters"] = _write(directory, name, "pos_emitters", g["e_pos"], "f32")
        arrays["pos_receivers"] = _write(directory, name, "pos_receivers", g["r_pos"], "f32")
        if g.get("use_path") is not None:
            arrays["use_path"] = _write(directory, name, "use_path", g["use_path"], "u8")
    manifest = {"format": FORMAT, "version": VERSION, "dims": [int(d) for d in dims], "m": m,
                "n": n, "nnz": nnz, "val_type": val_type, "arrays": arrays, "geometry": geo,
                "meta": meta or {}}
    with open(os.path.join(directory, f"{name}.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    return manifest


def load_problem(directory, name, verify=True):
    """Read a problem written by save_problem: returns the manifest with each
    array entry replaced by the numpy array (memory-mapped read-only)."""
    with open(os.path.join(directory, f"{name}.json")) as f:
        man = json.load(f)
    if man.get("format") != FORMAT or man.get("version") != VERSION:
        raise ValueError(f"not a {FORMAT} v{VERSION} manifest")
    out = dict(man)
    out["arrays"] = {}
    for key, e in man["arrays"].items():
        path = os.path.join(directory, e["file"])
        if verify and _sha(path) != e["sha256"]:
            raise ValueError(f"checksum mismatch for {e['file']}")
        arr = np.memmap(path, dtype=np.dtype(_DTYPES[e["dtype"]]), mode="r",
                        shape=tuple(e["shape"]))
        out["arrays"][key] = arr
    a = out["arrays"]
    if len(a["rowptr"]) != man["m"] + 1 or len(a["col"]) != man["nnz"] or \
            len(a["val"]) != man["nnz"] or len(a["b"]) != man["m"]:
        raise ValueError("array sizes disagree with the manifest")
    if a["rowptr"][0] != 0 or a["rowptr"][-1] != man["nnz"] or np.any(np.diff(a["rowptr"]) < 0):
        raise ValueError("rowptr is not a valid CSR pointer array")
    if man["nnz"] and (a["col"].min() < 0 or a["col"].max() >= man["n"]):
        raise ValueError("column index out of range")
    return out


def operator_from_problem(ctx, prob):
    """A SparseOperator (tval3_amd.api) on the device from a loaded problem."""
    from . import api
    a = prob["arrays"]
    vt = prob["val_type"]
    val = np.asarray(a["val"])
    if vt == "bf16":
        val = (val.astype(np.uint32) << 16).view(np.float32)
    return api.SparseOperator(ctx, prob["m"], prob["n"], np.asarray(a["rowptr"]),
                              np.asarray(a["col"]), val.astype(np.float64), dtype=vt)
