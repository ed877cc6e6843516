"""ctypes binding of the CPU oracle (oracle/_build/libtv3oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results against.  The product path never imports
this module.

The C sources it binds restate the reference algorithm; each C function cites
the reference file:line it follows (oracle/tv3_oracle.h).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libtv3oracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class OrcOpts(C.Structure):
    _fields_ = [("mu", C.c_double), ("mu0", C.c_double), ("beta", C.c_double),
                ("beta0", C.c_double), ("rate_cnt", C.c_double), ("tol", C.c_double),
                ("tol_inn", C.c_double), ("maxcnt", C.c_int), ("maxit", C.c_int),
                ("c", C.c_double), ("gamma", C.c_double), ("gam", C.c_double),
                ("rate_gam", C.c_double), ("nonneg", C.c_int), ("tv_iso", C.c_int),
                ("numInitIterations", C.c_int), ("skipMulRatio", C.c_double),
                ("maxRelativeChange", C.c_double)]


class OrcInfo(C.Structure):
    _fields_ = [("total_iters", C.c_int), ("outer_iters", C.c_int), ("rel_chg", C.c_double),
                ("rel_error", C.c_double), ("backtracks", C.c_int), ("fallbacks", C.c_int),
                ("status", C.c_int), ("n_matvec_T", C.c_int), ("loop_secs", C.c_double)]


_HIST_D = ["f", "lam1", "lam2", "lam3", "lam4", "lam5", "tau", "alpha", "relchg", "mu",
           "beta", "C"]
# line-search record of each iteration (orc_history in tv3_oracle.h)
_HIST_LS = ["f_trial", "f_acc", "C_in", "c_armij", "dd", "gam"]
# state at the end of the solve
_HIST_STATE = ["sigma_out", "delta_out", "dts_out", "d_out"]


class OrcHistory(C.Structure):
    _fields_ = [("cap", C.c_int), ("count", C.c_int)] + \
               [(k, C.POINTER(C.c_double)) for k in _HIST_D] + \
               [("cnt", C.POINTER(C.c_int)), ("outer", C.POINTER(C.c_int)),
                ("mulT", C.POINTER(C.c_int))] + \
               [(k, C.POINTER(C.c_double)) for k in _HIST_LS + _HIST_STATE]


_P = C.c_void_p


class OrcOp(C.Structure):
    _fields_ = [("kind", C.c_int), ("m", C.c_int64), ("n", C.c_int64),
                ("dense", _P), ("lda", C.c_int64),
                ("rowptr", _P), ("col", _P), ("val", _P),
                ("colptr", _P), ("row", _P), ("cval", _P),
                ("num_emitters", C.c_int), ("num_receivers", C.c_int), ("num_mps", C.c_int),
                ("rv_x", C.c_int), ("rv_y", C.c_int), ("rv_z", C.c_int),
                ("scale_factor", C.c_double),
                ("ex", _P), ("ey", _P), ("ez", _P), ("rx", _P), ("ry", _P), ("rz", _P),
                ("epos", _P), ("rpos", _P),
                ("origin", C.c_double * 3), ("voxel", C.c_double * 3),
                ("use_path", _P),
                ("num_paths", C.c_int64), ("path_e", _P), ("path_r", _P)]


DENSE, CSR, BRESENHAM, SIDDON = 0, 1, 2, 3


def _declare(L):
    L.orc_opts_default.argtypes = [C.POINTER(OrcOpts)]
    L.orc_set_threads.restype = C.c_int
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_build_paths.restype = C.c_int64
    L.orc_build_paths.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, _P]
    L.orc_bresenham_ray.restype = C.c_int
    L.orc_bresenham_ray.argtypes = [C.c_int] * 9 + [_P, C.c_int]
    L.orc_siddon_ray.restype = C.c_int
    L.orc_siddon_ray.argtypes = [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int,
                                 C.c_double]
    L.orc_siddon_csr.restype = C.c_int64
    L.orc_siddon_csr.argtypes = [C.POINTER(OrcOp), _P, _P, _P]
    L.orc_csr_transpose.argtypes = [C.c_int64, C.c_int64, _P, _P, _P, _P, _P, _P]
    for s in ("_f64", "_f32"):
        getattr(L, "orc_matvec" + s).argtypes = [C.POINTER(OrcOp), C.c_int, _P, _P]
        getattr(L, "orc_D" + s).argtypes = [C.c_int] * 3 + [_P] * 4
        getattr(L, "orc_Dt" + s).argtypes = [C.c_int] * 3 + [_P] * 4
        getattr(L, "orc_shrink" + s).restype = C.c_double
        getattr(L, "orc_shrink" + s).argtypes = [C.c_int64, C.c_int] + [_P] * 6 + \
            [C.c_double] + [_P] * 3
        getattr(L, "orc_scaleb" + s).restype = C.c_double
        getattr(L, "orc_scaleb" + s).argtypes = [C.c_int64, _P]
        getattr(L, "orc_siddon_walk" + s).restype = C.c_int
        getattr(L, "orc_siddon_walk" + s).argtypes = [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int,
                                                      _P, _P, C.c_int]
        getattr(L, "orc_siddon_walk_csr" + s).restype = C.c_int64
        getattr(L, "orc_siddon_walk_csr" + s).argtypes = [C.POINTER(OrcOp), _P, _P, _P]
        getattr(L, "orc_solve" + s).restype = C.c_int
        getattr(L, "orc_solve" + s).argtypes = [C.POINTER(OrcOp), C.POINTER(OrcOpts),
                                                C.c_int, C.c_int, C.c_int, _P, _P, _P,
                                                C.POINTER(OrcInfo), C.POINTER(OrcHistory)]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _sfx(dtype):
    return "_f64" if np.dtype(dtype) == np.float64 else "_f32"


def set_threads(n=0):
    """OpenMP threads of the oracle (n > 0 sets); returns the count in effect"""
    return lib().orc_set_threads(int(n))


def default_opts(**kw):
    o = OrcOpts()
    lib().orc_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Operator:
    """Host-side operator description; keeps its numpy arrays alive."""

    def __init__(self, kind, m, n, **arrays):
        self.keep = {}
        self.s = OrcOp()
        self.s.kind = kind
        self.s.m = m
        self.s.n = n
        for k, v in arrays.items():
            if isinstance(v, np.ndarray):
                self.keep[k] = v
                setattr(self.s, k, _ptr(v))
            elif k in ("origin", "voxel"):
                getattr(self.s, k)[:] = list(v)
            else:
                setattr(self.s, k, v)


def dense_op(A):
    A = np.asfortranarray(A, dtype=np.float64)
    return Operator(DENSE, A.shape[0], A.shape[1], dense=A, lda=A.shape[0])


def csr_op(m, n, rowptr, col, val, with_csc=True):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    kw = dict(rowptr=rowptr, col=col, val=val)
    if with_csc:
        cp, rw, cv = csr_transpose(m, n, rowptr, col, val)
        kw.update(colptr=cp, row=rw, cval=cv)
    return Operator(CSR, m, n, **kw)


def build_paths(E, R, mps, use_path=None):
    up = None if use_path is None else np.ascontiguousarray(use_path, np.uint8)
    cnt = lib().orc_build_paths(E, R, mps, _ptr(up), None, None)
    pe = np.zeros(max(cnt, 1), np.int32)
    pr = np.zeros(max(cnt, 1), np.int32)
    lib().orc_build_paths(E, R, mps, _ptr(up), _ptr(pe), _ptr(pr))
    return pe[:cnt], pr[:cnt]


def bresenham_op(geom):
    """geom: dict with E, R, mps, rv (x,y,z), scale, e_vox (E*mps,3), r_vox, use_path."""
    pe, pr = build_paths(geom["E"], geom["R"], geom["mps"], geom.get("use_path"))
    rvx, rvy, rvz = geom["rv"]
    ev = np.ascontiguousarray(geom["e_vox"], np.int32)
    rv = np.ascontiguousarray(geom["r_vox"], np.int32)
    return Operator(BRESENHAM, len(pe), rvx * rvy * rvz, num_emitters=geom["E"],
                    num_receivers=geom["R"], num_mps=geom["mps"], rv_x=rvx, rv_y=rvy, rv_z=rvz,
                    scale_factor=float(geom["scale"]),
                    ex=np.ascontiguousarray(ev[:, 0]), ey=np.ascontiguousarray(ev[:, 1]),
                    ez=np.ascontiguousarray(ev[:, 2]), rx=np.ascontiguousarray(rv[:, 0]),
                    ry=np.ascontiguousarray(rv[:, 1]), rz=np.ascontiguousarray(rv[:, 2]),
                    num_paths=len(pe), path_e=pe, path_r=pr)


def siddon_op(geom):
    """geom: dict with E, R, mps, rv, e_pos (E*mps,3), r_pos, origin, voxel, use_path."""
    pe, pr = build_paths(geom["E"], geom["R"], geom["mps"], geom.get("use_path"))
    rvx, rvy, rvz = geom["rv"]
    return Operator(SIDDON, len(pe), rvx * rvy * rvz, num_emitters=geom["E"],
                    num_receivers=geom["R"], num_mps=geom["mps"], rv_x=rvx, rv_y=rvy, rv_z=rvz,
                    epos=np.ascontiguousarray(geom["e_pos"], np.float64),
                    rpos=np.ascontiguousarray(geom["r_pos"], np.float64),
                    origin=geom["origin"], voxel=geom["voxel"],
                    num_paths=len(pe), path_e=pe, path_r=pr)


def siddon_csr(op):
    P = op.s.num_paths
    rowptr = np.zeros(P + 1, np.int32)
    nnz = lib().orc_siddon_csr(C.byref(op.s), _ptr(rowptr), None, None)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float64)
    lib().orc_siddon_csr(C.byref(op.s), _ptr(rowptr), _ptr(col), _ptr(val))
    return rowptr, col[:nnz], val[:nnz]


def bresenham_ray(p0, p1, rv):
    cap = 2 * sum(rv) + 8
    out = np.zeros(cap, np.int32)
    c = lib().orc_bresenham_ray(*[int(v) for v in p0], *[int(v) for v in p1], *rv, _ptr(out),
                                cap)
    return out[:c]


def siddon_ray(p0, p1, origin, voxel, rv, min_len=0.0):
    cap = 2 * sum(rv) + 8
    vox = np.zeros(cap, np.int32)
    ln = np.zeros(cap, np.float64)
    a = np.ascontiguousarray(p0, np.float64)
    b = np.ascontiguousarray(p1, np.float64)
    o = np.ascontiguousarray(origin, np.float64)
    h = np.ascontiguousarray(voxel, np.float64)
    c = lib().orc_siddon_ray(_ptr(a), _ptr(b), _ptr(o), _ptr(h), *rv, _ptr(vox), _ptr(ln), cap,
                             min_len)
    return vox[:c], ln[:c]


def siddon_walk(p0, p1, origin, voxel, rv, dtype=np.float64):
    """the tracked-plane traversal of the device products in dtype (float32
    or float64): (voxel indices, lengths)"""
    cap = 2 * sum(rv) + 8
    vox = np.zeros(cap, np.int32)
    ln = np.zeros(cap, dtype)
    a, b, o, h = (np.ascontiguousarray(v, dtype) for v in (p0, p1, origin, voxel))
    c = getattr(lib(), "orc_siddon_walk" + _sfx(dtype))(_ptr(a), _ptr(b), _ptr(o), _ptr(h), *rv,
                                                         _ptr(vox), _ptr(ln), cap)
    return vox[:c], ln[:c]


def siddon_walk_csr(op, dtype=np.float64):
    """CSR of a Siddon geometry by the tracked-plane traversal in dtype
    (positions, origin and voxel size rounded to dtype first)"""
    P = op.s.num_paths
    f = getattr(lib(), "orc_siddon_walk_csr" + _sfx(dtype))
    rowptr = np.zeros(P + 1, np.int32)
    nnz = f(C.byref(op.s), _ptr(rowptr), None, None)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), dtype)
    f(C.byref(op.s), _ptr(rowptr), _ptr(col), _ptr(val))
    return rowptr, col[:nnz], val[:nnz]


def csr_transpose(m, n, rowptr, col, val):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    nnz = int(rowptr[m])
    cp = np.zeros(n + 1, np.int32)
    rw = np.zeros(max(nnz, 1), np.int32)
    cv = np.zeros(max(nnz, 1), np.float64)
    lib().orc_csr_transpose(m, n, _ptr(rowptr), _ptr(col), _ptr(val), _ptr(cp), _ptr(rw),
                            _ptr(cv))
    return cp, rw[:nnz], cv[:nnz]


def matvec(op, trans, x, dtype=np.float64):
    x = np.ascontiguousarray(x, dtype)
    y = np.zeros(op.s.n if trans else op.s.m, dtype)
    getattr(lib(), "orc_matvec" + _sfx(dtype))(C.byref(op.s), int(trans), _ptr(x), _ptr(y))
    return y


def D(U, dims, dtype=np.float64):
    p, q, r = dims
    U = np.ascontiguousarray(U, dtype)
    Ux, Uy, Uz = (np.zeros_like(U) for _ in range(3))
    getattr(lib(), "orc_D" + _sfx(dtype))(p, q, r, _ptr(U), _ptr(Ux), _ptr(Uy), _ptr(Uz))
    return Ux, Uy, Uz


def Dt(X, Y, Z, dims, dtype=np.float64):
    p, q, r = dims
    X, Y, Z = (np.ascontiguousarray(a, dtype) for a in (X, Y, Z))
    res = np.zeros_like(X)
    getattr(lib(), "orc_Dt" + _sfx(dtype))(p, q, r, _ptr(X), _ptr(Y), _ptr(Z), _ptr(res))
    return res


def shrink(Ux, Uy, Uz, sx, sy, sz, beta, iso=False, dtype=np.float64):
    arrs = [np.ascontiguousarray(a, dtype) for a in (Ux, Uy, Uz, sx, sy, sz)]
    W = [np.zeros_like(arrs[0]) for _ in range(3)]
    lam1 = getattr(lib(), "orc_shrink" + _sfx(dtype))(len(arrs[0]), int(iso),
                                                      *[_ptr(a) for a in arrs], float(beta),
                                                      *[_ptr(w) for w in W])
    return W[0], W[1], W[2], lam1


def scaleb(b, dtype=np.float64):
    b = np.array(b, dtype=dtype)
    scl = getattr(lib(), "orc_scaleb" + _sfx(dtype))(len(b), _ptr(b))
    return b, scl


def solve(op, dims, b, opts=None, Ut=None, dtype=np.float64, hist_cap=4096, state=False):
    """Run the oracle TVAL3 loop.  Returns (U, info, history dict, status).
    The history holds the per-iteration record (including the line search's
    f_trial, f_acc, C_in, c_armij, dd, gam); with state=True also the final
    sigma (3 x n), delta (m), DtsAtd (n) and d (n) of the scaled problem."""
    p, q, r = dims
    o = opts if opts is not None else default_opts()
    b = np.ascontiguousarray(b, dtype)
    U = np.zeros(p * q * r, dtype)
    ut = None if Ut is None else np.ascontiguousarray(Ut, dtype)
    info = OrcInfo()
    h = OrcHistory()
    h.cap = hist_cap
    h.count = 0
    harr = {k: np.zeros(hist_cap, np.float64) for k in _HIST_D + _HIST_LS}
    harr["cnt"] = np.zeros(hist_cap, np.int32)
    harr["outer"] = np.zeros(hist_cap, np.int32)
    harr["mulT"] = np.zeros(hist_cap, np.int32)
    for k in _HIST_D + _HIST_LS:
        setattr(h, k, harr[k].ctypes.data_as(C.POINTER(C.c_double)))
    sarr = {}
    if state:
        n, m = p * q * r, int(op.s.m)
        sarr = {"sigma_out": np.zeros(3 * n), "delta_out": np.zeros(m),
                "dts_out": np.zeros(n), "d_out": np.zeros(n)}
        for k, v in sarr.items():
            setattr(h, k, v.ctypes.data_as(C.POINTER(C.c_double)))
    h.cnt = harr["cnt"].ctypes.data_as(C.POINTER(C.c_int))
    h.outer = harr["outer"].ctypes.data_as(C.POINTER(C.c_int))
    h.mulT = harr["mulT"].ctypes.data_as(C.POINTER(C.c_int))
    st = getattr(lib(), "orc_solve" + _sfx(dtype))(C.byref(op.s), C.byref(o), p, q, r, _ptr(b),
                                                   _ptr(U), _ptr(ut), C.byref(info),
                                                   C.byref(h))
    hist = {k: v[:h.count].copy() for k, v in harr.items()}
    if state:
        n = p * q * r
        hist["sigma"] = sarr["sigma_out"].reshape(3, n)
        hist["delta"] = sarr["delta_out"]
        hist["DtsAtd"] = sarr["dts_out"]
        hist["d"] = sarr["d_out"]
    return U, info, hist, st
