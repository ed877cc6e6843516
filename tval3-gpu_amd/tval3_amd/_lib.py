"""ctypes declarations of the C ABI in include/tval3_mi355x.h.

Loads tval3-gpu_amd/lib/libtval3_mi355x.so and fails loudly when it is
missing: there is no CPU or PyTorch fallback for any operation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libtval3_mi355x.so")

OK, E_ARG, E_SHAPE, E_FORMAT, E_NOMEM, E_HIP, E_NONFINITE, E_INTERNAL, E_UNSUPPORTED = range(9)
STATUS_NAMES = {0: "TV3_OK", 1: "TV3_E_ARG", 2: "TV3_E_SHAPE", 3: "TV3_E_FORMAT",
                4: "TV3_E_NOMEM", 5: "TV3_E_HIP", 6: "TV3_E_NONFINITE", 7: "TV3_E_INTERNAL",
                8: "TV3_E_UNSUPPORTED"}
F32, F16, BF16, F64 = 0, 1, 2, 3
DTYPES = {"f32": F32, "f16": F16, "bf16": BF16, "f64": F64}
BRESENHAM, SIDDON = 0, 1
ANISO, ISO = 0, 1
PLAIN, FUSED = 0, 1

_P = C.c_void_p


class TV3Error(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Opts(C.Structure):
    _fields_ = [("mu", C.c_double), ("mu0", C.c_double), ("beta", C.c_double),
                ("beta0", C.c_double), ("rate_cnt", C.c_double), ("tol", C.c_double),
                ("tol_inn", C.c_double), ("maxcnt", C.c_int32), ("maxit", C.c_int32),
                ("c", C.c_double), ("gamma", C.c_double), ("gam", C.c_double),
                ("rate_gam", C.c_double), ("isreal", C.c_int32), ("nonneg", C.c_int32),
                ("numInitIterations", C.c_uint32), ("skipMulRatio", C.c_double),
                ("maxRelativeChange", C.c_double), ("tv_mode", C.c_int32),
                ("solver", C.c_int32), ("warm_start", C.c_int32), ("graph_iters", C.c_int32),
                ("verbose", C.c_int32), ("tiles_per_wg", C.c_int32)]


class Info(C.Structure):
    _fields_ = [("total_iters", C.c_int32), ("outer_iters", C.c_int32),
                ("rel_chg", C.c_double), ("rel_error", C.c_double), ("secs", C.c_double),
                ("outer_secs", C.c_double), ("loop_secs", C.c_double),
                ("backtracks", C.c_int32), ("fallbacks", C.c_int32),
                ("n_matvec_N", C.c_int64), ("n_matvec_T", C.c_int64), ("status", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


HIST_D = ["f", "lam1", "lam2", "lam3", "lam4", "lam5", "tau", "alpha", "relchg", "mu", "beta",
          "C"]


class History(C.Structure):
    _fields_ = [("cap", C.c_int32), ("count", C.c_int32)] + \
               [(k, C.POINTER(C.c_double)) for k in HIST_D] + \
               [("cnt", C.POINTER(C.c_int32)), ("outer", C.POINTER(C.c_int32)),
                ("mulT", C.POINTER(C.c_int32)), ("f_trial", C.POINTER(C.c_double))]


class Geometry(C.Structure):
    _fields_ = [("num_emitters", C.c_int32), ("num_receivers", C.c_int32),
                ("num_mps", C.c_int32), ("rv_x", C.c_int32), ("rv_y", C.c_int32),
                ("rv_z", C.c_int32), ("scale_factor", C.c_float),
                ("x_emitters", _P), ("y_emitters", _P), ("z_emitters", _P),
                ("x_receivers", _P), ("y_receivers", _P), ("z_receivers", _P),
                ("pos_emitters", _P), ("pos_receivers", _P),
                ("origin", C.c_float * 3), ("voxel", C.c_float * 3), ("use_path", _P)]


class LayoutInfo(C.Structure):
    _fields_ = [("brick", C.c_int32), ("panels", C.c_int32), ("code_N", C.c_int32),
                ("code_T", C.c_int32), ("npan", C.c_int64), ("nseg", C.c_int64),
                ("max_seg", C.c_int32), ("reserved", C.c_int32)]


class BatchInfo(C.Structure):
    _fields_ = [("lanes_N", C.c_int32), ("lanes_T", C.c_int32), ("graph_iters", C.c_int32),
                ("graph_launches", C.c_int32), ("secs", C.c_double), ("loop_secs", C.c_double),
                ("host_secs", C.c_double), ("total_iters", C.c_int64), ("nnz", C.c_int64),
                ("voxels", C.c_int64), ("rays", C.c_int64), ("dense_groups", C.c_int32),
                ("splits_N", C.c_int32), ("splits_T", C.c_int32), ("panels_T", C.c_int32),
                ("dense_values", C.c_int64), ("panels_A", C.c_int32),
                ("panel_box_A", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) of every exported entry point of the header
SIGNATURES = [
    ("tv3_abi_version", C.c_int32, []),
    ("tv3_opts_default", None, [C.POINTER(Opts)]),
    ("tv3_ctx_create", C.c_int, [C.c_int32, C.POINTER(_P)]),
    ("tv3_ctx_destroy", None, [_P]),
    ("tv3_last_error", C.c_char_p, [_P]),
    ("tv3_ctx_stream", _P, [_P]),
    ("tv3_op_csr_create", C.c_int, [_P, C.c_int64, C.c_int64, C.c_int64, _P, _P, _P, C.c_int,
                                    _P, _P, _P, C.POINTER(_P)]),
    ("tv3_op_dense_create", C.c_int, [_P, C.c_int64, C.c_int64, _P, C.c_int64, C.c_int,
                                      C.POINTER(_P)]),
    ("tv3_op_dynamic_create", C.c_int, [_P, C.POINTER(Geometry), C.c_int, C.POINTER(_P)]),
    ("tv3_op_hybrid_create", C.c_int, [_P, C.POINTER(Geometry), C.c_int, C.c_int64, _P, _P, _P,
                                       C.c_int, _P, _P, _P, C.POINTER(_P)]),
    ("tv3_op_materialize", C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    ("tv3_op_destroy", None, [_P]),
    ("tv3_op_info", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                              C.POINTER(C.c_int64)]),
    ("tv3_op_csr_download", C.c_int, [_P, _P, _P, _P]),
    ("tv3_matvec", C.c_int, [_P, C.c_int, _P, _P, _P]),
    ("tv3_matvec_f64", C.c_int, [_P, C.c_int, _P, _P, _P]),
    ("tv3_matvec_multi", C.c_int, [_P, C.c_int, C.c_int32, _P, C.c_int64, _P, C.c_int64, _P]),
    ("tv3_solve", C.c_int, [_P, _P, C.POINTER(Opts), _P, _P, _P, _P, C.POINTER(Info),
                            C.POINTER(History)]),
    ("tv3_solve_device", C.c_int, [_P, _P, C.POINTER(Opts), _P, _P, _P, _P, C.POINTER(Info),
                                   C.POINTER(History)]),
    ("tv3_solve_f64", C.c_int, [_P, _P, C.POINTER(Opts), _P, _P, _P, _P, C.POINTER(Info),
                                C.POINTER(History)]),
    ("tv3_solve_batched", C.c_int, [_P, C.c_int32, _P, _P, C.POINTER(Opts), _P, _P, _P, _P, _P,
                                    C.POINTER(BatchInfo)]),
    ("tv3_D", C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    ("tv3_Dt", C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    ("tv3_shrink", C.c_int, [_P, C.c_int64, C.c_int, _P, _P, _P, _P, _P, _P, C.c_double, _P,
                             _P, _P, C.POINTER(C.c_double)]),
    ("tv3_D_f64", C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    ("tv3_Dt_f64", C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    ("tv3_shrink_apply", C.c_int, [_P, C.c_int64, C.c_int, C.c_int, _P, _P, _P, C.c_double,
                                   _P, _P, _P, _P, _P, _P, _P]),
    ("tv3_shrink_adjoint", C.c_int, [_P, C.c_int64, C.c_int, C.c_int, _P, _P, _P, C.c_double,
                                     _P, _P, _P, _P, _P, _P, _P, _P]),
    ("tv3_op_prepare", C.c_int, [_P, _P, C.POINTER(LayoutInfo)]),
    ("tv3_op_layout_force", C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.POINTER(LayoutInfo)]),
    ("tv3_op_layout_matvec", C.c_int, [_P, _P, C.c_int, _P, _P, _P]),
    ("tv3_op_layout_time", C.c_int, [_P, _P, C.c_int, C.c_int32, C.POINTER(C.c_double)]),
    ("tv3_bandwidth_probe", C.c_int, [_P, C.c_int64, C.c_int32, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]),
    ("tv3_matvec_noalloc", C.c_int, [_P, C.c_int, _P, _P, _P]),
    ("tv3_op_autotune", C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("tv3_op_set_lanes", C.c_int, [_P, C.c_int32, C.c_int32]),
]

_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} is missing "
                               "(run __graft_entry__.build() or make -C tval3-gpu_amd)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, ctx=None):
    if status != OK:
        msg = load().tv3_last_error(ctx)
        raise TV3Error(status, msg.decode() if msg else "")
