"""Python mirror of the reference's solver interface on top of the C ABI.

The reference exposes tval3_gpu_3d(U, A, b, opts, Ut, ...) overloads for a
sparse (CSR + CSC), dense and geometry ("dynamic") system matrix
(3d/tval3_gpu_3d.cpp:1673-1745) and a mat_vec_mul(trans, A, x, y, stream)
operator overload set (3d/mat_vec_mul_sparse.h:1703, 3d/mat_vec_mul_dense.h:25,
3d/mat_vec_mul_dynamic.h:682).  Here:

    ctx = Context(0)
    A = SparseOperator(ctx, m, n, rowptr, col, val)          # or DenseOperator,
    U, info, hist = tval3_gpu_3d(ctx, A, b, dims, opts)      #    DynamicOperator
    A.mat_vec_mul(x, y, trans=False)

Device memory comes from torch tensors on cuda:0 (plumbing only); every
computation runs in the HIP library.
"""
import ctypes as C

import numpy as np

from . import _lib as L


def _vp(t):
    """Device or host pointer of a torch tensor / numpy array."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data_as(C.c_void_p)
    return C.c_void_p(t.data_ptr())


def default_opts(**kw):
    o = L.Opts()
    L.load().tv3_opts_default(C.byref(o))
    for k, v in kw.items():
        if k == "solver" and isinstance(v, str):
            v = {"plain": L.PLAIN, "fused": L.FUSED}[v]
        if k == "tv_mode" and isinstance(v, str):
            v = {"aniso": L.ANISO, "iso": L.ISO}[v]
        setattr(o, k, v)
    return o


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        L.check(L.load().tv3_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    @property
    def stream(self):
        return L.load().tv3_ctx_stream(self.h)

    def close(self):
        if self.h:
            L.load().tv3_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st):
        L.check(st, self.h)

    def bandwidth_probe(self, nbytes=1 << 30, reps=10):
        """(copy GB/s, read GB/s) of streaming kernels over nbytes on this
        context's device (tv3_bandwidth_probe)"""
        cp, rd = C.c_double(), C.c_double()
        self.check(L.load().tv3_bandwidth_probe(self.h, int(nbytes), int(reps), C.byref(cp),
                                                C.byref(rd)))
        return cp.value, rd.value

    # building blocks -------------------------------------------------------
    def D(self, U, dims):
        import torch
        Ux, Uy, Uz = (torch.empty_like(U) for _ in range(3))
        d = (C.c_int32 * 3)(*dims)
        self.check(L.load().tv3_D(self.h, d, _vp(U), _vp(Ux), _vp(Uy), _vp(Uz),
                                  _torch_stream()))
        return Ux, Uy, Uz

    def Dt(self, X, Y, Z, dims):
        import torch
        r = torch.empty_like(X)
        d = (C.c_int32 * 3)(*dims)
        self.check(L.load().tv3_Dt(self.h, d, _vp(X), _vp(Y), _vp(Z), _vp(r), _torch_stream()))
        return r

    def shrink(self, Ux, Uy, Uz, sx, sy, sz, beta, iso=False):
        import torch
        torch.cuda.synchronize()
        W = [torch.empty_like(Ux) for _ in range(3)]
        lam1 = C.c_double()
        self.check(L.load().tv3_shrink(self.h, Ux.numel(), L.ISO if iso else L.ANISO, _vp(Ux),
                                       _vp(Uy), _vp(Uz), _vp(sx), _vp(sy), _vp(sz),
                                       float(beta), *[_vp(w) for w in W], C.byref(lam1)))
        return W[0], W[1], W[2], lam1.value


def _torch_stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Operator:
    kind = "abstract"

    def __init__(self, ctx, handle, keep=None):
        self.ctx = ctx
        self.h = handle
        self._keep = keep
        m, n, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        ctx.check(L.load().tv3_op_info(self.h, C.byref(m), C.byref(n), C.byref(nnz)))
        self.m, self.n, self.nnz = m.value, n.value, nnz.value

    def close(self):
        if self.h:
            L.load().tv3_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mat_vec_mul(self, x, y=None, trans=False):
        """y = A x (trans False) or A^T x on torch's current stream; y is
        overwritten.  float32 or float64 tensors on the context's device."""
        import torch
        out_len = self.n if trans else self.m
        if y is None:
            y = torch.empty(out_len, dtype=x.dtype, device=x.device)
        assert x.is_contiguous() and y.is_contiguous()
        assert x.numel() == (self.m if trans else self.n) and y.numel() == out_len
        f = L.load().tv3_matvec_f64 if x.dtype == torch.float64 else L.load().tv3_matvec
        self.ctx.check(f(self.h, 1 if trans else 0, _vp(x), _vp(y), _torch_stream()))
        return y

    def mat_vec_mul_noalloc(self, x, y, trans=False):
        """mat_vec_mul through tv3_matvec_noalloc: no device allocation (a
        dense A x keeps its partial sums in the library's fixed per-GPU
        scratch array); float32 tensors, y preallocated."""
        import torch
        assert x.dtype == torch.float32 and y.dtype == torch.float32
        assert x.is_contiguous() and y.is_contiguous()
        assert x.numel() == (self.m if trans else self.n)
        assert y.numel() == (self.n if trans else self.m)
        self.ctx.check(L.load().tv3_matvec_noalloc(self.h, 1 if trans else 0, _vp(x), _vp(y),
                                                   _torch_stream()))
        return y

    def mat_mat_mul(self, X, Y=None, trans=False):
        """Y[:, j] = A X[:, j] (or A^T X[:, j]) for the columns of X, a float32
        (length, nrhs) tensor with columns contiguous (X.t() contiguous, e.g.
        torch.empty(nrhs, length).t()); on torch's current stream."""
        import torch
        xl, yl = (self.m, self.n) if trans else (self.n, self.m)
        assert X.dtype == torch.float32 and X.dim() == 2 and X.shape[0] == xl
        assert X.stride(0) == 1, "columns of X must be contiguous"
        nrhs = X.shape[1]
        if Y is None:
            Y = torch.empty(nrhs, yl, dtype=torch.float32, device=X.device).t()
        assert Y.shape == (yl, nrhs) and Y.stride(0) == 1
        self.ctx.check(L.load().tv3_matvec_multi(self.h, 1 if trans else 0, nrhs, _vp(X),
                                                 X.stride(1) if nrhs > 1 else xl, _vp(Y),
                                                 Y.stride(1) if nrhs > 1 else yl,
                                                 _torch_stream()))
        return Y

    def materialize(self, dtype="f32"):
        h = C.c_void_p()
        self.ctx.check(L.load().tv3_op_materialize(self.h, L.DTYPES[dtype], C.byref(h)))
        op = Operator.__new__(SparseOperator)
        Operator.__init__(op, self.ctx, h)
        op.dtype = dtype
        return op

    def prepare(self, dims):
        """Build the fused solver's layout for a volume shape (renumbering,
        CSC panels, tuned launch shapes); returns its description."""
        d = np.ascontiguousarray(dims, np.int32)
        info = L.LayoutInfo()
        self.ctx.check(L.load().tv3_op_prepare(self.h, _vp(d), C.byref(info)))
        return {k: getattr(info, k) for k, _ in L.LayoutInfo._fields_ if k != "reserved"}

    def layout_force(self, dims, code_N=0, code_T=0, pan_nnz=16384, panels=0):
        """Override the layout's launch shapes.  panels: bit 0 = A x, bit 1 =
        A^T r in panel form (block sizes -code_N / -code_T, at most pan_nnz
        entries per panel; the other bits as tv3_op_layout_force documents
        them: panel sizes, and bits 8 / 9 packed entries of A / A^T); other
        directions take set_lanes codes."""
        d = np.ascontiguousarray(dims, np.int32)
        info = L.LayoutInfo()
        self.ctx.check(L.load().tv3_op_layout_force(self.h, _vp(d), code_N, code_T, pan_nnz,
                                                    panels, C.byref(info)))
        return {k: getattr(info, k) for k, _ in L.LayoutInfo._fields_ if k != "reserved"}

    def layout_mat_vec_mul(self, dims, x, y=None, trans=False):
        """A x or A^T x by the fused solver's kernels in its layout (natural
        order in and out) on torch's current stream."""
        import torch
        out_len = self.n if trans else self.m
        if y is None:
            y = torch.empty(out_len, dtype=torch.float32, device=x.device)
        assert x.dtype == torch.float32 and x.is_contiguous() and y.is_contiguous()
        assert x.numel() == (self.m if trans else self.n) and y.numel() == out_len
        d = np.ascontiguousarray(dims, np.int32)
        self.ctx.check(L.load().tv3_op_layout_matvec(self.h, _vp(d), 1 if trans else 0, _vp(x),
                                                     _vp(y), _torch_stream()))
        return y

    def layout_time(self, dims, trans=False, reps=20):
        """Average duration (s) of one launch of the layout's product kernel,
        HIP events on the context stream."""
        d = np.ascontiguousarray(dims, np.int32)
        us = C.c_double()
        self.ctx.check(L.load().tv3_op_layout_time(self.h, _vp(d), 1 if trans else 0, reps,
                                                   C.byref(us)))
        return us.value * 1e-6

    def autotune(self):
        a, b = C.c_int32(), C.c_int32()
        self.ctx.check(L.load().tv3_op_autotune(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_lanes(self, lanes_N=0, lanes_T=0):
        self.ctx.check(L.load().tv3_op_set_lanes(self.h, lanes_N, lanes_T))

    def download_csr(self):
        rowptr = np.zeros(self.m + 1, np.int32)
        col = np.zeros(max(self.nnz, 1), np.int32)
        val = np.zeros(max(self.nnz, 1), np.float64)
        self.ctx.check(L.load().tv3_op_csr_download(self.h, _vp(rowptr), _vp(col), _vp(val)))
        return rowptr, col[:self.nnz], val[:self.nnz]


def _to_dtype_bytes(val, dtype):
    if dtype == "f32":
        return np.ascontiguousarray(val, np.float32)
    if dtype == "f64":
        return np.ascontiguousarray(val, np.float64)
    if dtype == "f16":
        return np.ascontiguousarray(val, np.float16)
    if dtype == "bf16":
        v = np.ascontiguousarray(val, np.float32).view(np.uint32)
        # round to nearest even
        r = ((v >> 16) & 1) + 0x7FFF
        return ((v + r) >> 16).astype(np.uint16)
    raise ValueError(dtype)


class SparseOperator(Operator):
    """sparse_mm: CSR for A x and its CSC (given or built on the device) for
    A^T x (3d/mat_vec_mul_sparse.h:1608-1780)."""
    kind = "csr"

    def __init__(self, ctx, m, n, rowptr, col, val, dtype="f32", csc=None):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        vb = _to_dtype_bytes(val, dtype)
        cp = cr = cv = None
        if csc is not None:
            cp = np.ascontiguousarray(csc[0], np.int32)
            cr = np.ascontiguousarray(csc[1], np.int32)
            cv = _to_dtype_bytes(csc[2], dtype)
        h = C.c_void_p()
        ctx.check(L.load().tv3_op_csr_create(ctx.h, m, n, int(rowptr[-1]), _vp(rowptr), _vp(col),
                                             _vp(vb), L.DTYPES[dtype], _vp(cp), _vp(cr),
                                             _vp(cv), C.byref(h)))
        super().__init__(ctx, h)
        self.dtype = dtype


class DenseOperator(Operator):
    """dense_mm: column-major m x n (3d/mat_vec_mul_dense.h:13-73)."""
    kind = "dense"

    def __init__(self, ctx, A, dtype="f32"):
        A = np.asfortranarray(A)
        m, n = A.shape
        vb = np.asfortranarray(_to_dtype_bytes(A.reshape(-1, order="F"), dtype))
        h = C.c_void_p()
        ctx.check(L.load().tv3_op_dense_create(ctx.h, m, n, _vp(vb), m, L.DTYPES[dtype],
                                               C.byref(h)))
        super().__init__(ctx, h)
        self.dtype = dtype


def make_geometry(geom):
    """Build a tv3_geometry from the dict of tval3_amd.workload.transducer_geometry.
    Returns (struct, keepalive)."""
    g = L.Geometry()
    E, R, mps = geom["E"], geom["R"], geom.get("mps", 1)
    g.num_emitters, g.num_receivers, g.num_mps = E, R, mps
    g.rv_x, g.rv_y, g.rv_z = geom["rv"]
    g.scale_factor = float(geom.get("scale", 1.0))
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return _vp(a)

    ev = np.asarray(geom["e_vox"])
    rv = np.asarray(geom["r_vox"])
    g.x_emitters, g.y_emitters, g.z_emitters = (arr(ev[:, k], np.int32) for k in range(3))
    g.x_receivers, g.y_receivers, g.z_receivers = (arr(rv[:, k], np.int32) for k in range(3))
    g.pos_emitters = arr(geom["e_pos"], np.float32)
    g.pos_receivers = arr(geom["r_pos"], np.float32)
    g.origin[:] = [float(v) for v in geom["origin"]]
    g.voxel[:] = [float(v) for v in geom["voxel"]]
    up = geom.get("use_path")
    g.use_path = None if up is None else arr(up, np.uint8)
    return g, keep


class DynamicOperator(Operator):
    """geom_mm: ray weights generated inside every product
    (3d/mat_vec_mul_dynamic.h:330-967).  mode 'siddon' or 'bresenham'."""
    kind = "dynamic"

    def __init__(self, ctx, geom, mode="siddon"):
        g, keep = make_geometry(geom)
        h = C.c_void_p()
        md = L.SIDDON if mode == "siddon" else L.BRESENHAM
        ctx.check(L.load().tv3_op_dynamic_create(ctx.h, C.byref(g), md, C.byref(h)))
        super().__init__(ctx, h, keep)
        self.mode = mode


class HybridOperator(Operator):
    """sparse_geom_mm(hA_geo, hA_mat): A x by the traversal of the geometry,
    A^T x by the caller's matrix (3d/mat_vec_mul.h:104-120, products
    :732-756).  csr = (rowptr, col, val) and/or csc = (colptr, row, val) of
    A (num_paths x voxels); the CSC is used as given, a CSR alone is
    transposed on the device; with neither, the geometry's own matrix is
    built on the device."""
    kind = "hybrid"

    def __init__(self, ctx, geom, mode="siddon", dtype="f32", csr=None, csc=None):
        g, keep = make_geometry(geom)
        h = C.c_void_p()
        md = L.SIDDON if mode == "siddon" else L.BRESENHAM
        rp = ci = cvv = cp = cr = cv = None
        nnz = 0
        if csr is not None:
            rp = np.ascontiguousarray(csr[0], np.int32)
            ci = np.ascontiguousarray(csr[1], np.int32)
            cvv = _to_dtype_bytes(csr[2], dtype)
            nnz = int(rp[-1])
        if csc is not None:
            cp = np.ascontiguousarray(csc[0], np.int32)
            cr = np.ascontiguousarray(csc[1], np.int32)
            cv = _to_dtype_bytes(csc[2], dtype)
            nnz = int(cp[-1])
        keep += [a for a in (rp, ci, cvv, cp, cr, cv) if a is not None]
        ctx.check(L.load().tv3_op_hybrid_create(ctx.h, C.byref(g), md, nnz, _vp(rp), _vp(ci),
                                                _vp(cvv), L.DTYPES[dtype], _vp(cp), _vp(cr),
                                                _vp(cv), C.byref(h)))
        super().__init__(ctx, h, keep)
        self.mode = mode


def _history(cap):
    h = L.History()
    h.cap = cap
    h.count = 0
    arrs = {k: np.zeros(cap, np.float64) for k in L.HIST_D}
    arrs["cnt"] = np.zeros(cap, np.int32)
    arrs["outer"] = np.zeros(cap, np.int32)
    arrs["mulT"] = np.zeros(cap, np.int32)
    for k in L.HIST_D:
        setattr(h, k, arrs[k].ctypes.data_as(C.POINTER(C.c_double)))
    h.cnt = arrs["cnt"].ctypes.data_as(C.POINTER(C.c_int32))
    h.outer = arrs["outer"].ctypes.data_as(C.POINTER(C.c_int32))
    h.mulT = arrs["mulT"].ctypes.data_as(C.POINTER(C.c_int32))
    arrs["f_trial"] = np.zeros(cap, np.float64)
    h.f_trial = arrs["f_trial"].ctypes.data_as(C.POINTER(C.c_double))
    return h, arrs


def tval3_gpu_3d(ctx, A, b, dims, opts=None, Ut=None, U0=None, history=True, hist_cap=8192):
    """Solve min TV(U) s.t. A U ~ b (tval3_gpu_3d, 3d/tval3_gpu_3d.cpp:1088-1345).

    b, Ut, U0: numpy arrays (host path, like the reference) or torch tensors
    on the device (device path, no host transfers).  Returns (U, info dict,
    history dict or None)."""
    o = opts if opts is not None else default_opts()
    d = (C.c_int32 * 3)(*dims)
    info = L.Info()
    hp, harr = (_history(hist_cap) if history else (None, None))
    hptr = C.byref(hp) if hp is not None else None
    lib = L.load()
    n = int(np.prod(dims))
    if len(b) != A.m:
        raise ValueError(f"b has {len(b)} entries, A has {A.m} rows")
    if Ut is not None and len(Ut) != n:
        raise ValueError("Ut must have prod(dims) entries")
    if isinstance(b, np.ndarray):
        dt = b.dtype if b.dtype in (np.float32, np.float64) else np.float32
        bb = np.ascontiguousarray(b, dt)
        U = np.zeros(n, dt) if U0 is None else np.array(U0, dt)
        if U0 is not None:
            o.warm_start = 1
        ut = None if Ut is None else np.ascontiguousarray(Ut, dt)
        f = lib.tv3_solve_f64 if dt == np.float64 else lib.tv3_solve
        st = f(ctx.h, A.h, C.byref(o), d, _vp(bb), _vp(U), _vp(ut), C.byref(info), hptr)
    else:
        import torch
        assert b.dtype == torch.float32, "device solve takes float32 tensors"
        U = torch.zeros(n, dtype=torch.float32, device=b.device) if U0 is None else U0.clone()
        if U0 is not None:
            o.warm_start = 1
        torch.cuda.synchronize()
        st = lib.tv3_solve_device(ctx.h, A.h, C.byref(o), d, _vp(b), _vp(U), _vp(Ut),
                                  C.byref(info), hptr)
    ctx.check(st)
    hist = None
    if hp is not None:
        hist = {k: v[:hp.count].copy() for k, v in harr.items()}
    return U, info.as_dict(), hist


def tval3_gpu_3d_batched(ctx, ops, bs, dims, opts=None, Uts=None, history=False,
                         hist_cap=4096, lanes=None, splits=None, panels=None, panels_A=None,
                         panel_box_A=None):
    """Solve many problems together (tv3_solve_batched): ops a list of sparse
    operators (one weight type), bs their b (float32 torch tensors on the
    device), dims a list of (dim_y, dim_x, dim_z).  Every kernel of an
    iteration covers all problems; each problem keeps its own scalars and
    stopping.  Problems sharing one dense BF16 operator run its products as
    one matrix-core product.  lanes = (lanes_N, lanes_T) fixes the sparse
    products' launch shapes, splits = (splits_N, splits_T) the dense
    products' K splits (default: timed); panels = 1024, 2048 or 4096 puts A^T
    of the sparse problems in panel form with row blocks of that size, -1 never
    (default: timed where possible); panels_A likewise for the two-pass
    panel form of A (1024 or 2048; 4096 too with boxes of 8 tiles), panel_box_A = 1 or 8
    its tiles per panel (8: 16^3 voxels; default: timed with panels_A unset,
    else 1).  Returns (list of U tensors, list of
    info dicts, batch info dict, list of history dicts or None)."""
    import torch
    o = opts if opts is not None else default_opts()
    P = len(ops)
    if not (len(bs) == len(dims) == P) or P == 0:
        raise ValueError("ops, bs and dims must be non-empty lists of equal length")
    for k, (op, b) in enumerate(zip(ops, bs)):
        if b.dtype != torch.float32 or not b.is_contiguous() or b.numel() != op.m:
            raise ValueError(f"b[{k}] must be a contiguous float32 tensor of {op.m} entries")
    n = [int(np.prod(dk)) for dk in dims]
    if o.warm_start:
        raise ValueError("use the U0 of tval3_gpu_3d for warm starts; batched solves start "
                         "from A^T b")
    Us = [torch.zeros(nk, dtype=torch.float32, device=bs[0].device) for nk in n]
    Ph = C.c_void_p * P
    opsa = Ph(*[op.h for op in ops])
    d = (C.c_int32 * (3 * P))(*[int(v) for dk in dims for v in dk])
    ba = Ph(*[b.data_ptr() for b in bs])
    ua = Ph(*[u.data_ptr() for u in Us])
    uta = None
    if Uts is not None:
        uta = Ph(*[(ut.data_ptr() if ut is not None else None) for ut in Uts])
    infos = (L.Info * P)()
    hp = None
    keep = []
    if history:
        hp = (L.History * P)()
        for k in range(P):
            h, arrs = _history(hist_cap)
            hp[k] = h
            keep.append(arrs)
    bi = L.BatchInfo()
    if lanes is not None:
        bi.lanes_N, bi.lanes_T = lanes
    if splits is not None:
        bi.splits_N, bi.splits_T = splits
    if panels is not None:
        bi.panels_T = panels
    if panels_A is not None:
        bi.panels_A = panels_A
    if panel_box_A is not None:
        bi.panel_box_A = panel_box_A
    torch.cuda.synchronize()
    ctx.check(L.load().tv3_solve_batched(ctx.h, P, opsa, d, C.byref(o), ba, ua, uta, infos,
                                         hp, C.byref(bi)))
    hists = None
    if history:
        hists = [{k: v[:hp[j].count].copy() for k, v in keep[j].items()} for j in range(P)]
    return Us, [infos[k].as_dict() for k in range(P)], bi.as_dict(), hists
