"""CPU-side checks of the C ABI library: it loads, exports every entry point
the header declares, the ctypes mirrors match the header, and the option
defaults equal the reference's (2d/tval3_types.h:29-34).  No compute calls."""
import ctypes as C
import os
import re
import subprocess

import pytest

from tval3_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tval3_mi355x.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tv3_[A-Za-z0-9_]+)\s*\(", txt)) - {"tv3_status"})


def test_library_exports_every_header_symbol():
    lib = L.load()
    names = header_functions()
    assert len(names) >= 20
    out = subprocess.check_output(["nm", "-D", "--defined-only", L.LIB_PATH], text=True)
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)
    declared = {s[0] for s in L.SIGNATURES}
    assert set(names) == declared, set(names) ^ declared


def test_abi_version_and_defaults(orc):
    lib = L.load()
    assert lib.tv3_abi_version() == 1
    o = L.Opts()
    lib.tv3_opts_default(C.byref(o))
    ro = orc.default_opts()
    for k in ("mu", "mu0", "beta", "beta0", "rate_cnt", "tol", "tol_inn", "maxcnt", "maxit",
              "c", "gamma", "gam", "rate_gam", "nonneg"):
        assert getattr(o, k) == getattr(ro, k), k
    assert (o.mu, o.beta, o.maxit, o.maxcnt) == (256.0, 32.0, 1025, 10)
    assert o.skipMulRatio == 0.0 and o.solver == L.PLAIN and o.tv_mode == L.ANISO


def test_struct_layouts_match_header():
    # sizes computed from the C declarations by the host compiler
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "tval3_mi355x.h"
int main(void){
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(tv3_opts),
         sizeof(tv3_info), sizeof(tv3_history), sizeof(tv3_geometry),
         offsetof(tv3_opts, solver), offsetof(tv3_geometry, origin), sizeof(tv3_layout_info),
         offsetof(tv3_layout_info, max_seg), sizeof(tv3_batch_info),
         offsetof(tv3_batch_info, panels_T), offsetof(tv3_batch_info, panels_A),
         offsetof(tv3_opts, tiles_per_wg));
  return 0;
}'''
    tmp = os.path.join(ROOT, "tests", "_layout_probe")
    with open(tmp + ".c", "w") as f:
        f.write(src)
    try:
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), tmp + ".c", "-o", tmp])
        got = [int(v) for v in subprocess.check_output([tmp], text=True).split()]
    finally:
        for p in (tmp, tmp + ".c"):
            if os.path.exists(p):
                os.remove(p)
    exp = [C.sizeof(L.Opts), C.sizeof(L.Info), C.sizeof(L.History), C.sizeof(L.Geometry),
           L.Opts.solver.offset, L.Geometry.origin.offset, C.sizeof(L.LayoutInfo),
           L.LayoutInfo.max_seg.offset, C.sizeof(L.BatchInfo), L.BatchInfo.panels_T.offset,
           L.BatchInfo.panels_A.offset, L.Opts.tiles_per_wg.offset]
    assert got == exp


def test_bf16_rounding_helper():
    import numpy as np
    from tval3_amd.api import _to_dtype_bytes
    v = np.array([1.0, 1.00390625, -2.5, 3.0e-3], np.float32)
    b = _to_dtype_bytes(v, "bf16")
    back = (b.astype(np.uint32) << 16).view(np.float32)
    assert back[0] == 1.0 and back[2] == -2.5
    assert np.allclose(back, v, rtol=2 ** -8)


def test_cpp_dropin_shim_compiles_and_links():
    """include/tval3_gpu_3d.hpp: the reference-shaped C++ overload set over the
    C ABI compiles (every overload instantiated, float and double) and links
    against the HIP library.  No GPU call is made."""
    from tval3_amd import _lib as L
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("library not built")
    src = r'''
#include "tval3_gpu_3d.hpp"
template <class T> int use(bool run) {
    mat_host<T> U(8, 8, 8), b(100), Ut(8, 8, 8), Ad(100 * 512);
    sparse_mat_host<T> A(100, 512, 0, sparse_mat_csr);
    geometry_host g;
    tval3_options<T> o;
    o.tv_mode = TV3_TV_ISO;
    if (run) {
        tval3_info<T> i1 = tval3_gpu_3d(U, A, b, o, Ut, true, 0, 1, nullptr);
        tval3_info<T> i2 = tval3_gpu_3d(U, Ad, b, o, Ut);
        tval3_info<T> i3 = tval3_gpu_3d(U, g, b, o, Ut);
        tval3_info<T> i4 = tval3_gpu_3d(U, g, A, b, o, Ut);
        tval3_operator op(A);
        mat_vec_mul(1, op, (const T *)nullptr, (T *)nullptr, nullptr, true);
        return i1.total_iters + i2.total_iters + i3.total_iters + i4.total_iters;
    }
    return 0;
}
int main(int argc, char **) { return use<float>(argc > 5) + use<double>(argc > 5); }
'''
    tmp = os.path.join(ROOT, "tests", "_shim_probe")
    with open(tmp + ".cpp", "w") as f:
        f.write(src)
    try:
        libdir = os.path.dirname(L.LIB_PATH)
        subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-I",
                               os.path.join(ROOT, "include"), tmp + ".cpp", "-o", tmp,
                               "-L" + libdir, "-ltval3_mi355x", "-Wl,-rpath-link,/opt/rocm/lib",
                               "-Wl,--allow-shlib-undefined"])
    finally:
        for p in (tmp, tmp + ".cpp"):
            if os.path.exists(p):
                os.remove(p)
