"""ORACLE — test infrastructure only: ctypes wrapper of build/liboracle_cocr.so.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import this module,
and only as the checker or the timed CPU baseline; the product never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "build" / "liboracle_cocr.so"
EXE = HERE / "build" / "cocr_fp_oracle"

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OrcStats(C.Structure):
    _fields_ = [("req", C.c_int), ("iter", C.c_int), ("residual_rel", C.c_double),
                ("seconds", C.c_double), ("n_spmv", C.c_int), ("n_precond", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class OrcDfp(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int),
                ("iptr_g", C.POINTER(C.c_int)), ("jptr_g", C.POINTER(C.c_int)),
                ("gg_g", C.POINTER(C.c_double)),
                ("iptr_gt", C.POINTER(C.c_int)), ("jptr_gt", C.POINTER(C.c_int)),
                ("gg_gt", C.POINTER(C.c_double)),
                ("diag_a", C.POINTER(C.c_double)), ("diag_gagt", C.POINTER(C.c_double)),
                ("help", C.POINTER(C.c_double))]


_lib = None


def build():
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not SO.exists():
        build()
    L = C.CDLL(str(SO))
    L.orc_spmv.argtypes = [C.c_int, i32p, i32p, i32p, i32p, f64p, f64p, f64p, f64p, C.c_int,
                           C.c_void_p]
    L.orc_spmv.restype = None
    L.orc_scal_cmplx.argtypes = [f64p, f64p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
    L.orc_scal_cmplx.restype = None
    L.orc_build_g.argtypes = [C.POINTER(OrcDfp), C.c_int, C.c_int, C.c_int, i32p, i32p, f64p,
                              i32p, C.c_void_p, C.c_void_p, C.c_void_p, i32p]
    L.orc_build_diag.argtypes = [C.POINTER(OrcDfp), i32p, i32p, i32p, i32p, f64p, f64p]
    L.orc_apply_precond.argtypes = [C.POINTER(OrcDfp), f64p, f64p, C.c_int]
    L.orc_apply_precond.restype = None
    L.orc_dfp_free.argtypes = [C.POINTER(OrcDfp)]
    L.orc_dfp_free.restype = None
    L.orc_pcocr_2x2_folded.argtypes = [C.c_int, i32p, i32p, i32p, i32p, f64p, f64p, f64p, f64p,
                                       C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, i32p,
                                       i32p, f64p, i32p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       i32p, C.c_int, C.POINTER(OrcStats), C.c_void_p, C.c_int]
    L.orc_pcocr_prebuilt_g.argtypes = [C.POINTER(OrcDfp), C.c_int, i32p, i32p, i32p, i32p, f64p,
                                       f64p, f64p, f64p, C.c_double, C.c_int, C.c_int,
                                       C.POINTER(OrcStats), C.c_void_p, C.c_int, C.c_void_p]
    L.orc_cocr_fp_main.argtypes = [C.c_char_p, C.c_int, C.POINTER(OrcStats)]
    _lib = L
    return L


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def spmv(sys, x, nthreads=1):
    y = np.empty(2 * sys.nb)
    lib().orc_spmv(sys.nb, _i(sys.ig), _i(sys.jg), _i(sys.idi), _i(sys.ijg), _d(sys.di),
                   _d(sys.gg), _d(x), y, nthreads, None)
    return y


def scal_cmplx(x, y, nthreads=1):
    re, im = C.c_double(), C.c_double()
    lib().orc_scal_cmplx(_d(x), _d(y), len(x) // 2, nthreads, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


class Dfp:
    """The folded preconditioner of the oracle (G, G^T, diagonals) with numpy views."""

    def __init__(self, ma, sys=None):
        self.p = OrcDfp()
        rc = lib().orc_build_g(C.byref(self.p), ma.kpar, len(ma.edges) // 2, ma.kuzlov,
                               _i(ma.nvkat), _i(ma.nver), _d(ma.sig_flags), _i(ma.edges), None,
                               None, None, _i(ma.is_node_bound))
        assert rc == 0
        if sys is not None:
            self.build_diag(sys)

    def build_diag(self, sys):
        rc = lib().orc_build_diag(C.byref(self.p), _i(sys.ig), _i(sys.jg), _i(sys.idi),
                                  _i(sys.ijg), _d(sys.di), _d(sys.gg))
        assert rc == 0

    @property
    def m(self):
        return self.p.m

    def g_csr(self):
        n, m = self.p.n, self.p.m
        ip = np.ctypeslib.as_array(self.p.iptr_g, (m + 1,)).copy()
        nnz = ip[-1]
        jp = np.ctypeslib.as_array(self.p.jptr_g, (nnz,)).copy()
        gv = np.ctypeslib.as_array(self.p.gg_g, (nnz,)).copy()
        return ip, jp, gv

    def diagonals(self):
        n, m = self.p.n, self.p.m
        da = np.ctypeslib.as_array(self.p.diag_a, (2 * n,)).copy()
        dn = np.ctypeslib.as_array(self.p.diag_gagt, (2 * m,)).copy()
        return da, dn

    def apply(self, x, nthreads=1):
        y = np.empty(2 * self.p.n)
        lib().orc_apply_precond(C.byref(self.p), _d(x), y, nthreads)
        return y

    def __del__(self):
        try:
            lib().orc_dfp_free(C.byref(self.p))
        except Exception:
            pass


def pcocr(sys, ma, eps=1e-8, maxiter=10000, nthreads=1, x0=None, trace=False):
    """Whole solver (PCOCR_2x2_Folded restated). Returns (x, stats, trace-or-None)."""
    x = np.zeros(2 * sys.nb) if x0 is None else _d(x0).copy()
    st = OrcStats()
    tr = np.zeros(maxiter + 1) if trace else None
    lib().orc_pcocr_2x2_folded(
        2 * sys.nb, _i(sys.ig), _i(sys.jg), _i(sys.idi), _i(sys.ijg), _d(sys.di), _d(sys.gg),
        _d(sys.pr), x, float(eps), int(maxiter), ma.kpar, sys.nb, ma.kuzlov, _i(ma.nvkat),
        _i(ma.nver), _d(ma.sig_flags), _i(ma.edges), None, None, None, _i(ma.is_node_bound),
        int(nthreads), C.byref(st),
        None if tr is None else tr.ctypes.data_as(C.c_void_p), 0 if tr is None else len(tr))
    d = st.as_dict()
    return x, d, (None if tr is None else tr[:d["iter"]])


def pcocr_with_mrs(sys, dfp, eps=1e-8, maxiter=10000, nthreads=1):
    """Solver on a prebuilt Dfp; returns (x, y_mrs, stats, trace)."""
    x = np.zeros(2 * sys.nb)
    y = np.zeros(2 * sys.nb)
    tr = np.zeros(maxiter + 1)
    st = OrcStats()
    lib().orc_pcocr_prebuilt_g(C.byref(dfp.p), 2 * sys.nb, _i(sys.ig), _i(sys.jg), _i(sys.idi),
                               _i(sys.ijg), _d(sys.di), _d(sys.gg), _d(sys.pr), x, float(eps),
                               int(maxiter), int(nthreads), C.byref(st),
                               tr.ctypes.data_as(C.c_void_p), len(tr),
                               y.ctypes.data_as(C.c_void_p))
    d = st.as_dict()
    return x, y, d, tr[:d["iter"]]


def cocr_fp_dir(path, nthreads=0):
    st = OrcStats()
    rc = lib().orc_cocr_fp_main(str(path).encode(), int(nthreads), C.byref(st))
    return rc, st.as_dict()
