"""ctypes binding of lib/libmtcocr.so (declared in include/mtcocr.h and include/mtmodel.h).

The library is built in-tree by ``mtcalc_with_dfp_cocr_amd.build``. There is no fallback: if
it is missing, every entry point raises, so nothing can silently run on another path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libmtcocr.so"

_lib = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class MtcocrStats(C.Structure):
    _fields_ = [
        ("req", C.c_int),
        ("iter", C.c_int),
        ("outer", C.c_int),
        ("flags", C.c_int),
        ("residual_rel", C.c_double),
        ("true_rel_res", C.c_double),
        ("seconds", C.c_double),
        ("seconds_io", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


# exported symbols and their signatures; tests check that every declaration of include/*.h
# is exported
SIGNATURES = {
    # mtmodel.h
    "mtmodel_num_edges": (C.c_int64, [C.c_int, C.c_int, C.c_int]),
    "mtmodel_topology": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.POINTER(C.c_int64)]),
    "mtmodel_portrait_size": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
    "mtmodel_portrait": (C.c_int, [C.c_int, C.c_int, C.c_int, i32p, i32p, i32p, i32p]),
    "mtmodel_harm1d": (C.c_int, [C.c_int, f64p, f64p, C.c_double, C.c_int, C.POINTER(C.c_int),
                                 C.c_void_p, C.c_void_p, C.c_void_p]),
    "mtmodel_assemble": (C.c_int, [C.c_int, C.c_int, C.c_int, f64p, f64p, f64p, f64p, f64p, f64p,
                                   C.c_double, C.c_int, C.c_int, f64p, f64p, f64p, i32p, i32p,
                                   i32p, i32p, f64p, f64p, f64p]),
    "mtmodel_assemble_rhs": (C.c_int, [C.c_int, C.c_int, C.c_int, f64p, f64p, f64p, f64p, f64p,
                                       C.c_double, C.c_int, C.c_int, f64p, f64p, f64p, f64p]),
    "mtmodel_assemble_ks": (C.c_int, [C.c_int, C.c_int, C.c_int, f64p, f64p, f64p, f64p, f64p,
                                      i32p, i32p, f64p, f64p, f64p, f64p]),
    "mtmodel_element_matrices": (C.c_int, [C.c_double, C.c_double, C.c_double, f64p, f64p]),
    "mtmodel_normal_dofs": (C.c_int, [C.c_int, C.c_int, C.c_int, f64p, f64p, f64p, C.c_int,
                                      C.c_int, f64p, f64p, f64p, f64p]),
    "mtmodel_cell_mass_forms": (C.c_int, [C.c_int, C.c_int, C.c_int, f64p, f64p, f64p, f64p,
                                          f64p, f64p]),
    # mtcocr.h
    "mtcocr_error_string": (C.c_char_p, [C.c_int]),
    "mtcocr_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "mtcocr_destroy": (C.c_int, [C.c_void_p]),
    "mtcocr_set_matrix_block2x2": (C.c_int, [C.c_void_p, C.c_int, i32p, i32p, i32p, i32p, f64p,
                                             f64p]),
    "mtcocr_set_matrix_ks": (C.c_int, [C.c_void_p, C.c_int, i32p, i32p, f64p, f64p, f64p, f64p]),
    "mtcocr_set_gradient_mesh": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, i32p, i32p,
                                           f64p, i32p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           i32p]),
    "mtcocr_solve": (C.c_int, [C.c_void_p, f64p, f64p, C.c_double, C.c_int, C.c_int,
                               C.POINTER(MtcocrStats)]),
    "mtcocr_solve_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, f64p, f64p, C.c_double,
                                     C.c_int, C.c_int, C.POINTER(MtcocrStats)]),
    "mtcocr_set_rhs": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, f64p, C.c_void_p]),
    "mtcocr_solve_resident": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int,
                                        C.POINTER(MtcocrStats)]),
    "mtcocr_get_solution": (C.c_int, [C.c_void_p, f64p]),
    "mtcocr_pcocr_2x2_folded": (C.c_int, [C.c_int, i32p, i32p, i32p, i32p, f64p, f64p, f64p, f64p,
                                          C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_void_p, i32p, i32p, f64p, i32p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, i32p,
                                          C.POINTER(MtcocrStats)]),
    "mtcocr_set_mesh_order": (C.c_int, [C.c_void_p, C.c_int, C.c_int, i32p, C.c_int, i32p,
                                        C.c_int]),
    "mtcocr_mesh_order": (C.c_int, [C.c_int, C.c_int, i32p, C.c_int, i32p, C.c_int, i32p]),
    "mtcocr_build_gmatrix": (C.c_int, [C.c_int, C.c_int, C.c_int, i32p, i32p, f64p, i32p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, i32p,
                                       C.POINTER(C.c_int), C.POINTER(C.c_int64), C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p]),
    "mtcocr_bench_copy": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    "mtcocr_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mtcocr_get_launch": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]),
    "mtcocr_set_stencil_sweep": (C.c_int, [C.c_void_p, C.c_int]),
    "mtcocr_get_stencil_sweep": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "mtcocr_set_stencil_band": (C.c_int, [C.c_void_p, C.c_int]),
    "mtcocr_get_stencil_band": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "mtcocr_set_xp_period": (C.c_int, [C.c_void_p, C.c_int]),
    "mtcocr_get_xp_period": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "mtcocr_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]),
    "mtcocr_bench_spmv": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mtcocr_bench_update": (C.c_int, [C.c_void_p, C.c_int, C.c_int,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mtcocr_bench_flush": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "mtcocr_apply": (C.c_int, [C.c_void_p, C.c_int, C.c_int, f64p, f64p]),
    "mtcocr_apply_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                       C.c_void_p]),
    "mtcocr_matrix_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mtcocr_sigma_gradient": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, f64p, f64p, f64p,
                                        C.c_int, f64p, f64p, f64p, f64p]),
}


def lib() -> C.CDLL:
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m mtcalc_with_dfp_cocr_amd.build` "
            "(the HIP library is required; there is no CPU fallback)")
    L = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int, what: str) -> int:
    """Raise on argument / device errors (codes <= -10 except the stall code -14)."""
    if rc <= -10 and rc != -14:
        msg = lib().mtcocr_error_string(rc).decode()
        raise RuntimeError(f"{what} failed ({rc}): {msg}")
    return rc


def ptr(a):
    """ctypes void pointer of a numpy array, or None."""
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)
