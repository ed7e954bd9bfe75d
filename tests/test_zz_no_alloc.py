"""The allocation-free entry point mtcocr_apply_noalloc against mtcocr_apply on the same
handle: SpMV and folded preconditioner, bit for bit, on two calls in a row (the config-1
system, then a smaller one on another handle, which reuses the same scratch array), inputs and
outputs in device memory that the test allocates itself."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)


class Hip:
    """The few runtime calls the test needs for its own device buffers."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")

    def check(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), C.c_size_t(nbytes)), "hipMalloc")
        return p

    def h2d(self, dst, a):
        self.check(self.L.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1),
                   "hipMemcpy H2D")

    def d2h(self, a, src):
        self.check(self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, C.c_size_t(a.nbytes), 2),
                   "hipMemcpy D2H")

    def sync(self):
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self, p):
        self.L.hipFree(p)


def test_apply_noalloc_matches_apply(M, c1_system, c1_mesh, small_model):
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd._lib import lib
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    hip = Hip()
    L = lib()
    small = M.assemble(small_model, 3.0, 1)
    cases = [(c1_system, c1_mesh), (small, M.mesh_arrays(small_model))]
    assert small.nb < c1_system.nb
    rng = np.random.default_rng(31)
    for s, ma in cases:
        say(f"apply_noalloc on nb = {s.nb}")
        n = 2 * s.nb
        x = rng.standard_normal(n)
        dx, dy = hip.malloc(8 * n), hip.malloc(8 * n)
        try:
            hip.h2d(dx, x)
            with Solver(0) as H:
                H.set_matrix(s)
                H.set_gradient(ma)
                for op, name in ((0, "spmv"), (1, "precond")):
                    ref = H.apply(name, x, "fp64")
                    hip.h2d(dy, np.full(n, np.nan))
                    rc = L.mtcocr_apply_noalloc(H._h, op, 1.0, dx, dy, None)
                    assert rc == 0, L.mtcocr_error_string(rc).decode()
                    hip.sync()
                    got = np.empty(n)
                    hip.d2h(got, dy)
                    assert np.array_equal(got, ref), f"{name}: max diff {np.abs(got - ref).max()}"
                # refusals: a block2x2 matrix takes omega = 1 only; null pointers
                assert L.mtcocr_apply_noalloc(H._h, 0, 2.0, dx, dy, None) == -10
                assert L.mtcocr_apply_noalloc(H._h, 0, 1.0, None, dy, None) == -10
        finally:
            hip.free(dx)
            hip.free(dy)
