"""Shared fixtures. GPU tests carry @pytest.mark.gpu and run only on a GPU box (-m gpu)."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


# True relative residual ||b - A x|| / ||b|| allowed for the x an FP64 solve returns. The stop
# test is on the MRS-smoothed residual (< eps = 1e-8); the raw iterate's residual at the stop
# point depends on the rounding path. The oracle's own x reaches 2.0e-8 on 8 threads and
# 1.06e-7 on 16 on the benchmark model's system 13 (19.3 Hz, Y), and 1.65e-7 on one thread on
# system 19 (3.98 Hz, Y) of the 48^3 model of the four-model workload; GPU launch shapes spread
# the same way (2.9e-7 there, 1.5e-7 on system 13 at the autotuned shapes). The bound holds
# the method's attainable accuracy at the stop point; iterations and impedances are held
# against the oracle tightly beside it.
FP64_TRUE_BOUND = 5e-7

_capman = None


def pytest_configure(config):
    global _capman
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    _capman = config.pluginmanager.getplugin("capturemanager")


def say(msg: str) -> None:
    """A progress line of a GPU test, printed past pytest's output capture (as `-s` would), so
    that a test running for minutes reports its steps as they end."""
    line = f"[gpu-test] {msg}"
    if _capman is not None:
        with _capman.global_and_fixture_disabled():
            print(line, flush=True)
    else:
        print(line, flush=True)


_gpu = None


def gpu_available() -> bool:
    """HIP device count through the runtime library (no torch import)."""
    global _gpu
    if _gpu is None:
        try:
            import ctypes

            hip = ctypes.CDLL("libamdhip64.so")
            n = ctypes.c_int(0)
            _gpu = hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
        except OSError:
            _gpu = False
    return _gpu


@pytest.fixture(scope="session")
def M():
    from mtcalc_with_dfp_cocr_amd import model

    return model


@pytest.fixture(scope="session")
def O():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def tiny_model(M):
    """6x6x6 cells, 1 padding cell, 2 air cells: 882 edges (dense checks)."""
    return M.random_model(6, 6, 6, seed=3, h=100.0, n_pad=1, n_air=2, name="tiny")


@pytest.fixture(scope="session")
def small_model(M):
    """12x10x10 cells with padding and air: a few thousand edges."""
    return M.random_model(12, 10, 10, seed=5, h=60.0, n_pad=2, n_air=3, name="small")


@pytest.fixture(scope="session")
def c1_model(M):
    return M.config1_model()


@pytest.fixture(scope="session")
def c1_system(M, c1_model):
    return M.assemble(c1_model, 1.0, 1)


@pytest.fixture(scope="session")
def c1_mesh(M, c1_model):
    return M.mesh_arrays(c1_model)
