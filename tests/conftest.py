"""Shared fixtures. GPU tests carry @pytest.mark.gpu and run only on a GPU box (-m gpu)."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


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
