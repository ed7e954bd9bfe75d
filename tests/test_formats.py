"""File contract and C-ABI surface (CPU only).

* the oracle's file-driven process (a restatement of COCR_FP's main) reads a working directory
  written in the reference's formats and gives the same solution as the in-memory oracle;
* capitalised file names written by the reference's mesh builder are accepted;
* a missing input file is an error;
* the native library exports every function declared in include/*.h (no compute calls).
"""
import ctypes
import os
import re
import shutil
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def small_dir(M, small_model, tmp_path_factory):
    s = M.assemble(small_model, 1.0, 1)
    d = M.write_cocr_fp_dir(tmp_path_factory.mktemp("w") / "Nu1_X", small_model, s)
    return s, d


def test_oracle_process_matches_in_memory(O, M, small_model, small_dir):
    s, d = small_dir
    rc, st = O.cocr_fp_dir(d, 1)
    assert rc == 0 and st["req"] == 0
    v3 = np.fromfile(d / "v3.dat", np.float64)
    x, st2, _ = O.pcocr(s, M.mesh_arrays(small_model), nthreads=1)
    assert np.array_equal(v3, x) and st["iter"] == st2["iter"]
    kit = (d / "kit").read_text()
    m = re.fullmatch(r"r=(\S+)\teps=(\S+)\tkit=(\d+)\ttime=(\S+) sec\n", kit)
    assert m and int(m.group(3)) == st["iter"] and m.group(2) == "1.000000e-08"


def test_capitalised_names_and_missing_files(O, M, small_model, small_dir, tmp_path):
    s, d = small_dir
    e = tmp_path / "caps"
    shutil.copytree(d, e)
    for a, b in [("sig3d", "Sig3d"), ("l13d.dat", "L13d.dat"), ("mu3d", "mu3D"),
                 ("dpr3d", "dpr3D")]:
        os.rename(e / a, e / b)
    rc, st = O.cocr_fp_dir(e, 1)
    assert rc == 0 and st["req"] == 0
    x, _, _ = O.pcocr(s, M.mesh_arrays(small_model), nthreads=1)
    np.testing.assert_array_equal(np.fromfile(e / "v3.dat"), x)
    os.remove(e / "gg")
    rc, _ = O.cocr_fp_dir(e, 1)
    assert rc != 0


def test_sigma_file_rounding(M, small_model, small_dir):
    """sig3d carries %.7e text; the in-memory model is rounded the same way, so material
    conductivities read back from the file are the ones assembled."""
    s, d = small_dir
    mats = np.loadtxt(d / "sig3d")
    ma = M.mesh_arrays(small_model)
    np.testing.assert_array_equal(mats[:, 1], ma.materials[:, 0])
    np.testing.assert_array_equal(mats[ma.nvkat, 1], small_model.sigma)


def _declared(header: Path):
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(mt(?:cocr|model)_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from mtcalc_with_dfp_cocr_amd._lib import LIB_PATH, SIGNATURES

    lib = ctypes.CDLL(str(LIB_PATH))
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        names += _declared(h)
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
        assert n in SIGNATURES, n


def test_host_argument_errors_without_gpu():
    """Argument checks that run before any device call."""
    from mtcalc_with_dfp_cocr_amd._lib import lib

    L = lib()
    assert L.mtcocr_destroy(None) == -10
    assert L.mtcocr_error_string(-3).decode() == "maximum number of iterations reached"
    assert L.mtmodel_num_edges(1, 5, 5) == -10
    assert L.mtmodel_num_edges(2, 2, 2) == 12
    z3, z2, out = np.zeros(3), np.zeros(2), np.zeros(8)
    assert L.mtcocr_sigma_gradient(None, 3, 3, 3, z3, z3, z3, 1, z2, np.zeros(108),
                                   np.zeros(108), out) == -10


def test_headers_compile_as_c99(tmp_path):
    """include/*.h are plain C: a C99 translation unit that includes both compiles warning-free
    (what a cgo / FFI binding of the reference side would include)."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "abi.c"
    src.write_text('#include "mtcocr.h"\n#include "mtmodel.h"\n'
                   "int main(void) {\n"
                   "  int (*create)(int, mtcocr_handle**) = mtcocr_create;\n"
                   "  int64_t (*edges)(int, int, int) = mtmodel_num_edges;\n"
                   "  (void)create;\n  (void)edges;\n  return 0;\n}\n")
    r = subprocess.run([cc, "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        f"-I{ROOT / 'include'}", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
