"""Build the native library and the drop-in executable, in-tree.

Outputs (git-ignored, shipped with the tree):
  mtcalc_with_dfp_cocr_amd/lib/libmtcocr.so   HIP solver + model layer + C-ABI (include/*.h)
  mtcalc_with_dfp_cocr_amd/lib/COCR_FP        drop-in executable for the reference's COCR_FP

HIP sources are compiled for gfx950 only; host C++ is compiled with g++ and OpenMP (libgomp,
the OpenMP runtime PyTorch already loads).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "lib"
OBJ = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-Wall", "-Wno-unused-result",
             f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build(verbose: bool = True) -> Path:
    """Compile everything; returns the path of libmtcocr.so."""
    OBJ.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    objs = []
    for src in ["model.cpp", "host_prep.cpp"]:
        o = OBJ / (src + ".o")
        _run([CXX, *CXX_FLAGS, "-c", CSRC / src, "-o", o])
        objs.append(o)
    o = OBJ / "solver.hip.o"
    _run([HIPCC, *HIP_FLAGS, "-c", CSRC / "solver.hip", "-o", o])
    objs.append(o)
    so = LIB / "libmtcocr.so"
    tmp = LIB / "libmtcocr.so.tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lgomp",
          "-Wl,-soname,libmtcocr.so"])
    os.replace(tmp, so)
    exe_o = OBJ / "cocr_fp_main.cpp.o"
    _run([CXX, *CXX_FLAGS, "-c", CSRC / "cocr_fp_main.cpp", "-o", exe_o])
    _run([CXX, "-o", LIB / "COCR_FP", exe_o, f"-L{LIB}", "-lmtcocr", "-Wl,-rpath,$ORIGIN",
          "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-lgomp"])
    return so


def build_oracle() -> None:
    """Compile the oracle (test infrastructure) with its own Makefile."""
    _run(["make", "-C", ROOT / "oracle", "-s"])


if __name__ == "__main__":
    build()
    if "--oracle" in sys.argv:
        build_oracle()
