"""Linux restatement of CalcStarter's orchestration for the regular-mesh models of this package.

The reference's launcher (Code/CalcStarter/main.cpp:179-325) creates one working directory per
frequency and current direction, `Calculation/Nu{k+1}_{X,Y}` (main.cpp:229-246; direction 1 is
'X', 0 is 'Y', main.cpp:28), runs RegularMeshBuilder, Harm1D, BuildMatrix, COCR_FP and OutputNu
in each as separate Windows processes with that directory as cwd and no arguments
(main.cpp:114-131,283-301), for every frequency the X direction first (main.cpp:304-313), then
writes `Results/freq` and `Results/frec` (main.cpp:199-227) and runs ImpedanceZToEdsAll.

Here the stages in front of the solver come from the model layer (mesh numbering, Harm1D and
the BuildMatrix assembly, restated in csrc/model.cpp), which writes every file COCR_FP reads;
the solver is this package's drop-in `COCR_FP` executable, either once per directory exactly as
the reference calls it, or once with `--batch` over all directories (one mesh, every system in
one batched launch); the receiver fields and the impedance come from `receivers`, in FP64.
OutputNu's smoothed field files and ImpedanceZToEdsAll's EDS tables are not produced: the
impedance table `Results/impedance.txt` takes their place (format in `collect`).
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

import numpy as np

from . import model as M
from .receivers import ReceiverOperator, apparent_resistivity, impedance

EXE = Path(__file__).resolve().parent / "lib" / "COCR_FP"
# direction letters of CalcStarter (main.cpp:28) and the polarization each one drives
LETTER = {1: "X", 0: "Y"}


def nu_dir(root, k: int, alpha: int) -> Path:
    """Working directory of frequency index k (0-based) and direction alpha (1: X, 0: Y)."""
    return Path(root) / "Calculation" / f"Nu{k + 1}_{LETTER[alpha]}"


def write_freq_files(root, freqs) -> None:
    """Results/freq: one frequency per line; Results/frec: the count, then the frequencies in
    reverse order (main.cpp:199-227), both "%.13e"."""
    res = Path(root) / "Results"
    res.mkdir(parents=True, exist_ok=True)
    (res / "freq").write_text("".join(f"{f:.13e}\n" for f in freqs))
    (res / "frec").write_text(f"{len(freqs)}\n" + "".join(f"{f:.13e}\n" for f in freqs[::-1]))


def prepare(root, m: M.Model, freqs, eps: float = 1e-8, maxiter: int = 10000,
            nthreads: int = 1) -> list[Path]:
    """Create Calculation/Nu{k}_{X,Y} with every COCR_FP input for each frequency and the
    Results/freq, Results/frec files. Returns the directories in the reference's run order
    (per frequency: X, then Y)."""
    ma = M.mesh_arrays(m)
    dirs = []
    for k, f in enumerate(freqs):
        for alpha in (1, 0):
            s = M.assemble(m, float(f), alpha)
            dirs.append(M.write_cocr_fp_dir(nu_dir(root, k, alpha), m, s, eps=eps,
                                            maxiter=maxiter, nthreads=nthreads, mesh=ma))
    write_freq_files(root, list(freqs))
    return dirs


def run(dirs, batch: bool = True, strict: bool = False, exe=None, timeout=None) -> int:
    """Run the solver over the working directories: one `COCR_FP --batch` process over all of
    them (they must share one mesh), or one process per directory with that directory as cwd
    and no arguments, as CalcStarter calls it. Returns the first non-zero exit code, or 0.
    With strict, a system that did not converge makes the exit code non-zero (2)."""
    exe = str(exe or EXE)
    if not os.path.exists(exe):
        raise RuntimeError(f"COCR_FP executable not found at {exe}: build the package first")
    flags = ["--strict"] if strict else []
    if batch:
        dirs = [str(d) for d in dirs]
        for i in range(0, len(dirs), 64):  # a --batch launch takes up to 64 directories
            r = subprocess.run([exe, *flags, "--batch", *dirs[i:i + 64]], timeout=timeout)
            if r.returncode != 0:
                return r.returncode
        return 0
    for d in dirs:
        r = subprocess.run([exe, *flags], cwd=str(d), timeout=timeout)
        if r.returncode != 0:
            return r.returncode
    return 0


def collect(root, m: M.Model, freqs, receivers) -> np.ndarray:
    """Read every v3.dat, evaluate E and H at the receivers (normal field included) and the
    impedance Z = E H^-1 from the two polarizations. Writes Results/impedance.txt, one line per
    frequency and receiver:
        f  x  y  Re Zxx  Im Zxx  Re Zxy  Im Zxy  Re Zyx  Im Zyx  Re Zyy  Im Zyy
        rho_xy  phase_xy  rho_yx  phase_yx
    (ohm, ohm m, degrees; "%.10e"). Returns Z as an array (frequencies, receivers, 2, 2)."""
    R = ReceiverOperator(m, receivers)
    rec = np.atleast_2d(np.asarray(receivers, dtype=np.float64))
    out = []
    lines = []
    for k, f in enumerate(freqs):
        fields = {}
        for alpha in (1, 0):
            v = np.fromfile(nu_dir(root, k, alpha) / "v3.dat", np.float64)
            fields[alpha] = R.fields(v[0::2] + 1j * v[1::2], float(f), alpha)
        Z = impedance(fields[1][0], fields[1][1], fields[0][0], fields[0][1])
        rho, ph = apparent_resistivity(Z, float(f))
        out.append(Z)
        for r in range(len(rec)):
            z = Z[r]
            vals = [f, rec[r, 0], rec[r, 1]]
            for i, j in ((0, 0), (0, 1), (1, 0), (1, 1)):
                vals += [z[i, j].real, z[i, j].imag]
            vals += [rho[r, 0, 1], ph[r, 0, 1], rho[r, 1, 0], ph[r, 1, 0]]
            lines.append("\t".join(f"{x:.10e}" for x in vals))
    res = Path(root) / "Results"
    res.mkdir(parents=True, exist_ok=True)
    (res / "impedance.txt").write_text("\n".join(lines) + "\n")
    return np.stack(out)


def example_model(folder, step_xy: float | None = None, step_z: float | None = None):
    """The model of a reference calculation folder (CalculationExample: settings.cfg, objects,
    z_sig_2d, xyzVectorB, frequencies; read as RegularMeshBuilder and CalcStarter read them,
    RegularMeshBuilder/main.cpp:1116-1264, CalcStarter/main.cpp:134-176), through the restated
    mesh builder. step_xy / step_z override the settings' start steps. Returns the model, the
    frequencies (Hz, file order) and the receivers (R, 3)."""
    folder = Path(folder)
    s = M.read_settings(folder / "settings.cfg")
    if step_xy:
        s.step_xy = float(step_xy)
    if step_z:
        s.step_z = float(step_z)
    tok = (folder / "frequencies").read_text().split()
    freqs = [float(v) for v in tok[1:1 + int(tok[0])]]
    rec = M.read_receivers(folder / "xyzVectorB")
    m = M.box_model(rec, s, M.read_objects(folder / "objects"), M.read_layers(folder / "z_sig_2d"),
                    name=folder.name or "example")
    return m, freqs, rec


def read_solver_settings(path) -> tuple[float, int]:
    """The SLAE tolerance and iteration cap that follow the mesh numbers in settings.cfg
    (CalcStarter/main.cpp:134-149); RegularMeshBuilder passes them on through issc.txt to
    BuildMatrix, which writes them into kuslau (RegularMeshBuilder/main.cpp:2137,
    BuildMatrix/t_global_slae.cpp:174-190)."""
    tok = Path(path).read_text().split()
    return float(tok[8]), int(tok[9])


RZ_COLUMNS = ("northing (m)", "easting (m)", "periods (s)", "Zxx (Ohm)", "Zxy (Ohm)", "Zyx (Ohm)",
              "Zyy (Ohm)", "Rxx (Ohm)", "Rxy (Ohm)", "Ryx (Ohm)", "Ryy (Ohm)")


def write_rz_table(path, freqs, receivers, Z) -> None:
    """The impedance table in the layout of the reference's published result
    (CalculationResults/625_500_reg_R_Z.txt): a header line, then per receiver (in the given
    order) and period (ascending) the receiver's northing and easting, the period, |Z| of the
    four components and the four apparent resistivities |Z|^2/(omega mu0). The table is in a
    (north, east, down) frame: northing = the model's x, easting = the model's -y."""
    rec = np.atleast_2d(np.asarray(receivers, dtype=np.float64))
    freqs = np.asarray(freqs, dtype=np.float64)
    order = np.argsort(1.0 / freqs, kind="stable")
    lines = ["\t".join(RZ_COLUMNS)]
    for r in range(len(rec)):
        for k in order:
            z = np.abs(Z[k, r])
            rho = z ** 2 / (M.omega_of(float(freqs[k])) * M.MU_0)
            vals = [rec[r, 0], -rec[r, 1] + 0.0, 1.0 / freqs[k], z[0, 0], z[0, 1], z[1, 0], z[1, 1],
                    rho[0, 0], rho[0, 1], rho[1, 0], rho[1, 1]]
            lines.append("\t".join(f"{v:.6e}" for v in vals))
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    Path(path).write_text("\n".join(lines) + "\n")


def main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(description="CalcStarter-equivalent run of a synthetic model: "
                                 "working directories, the COCR_FP solver, impedance table")
    ap.add_argument("root", help="calculation root (Calculation/ and Results/ are created here)")
    ap.add_argument("--model", choices=("config1", "random"), default="config1")
    ap.add_argument("--example", metavar="DIR",
                    help="a reference calculation folder (settings.cfg, objects, z_sig_2d, "
                         "xyzVectorB, frequencies) instead of --model; also writes "
                         "Results/R_Z.txt in the reference's published table layout")
    ap.add_argument("--step-xy", type=float, default=None,
                    help="with --example: horizontal start step overriding settings.cfg")
    ap.add_argument("--cells", type=int, nargs=3, default=(32, 32, 24),
                    help="random model: cells in x, y, z")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--freqs", type=float, nargs="+", default=[0.1, 1.0, 10.0])
    ap.add_argument("--eps", type=float, default=None,
                    help="SLAE tolerance (default: settings.cfg's with --example, else 1e-8)")
    ap.add_argument("--maxiter", type=int, default=None)
    ap.add_argument("--per-directory", action="store_true",
                    help="one solver process per directory (the reference's way)")
    ap.add_argument("--strict", action="store_true")
    a = ap.parse_args(argv)
    eps, maxiter = 1e-8, 10000
    if a.example:
        m, freqs, rec = example_model(a.example, step_xy=a.step_xy)
        eps, maxiter = read_solver_settings(Path(a.example) / "settings.cfg")
    else:
        if a.model == "config1":
            m = M.config1_model()
        else:
            m = M.random_model(*a.cells, seed=a.seed, h=50.0, name="random")
        freqs = a.freqs
        g = np.arange(-200.0, 201.0, 100.0)
        rx, ry = np.meshgrid(g, g, indexing="ij")
        rec = np.stack([rx.ravel(), ry.ravel(), np.zeros(rx.size)], axis=1)
    eps = a.eps if a.eps is not None else eps
    maxiter = a.maxiter if a.maxiter is not None else maxiter
    dirs = prepare(a.root, m, freqs, eps, maxiter)
    rc = run(dirs, batch=not a.per_directory, strict=a.strict)
    if rc != 0:
        return rc
    Z = collect(a.root, m, freqs, rec)
    if a.example:
        write_rz_table(Path(a.root) / "Results" / "R_Z.txt", freqs, rec, Z)
    print(f"results in {Path(a.root) / 'Results'}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
