"""CalcStarter-equivalent launcher (host side): the working-directory tree, the frequency
files, and the inputs written for the solver."""
import numpy as np


def test_prepare_tree_and_frequency_files(M, small_model, tmp_path):
    from mtcalc_with_dfp_cocr_amd import launcher as L

    freqs = [0.1, 2.0]
    dirs = L.prepare(tmp_path, small_model, freqs, eps=1e-9, maxiter=500)
    # per frequency X then Y, named as CalcStarter names them (main.cpp:28,231)
    assert [d.name for d in dirs] == ["Nu1_X", "Nu1_Y", "Nu2_X", "Nu2_Y"]
    assert all(d.parent.name == "Calculation" for d in dirs)
    assert (tmp_path / "Results" / "freq").read_text() == "1.0000000000000e-01\n2.0000000000000e+00\n"
    assert (tmp_path / "Results" / "frec").read_text() == (
        "2\n2.0000000000000e+00\n1.0000000000000e-01\n")
    for d, (f, a) in zip(dirs, [(0.1, 1), (0.1, 0), (2.0, 1), (2.0, 0)]):
        s = M.read_block_system(d)
        ref = M.assemble(small_model, f, a)
        assert s.nb == ref.nb and abs(s.nu - f) <= 1e-12 * f
        assert np.array_equal(s.pr, ref.pr) and np.array_equal(s.gg, ref.gg)
        assert (d / "alfa").read_text().strip() == str(a)
        n, eps, maxiter = (d / "kuslau").read_text().split()[:3]
        assert int(n) == 2 * ref.nb and float(eps) == 1e-9 and int(maxiter) == 500


def test_example_folder_and_rz_table(M, tmp_path):
    """A reference calculation folder (the example's files, written back from the fixture in
    their own text formats) gives the restated builder's example mesh; the R_Z table carries,
    per receiver and ascending period, |Z| and rho_a in the published table's columns, in its
    (north, east, down) frame."""
    import json
    from pathlib import Path

    from mtcalc_with_dfp_cocr_amd import launcher as L

    d = json.loads((Path(__file__).resolve().parent / "golden" / "dtm1_example.json").read_text())
    ex = tmp_path / "CalculationExample"
    ex.mkdir()
    s = d["settings"]
    (ex / "settings.cfg").write_text(
        "".join(f"{s[i]:g} {s[i + 1]:g}\n" for i in range(0, 8, 2)) + "0.0001\n10000\n\nnotes\n")
    (ex / "objects").write_text(f"{len(d['objects'])}\n" + "".join(
        " ".join(f"{v:g}" for v in o) + "\n" for o in d["objects"]))
    (ex / "z_sig_2d").write_text(f"{len(d['layers'])}\n" + "".join(
        f"{z:g} {sg:g}\n" for z, sg in d["layers"]))
    (ex / "xyzVectorB").write_text(f"{len(d['receivers'])}\n" + "".join(
        "\t".join(f"{v:g}" for v in r) + "\n" for r in d["receivers"]))
    (ex / "frequencies").write_text(f"{len(d['frequencies'])}\n" + "".join(
        f"{f:g}\n" for f in d["frequencies"]))
    m, freqs, rec = L.example_model(ex)
    assert m.n_edges == 135152 and freqs == d["frequencies"] and len(rec) == 53
    assert L.read_solver_settings(ex / "settings.cfg") == (1e-4, 10000)
    rng = np.random.default_rng(0)
    Z = rng.standard_normal((len(freqs), 2, 2, 2)) + 1j * rng.standard_normal((len(freqs), 2, 2, 2))
    out = tmp_path / "R_Z.txt"
    L.write_rz_table(out, freqs, rec[:2], Z)
    lines = out.read_text().splitlines()
    assert lines[0].split("\t") == list(L.RZ_COLUMNS) and len(lines) == 1 + 2 * len(freqs)
    rows = np.array([[float(v) for v in ln.split("\t")] for ln in lines[1:]])
    periods = np.sort(1.0 / np.array(freqs))
    for r in range(2):
        blk = rows[r * len(freqs):(r + 1) * len(freqs)]
        np.testing.assert_allclose(blk[:, 2], periods, rtol=1e-6)
        assert np.all(blk[:, 0] == rec[r, 0]) and np.all(blk[:, 1] == -rec[r, 1])
        for k, T in enumerate(blk[:, 2]):
            f = 1.0 / T
            i = int(np.argmin(np.abs(np.array(freqs) - f)))
            z = np.abs(Z[i, r])
            np.testing.assert_allclose(blk[k, 3:7], [z[0, 0], z[0, 1], z[1, 0], z[1, 1]],
                                       rtol=1e-6)
            np.testing.assert_allclose(blk[k, 7:11], blk[k, 3:7] ** 2 / (M.omega_of(f) * M.MU_0),
                                       rtol=1e-5)
