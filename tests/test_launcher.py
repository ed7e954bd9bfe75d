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
