"""GPU parity: the HIP path through the C-ABI against the oracle (FP64 CPU restatement).

Tolerances:
  SpMV / preconditioner, FP64 storage: 1e-13 / 1e-11 of max|y| (summation order only)
  SpMV / preconditioner, FP32 storage: 2e-6 / 2e-5 of max|y|
  full FP64 solve: iteration count within 2% (+2) of the oracle, MRS residual < eps, true
    residual <= 3e-8 (the reference's own returned x reaches 1.0e-8..2.9e-8), x within 1e-5
    relative of the oracle's x (rounding growth through COCR on a system with 1e-8 S/m air)
  mixed precision: true FP64 residual < eps = 1e-8 (systems whose FP32 refinement stalls are
    finished by the FP64 solver and flagged), ||x_mixed - x_oracle|| <= 1e-5 ||x_oracle||
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import gpu_available, say
from helpers import identity_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S(c1_system, c1_mesh):
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say("create solver, config-1 system (45000 edges)")
    s = Solver(0)
    s.set_matrix(c1_system)
    s.set_gradient(c1_mesh)
    yield s
    s.close()


@pytest.fixture(scope="module")
def oracle_c1(O, c1_system, c1_mesh):
    x, st, tr = O.pcocr(c1_system, c1_mesh, eps=1e-8, nthreads=1, trace=True)
    return x, st


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("mixed", 2e-6)])
def test_spmv_parity(S, O, M, c1_system, prec, tol):
    say(f"spmv {prec}")
    x = np.random.default_rng(11).standard_normal(2 * c1_system.nb)
    ref = O.spmv(c1_system, x)
    got = S.apply("spmv", x, prec)
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("nsys", [4, 8, 16, 32, 64])
@pytest.mark.parametrize("layout", ["ell", "rowpack", "rowblock"])
def test_spmv_parity_batch_widths(O, M, c1_system, c1_mesh, nsys, layout):
    """Every SpMV layout at every batch width with a row-group kernel (4..64 lanes per row):
    system 0 of the batch against the oracle, FP64 and FP32 vectors. Row blocks run up to 16
    systems (the row pack above) and are checked both in the file's edge order and in the
    pencil order."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say(f"spmv layout {layout}, {nsys} systems")
    x = np.random.default_rng(13).standard_normal(2 * c1_system.nb)
    ref = O.spmv(c1_system, x)
    with Solver(0) as B:
        B.set_matrix(c1_system)
        B.set_gradient(c1_mesh)
        B.set_rhs(np.tile(c1_system.pr, (nsys, 1)))
        B.set_launch(spmv_layout=layout)
        runs = "rowpack" if (layout == "rowblock" and nsys > 16) else layout
        for prec, tol in (("fp64", 1e-13), ("mixed", 2e-6)):
            assert B.get_launch(prec)["spmv_layout"] == runs
            got = B.apply("spmv", x, prec)
            assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    if layout == "rowblock" and nsys <= 16:
        with Solver(0) as B:
            B.set_mesh_order(c1_mesh, 16)
            B.set_matrix(c1_system)
            B.set_gradient(c1_mesh)
            B.set_rhs(np.tile(c1_system.pr, (nsys, 1)))
            B.set_launch(spmv_layout=layout)
            for prec, tol in (("fp64", 1e-13), ("mixed", 2e-6)):
                assert B.get_launch(prec)["spmv_layout"] == layout
                got = B.apply("spmv", x, prec)
                assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("nsys", [5, 8])
def test_spmv_stencil_lds_8_systems(M, c1_model, c1_mesh, nsys):
    """The LDS-staged stencil SpMV at batch width 8 (5 systems padded to 8, or 8): its 16-lane
    rows hold 8 systems x (re, im), each lane one component, with the same FMAs in the same
    order as the 16-system kernel's, so system 0's product equals that of a 16-system batch
    bit for bit, FP64 and FP32 vectors, at a sweep of 3 levels and of whole columns; and
    matches an exact K + i omega S product."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import ORDER_GRID, Solver

    say(f"stencil_lds SpMV at width 8, {nsys} systems")
    freqs = M.frequencies(1e-2, 1e2, 8)
    cases = [(f, a) for f in freqs for a in (1, 0)]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    x = np.random.default_rng(23).standard_normal(2 * p.nb)
    import scipy.sparse as sp

    rows = np.repeat(np.arange(p.nb), np.diff(p.ig))
    full = lambda d, g: (sp.csr_matrix((g, (rows, p.jg)), shape=(p.nb, p.nb))
                         + sp.csr_matrix((g, (p.jg, rows)), shape=(p.nb, p.nb)) + sp.diags(d))
    ref = M.to_interleaved((full(kd, kg) + 1j * om[0] * full(sd, sg)) @ M.to_complex(x))
    with Solver(0) as B:
        B.set_mesh_order(c1_mesh, ORDER_GRID)
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(c1_mesh)
        B.set_launch(spmv_layout="stencil_lds")
        for kc in (3, len(c1_model.z)):
            B.set_stencil_sweep(kc)
            for prec, tol in (("fp64", 1e-13), ("mixed", 2e-6)):
                B.set_rhs(prs, om)
                y16 = B.apply("spmv", x, prec)
                B.set_rhs(prs[:nsys], om[:nsys])
                assert B.get_launch(prec)["spmv_layout"] == "stencil_lds"
                y8 = B.apply("spmv", x, prec)
                assert np.abs(y8 - ref).max() <= tol * np.abs(ref).max(), (kc, prec)
                assert np.array_equal(y8, y16), (kc, prec)


@pytest.mark.parametrize("layout", ["stencil", "stencil_lds"])
@pytest.mark.parametrize("nsys", [9, 16])
def test_spmv_stencil_layout(O, M, c1_model, c1_mesh, nsys, layout):
    """The structured stencil SpMVs (matrix and vectors in the grid order of the box mesh; wave
    tasks, or 4 x 4 tiles with x staged in LDS): system 0 of a batch of 9 or 16 (both run
    16-lane rows) against an exact K + i omega S product, FP64 and FP32 vectors, at a sweep of
    3 levels (chunks that do not divide the mesh) and of whole columns; at 8 systems the
    LDS-staged layout stays, the wave-task one falls back to the row blocks or the row pack."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import ORDER_GRID, Solver

    say(f"{layout} SpMV, {nsys} systems")
    freqs = M.frequencies(1e-2, 1e2, 8)
    cases = [(f, a) for f in freqs for a in (1, 0)][:nsys]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    x = np.random.default_rng(17).standard_normal(2 * p.nb)
    import scipy.sparse as sp

    rows = np.repeat(np.arange(p.nb), np.diff(p.ig))
    full = lambda d, g: (sp.csr_matrix((g, (rows, p.jg)), shape=(p.nb, p.nb))
                         + sp.csr_matrix((g, (p.jg, rows)), shape=(p.nb, p.nb)) + sp.diags(d))
    A0 = full(kd, kg) + 1j * om[0] * full(sd, sg)  # system 0 of the batch: K + i omega_0 S
    ref = M.to_interleaved(A0 @ M.to_complex(x))
    with Solver(0) as B:
        B.set_mesh_order(c1_mesh, ORDER_GRID)
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(c1_mesh)
        B.set_rhs(prs, om)
        B.set_launch(spmv_layout=layout)
        for kc in (3, len(c1_model.z)):
            B.set_stencil_sweep(kc)
            for prec, tol in (("fp64", 1e-13), ("mixed", 2e-6)):
                assert B.get_launch(prec)["spmv_layout"] == layout
                got = B.apply("spmv", x, prec)  # system 0 of the batch
                assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), (kc, prec)
                if layout == "stencil_lds":
                    # each row's products in the same order as the wave-task kernel: equal bits
                    B.set_launch(spmv_layout="stencil")
                    assert np.array_equal(B.apply("spmv", x, prec), got), (kc, prec)
                    B.set_launch(spmv_layout=layout)
        # 8 systems: the LDS-staged stencil serves them too (8 systems x (re, im) per 16-lane
        # row); the wave-task stencil does not, and falls back to the row blocks where their
        # halos fit (not in the grid order here), else the row pack
        B.set_rhs(prs[:8], om[:8])
        expect = ("stencil_lds",) if layout == "stencil_lds" else ("rowblock", "rowpack")
        assert B.get_launch("fp64")["spmv_layout"] in expect


def test_stencil_batched_solve(O, M, c1_model, c1_mesh):
    """A batch of 16 systems (8 frequencies x 2 polarizations) solved with the stencil SpMV in the
    grid order: every system against the oracle (iterations, x), against the same batch with
    the row-block SpMV in the pencil order (iterations within 2% + 2), and against the
    LDS-staged stencil kernel (iterations within 2% + 2, x within 1e-5)."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import ORDER_GRID, Solver

    say("stencil batched solve, 16 systems")
    freqs = M.frequencies(1e-2, 1e2, 8)
    cases = [(f, a) for f in freqs for a in (1, 0)]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    runs = {}
    for order, layout in ((ORDER_GRID, "stencil"), (ORDER_GRID, "stencil_lds"),
                          (16, "rowblock")):
        with Solver(0) as B:
            B.set_mesh_order(c1_mesh, order)
            B.set_matrix_ks(p, kd, kg, sd, sg)
            B.set_gradient(c1_mesh)
            B.set_launch(spmv_layout=layout)
            runs[layout] = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
    xs, ss = runs["stencil"]
    xr, sr = runs["rowblock"]
    for k in (0, 5, 10, 15):
        f, a = cases[k]
        xo, so, _ = O.pcocr(M.assemble(c1_model, f, a), c1_mesh, eps=1e-8)
        say(f"case {f:.3g} Hz alpha={a}: stencil {ss[k]['iter']} it, oracle {so['iter']} it")
        assert ss[k]["req"] == 0
        assert abs(ss[k]["iter"] - so["iter"]) <= 2 + 0.02 * so["iter"]
        assert rel(xs[k], xo) <= 1e-5
    xl, sl = runs["stencil_lds"]
    for k in range(len(cases)):
        assert ss[k]["req"] == 0 and abs(ss[k]["iter"] - sr[k]["iter"]) <= 2 + 0.02 * sr[k]["iter"]
        # the same products per row; the dot products' partial sums group rows differently
        assert sl[k]["req"] == 0 and abs(sl[k]["iter"] - ss[k]["iter"]) <= 2 + 0.02 * ss[k]["iter"]
        assert rel(xl[k], xs[k]) <= 1e-5


@pytest.mark.parametrize("shape,nfreq", [(None, 8), ((13, 10, 9), 8), (None, 4)])
def test_fused_grid_preconditioner(O, M, c1_model, shape, nfreq, monkeypatch):
    """A batch of 8 or 16 systems on a box mesh in the grid order runs the preconditioner as one
    fused kernel (node and edge passes together, help in LDS, a alternating between two
    buffers). Against the node and edge kernels (MTCOCR_PC_SPLIT=1) on the same batch, whose
    arithmetic per element it shares (only the dot products' partial sums group differently):
    x after three iterations within ten times the split path's own rounding spread (the same
    path with another SpMV layout); solved, FP64 iterations within 2% + 2 and x within 1e-5, mixed
    true residuals below 1e-8; and the oracle on two systems. Config 1's box (25^3 nodes) and a
    14 x 11 x 10-node box: neither is a multiple of the 8 x 4 node tile."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import ORDER_GRID, Solver

    m = c1_model if shape is None else M.random_model(*shape, seed=7, h=100.0, n_pad=2, n_air=3)
    ma = M.mesh_arrays(m)
    say(f"fused grid preconditioner on {m.shape_cells} cells, {2 * nfreq} systems")
    freqs = M.frequencies(1e-2, 1e2, nfreq)
    cases = [(f, a) for f in freqs for a in (1, 0)]
    p, kd, kg, sd, sg = M.assemble_ks(m)
    prs = np.stack([M.assemble_rhs(m, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    runs = {}
    for split in ("0", "1"):
        monkeypatch.setenv("MTCOCR_PC_SPLIT", split)
        with Solver(0) as B:
            B.set_mesh_order(ma, ORDER_GRID)
            B.set_matrix_ks(p, kd, kg, sd, sg)
            B.set_gradient(ma)
            for prec in ("fp64", "mixed"):
                runs[split, prec] = B.solve_batch(om, prs, eps=1e-8, precision=prec)
            runs[split, "3 iterations"] = B.solve_batch(om, prs, eps=1e-8, maxiter=3)
            if split == "1":
                # the same path with another SpMV kernel: the partial sums of (z, s) group
                # differently, which sets the rounding spread of x after 3 iterations
                cur = B.get_launch()["spmv_layout"]
                B.set_launch(spmv_layout="ell" if cur != "ell" else "rowpack")
                runs["1b", "3 iterations"] = B.solve_batch(om, prs, eps=1e-8, maxiter=3)
    # after three iterations x depends on w through s = s - alpha w: a wrong w would move it by
    # O(1); the fused path stays within the rounding spread of the split path's own kernels
    # (the only difference is the grouping of the (a, w) partial sums)
    (x3f, s3f), (x3s, s3s) = runs["0", "3 iterations"], runs["1", "3 iterations"]
    x3b, s3b = runs["1b", "3 iterations"]
    d3 = max(rel(x3f[k], x3s[k]) for k in range(len(cases)))
    dref = max(rel(x3b[k], x3s[k]) for k in range(len(cases)))
    say(f"  after 3 iterations: x max rel diff {d3:.2e} (split path, two SpMV layouts: "
        f"{dref:.2e})")
    assert [q["iter"] for q in s3f] == [q["iter"] for q in s3s] == [q["iter"] for q in s3b]
    assert d3 <= max(10 * dref, 1e-12)
    for prec in ("fp64", "mixed"):
        (xf, sf), (xs, ss) = runs["0", prec], runs["1", prec]
        d = [rel(xf[k], xs[k]) for k in range(len(cases))]
        say(f"  {prec}: fused iterations {[q['iter'] for q in sf]}, split "
            f"{[q['iter'] for q in ss]}, x max rel diff {max(d):.2e}")
        for k in range(len(cases)):
            assert sf[k]["req"] == 0 and ss[k]["req"] == 0
            if prec == "fp64":
                assert abs(sf[k]["iter"] - ss[k]["iter"]) <= 2 + 0.02 * ss[k]["iter"]
                assert d[k] <= 1e-5, (prec, k, d[k])
            else:
                # the mixed solver's passes and finish follow its stall tests, which rounding
                # moves, and its answer is held by its FP64 true residual
                assert sf[k]["true_rel_res"] < 1e-8 and ss[k]["true_rel_res"] < 1e-8
    # against the oracle: the small box is far less well conditioned (its rounding spread after
    # three iterations is 1e5 times config 1's), so its x is held looser
    xf, sf = runs["0", "fp64"]
    for k in (0, len(cases) - 1):
        f, a = cases[k]
        xo, so, _ = O.pcocr(M.assemble(m, f, a), ma, eps=1e-8)
        assert abs(sf[k]["iter"] - so["iter"]) <= 2 + 0.02 * so["iter"]
        assert rel(xf[k], xo) <= (1e-5 if shape is None else 1e-4)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("mixed", 2e-5)])
def test_preconditioner_parity(S, O, M, c1_system, c1_mesh, prec, tol):
    say(f"preconditioner {prec}")
    d = O.Dfp(c1_mesh, c1_system)
    x = np.random.default_rng(12).standard_normal(2 * c1_system.nb)
    ref = d.apply(x)
    got = S.apply("precond", x, prec)
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


def test_fp64_solve_parity(S, O, M, c1_system, oracle_c1):
    say("fp64 solve")
    xo, so = oracle_c1
    x, st = S.solve(c1_system.pr, eps=1e-8, maxiter=10000, precision="fp64")
    say(f"gpu {st} oracle iter {so['iter']}")
    assert st["req"] == 0 and so["req"] == 0
    assert abs(st["iter"] - so["iter"]) <= 2 + 0.02 * so["iter"]
    assert st["residual_rel"] < 1e-8
    A = M.sparse_matrix(c1_system)
    b = M.to_complex(c1_system.pr)
    true_rel = np.linalg.norm(b - A @ M.to_complex(x)) / np.linalg.norm(b)
    assert true_rel <= 3e-8
    assert st["true_rel_res"] == pytest.approx(true_rel, rel=1e-3)
    assert rel(x, xo) <= 1e-5


def test_fp64_deterministic(S, c1_system):
    say("determinism")
    x1, s1 = S.solve(c1_system.pr, eps=1e-8, precision="fp64")
    x2, s2 = S.solve(c1_system.pr, eps=1e-8, precision="fp64")
    assert s1["iter"] == s2["iter"] and np.array_equal(x1, x2)


def test_mixed_solve(S, O, M, c1_system, oracle_c1):
    say("mixed solve")
    xo, so = oracle_c1
    x, st = S.solve(c1_system.pr, eps=1e-8, maxiter=10000, precision="mixed")
    say(f"mixed {st}")
    assert st["req"] == 0 and st["true_rel_res"] < 1e-8
    A = M.sparse_matrix(c1_system)
    b = M.to_complex(c1_system.pr)
    true_rel = np.linalg.norm(b - A @ M.to_complex(x)) / np.linalg.norm(b)
    assert true_rel < 1e-8
    assert rel(x, xo) <= 1e-5


def test_batched_ks_matches_single(O, M, c1_model, c1_mesh):
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say("batched K/S: 2 frequencies x 2 polarizations")
    freqs = [0.1, 10.0]
    cases = [(f, a) for f in freqs for a in (1, 0)]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    with Solver(0) as B:
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(c1_mesh)
        xb, sb = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
        xm, sm = B.solve_batch(om, prs, eps=1e-8, precision="mixed")
    for k, (f, a) in enumerate(cases):
        s = M.assemble(c1_model, f, a)
        xo, so, _ = O.pcocr(s, c1_mesh, eps=1e-8)
        say(f"case {f} Hz alpha={a}: batch {sb[k]['iter']} it, oracle {so['iter']} it, "
            f"mixed {sm[k]['iter']} it / {sm[k]['outer']} passes")
        assert sb[k]["req"] == 0 and sm[k]["req"] == 0
        assert abs(sb[k]["iter"] - so["iter"]) <= 2 + 0.02 * so["iter"]
        assert rel(xb[k], xo) <= 1e-5
        A = M.sparse_matrix(s)
        b = M.to_complex(s.pr)
        assert np.linalg.norm(b - A @ M.to_complex(xm[k])) / np.linalg.norm(b) < 1e-8


def test_wide_batch_runs_in_groups(O, M, c1_model, c1_mesh):
    """20 systems (batch width 32): the FP64 solver iterates them as a group of 16 and a group
    of 4. The first group's answers are bit-identical to a batch of those 16 alone; every
    system converges, the last ones agree with the oracle; the mixed solver converges too."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say("wide batch: 20 systems in groups of 16")
    freqs = M.frequencies(1e-2, 1e2, 10)
    cases = [(f, a) for f in freqs for a in (1, 0)]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    with Solver(0) as B:
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(c1_mesh)
        xw, sw = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
        x16, s16 = B.solve_batch(om[:16], prs[:16], eps=1e-8, precision="fp64")
        xm, sm = B.solve_batch(om, prs, eps=1e-8, precision="mixed")
    assert all(q["req"] == 0 for q in sw) and all(q["req"] == 0 for q in sm)
    assert np.array_equal(xw[:16], x16)
    assert [q["iter"] for q in sw[:16]] == [q["iter"] for q in s16]
    for k in (16, 19):
        f, a = cases[k]
        xo, so, _ = O.pcocr(M.assemble(c1_model, f, a), c1_mesh, eps=1e-8)
        say(f"case {f:.3g} Hz alpha={a}: group {sw[k]['iter']} it, oracle {so['iter']} it")
        assert abs(sw[k]["iter"] - so["iter"]) <= 2 + 0.02 * so["iter"]
        assert rel(xw[k], xo) <= 1e-5
        s = M.assemble(c1_model, f, a)
        res = M.to_complex(s.pr) - M.sparse_matrix(s) @ M.to_complex(xm[k])
        assert np.linalg.norm(res) / np.linalg.norm(M.to_complex(s.pr)) < 1e-8


def test_special_cases(S, M, c1_system, c1_mesh):
    say("special cases")
    x, st = S.solve(np.zeros_like(c1_system.pr), precision="fp64")
    assert st["req"] == -1 and not x.any()
    x, st = S.solve(c1_system.pr, maxiter=1, precision="fp64")
    assert st["req"] == -3 and st["iter"] == 2
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    m, ma, s = identity_system(M)
    with Solver(0) as T:
        T.set_matrix(s)
        T.set_gradient(ma)
        x, st = T.solve(s.pr, precision="fp64")
        assert st["req"] == -2 and st["iter"] == 1


@pytest.mark.parametrize("period", [8, 16])
def test_iteration_limit_at_every_xp_position(O, M, c1_model, c1_mesh, period):
    """The FP64 solver applies x and p `period` iterations at a time; a system that halts at
    its iteration limit returns x of the last iteration whatever its position in that window.
    A batch of 16 systems (stencil SpMV, fused preconditioner) halted at maxiter = 57 ..
    57 + period: each checked system's x against the oracle's x at the same limit, closer to it
    than to the oracle's x one iteration before or after by a factor of 100."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import ORDER_GRID, Solver

    freqs = M.frequencies(1e-2, 1e2, 8)
    cases = [(f, a) for f in freqs for a in (1, 0)]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    limits = tuple(range(57, 58 + period))
    gpu = {}
    with Solver(0) as B:
        B.set_mesh_order(c1_mesh, ORDER_GRID)
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(c1_mesh)
        B.set_xp_period(period)
        for mi in limits:
            gpu[mi] = B.solve_batch(om, prs, eps=1e-8, maxiter=mi, precision="fp64")
        assert B.get_xp_period() == period
    for k in (0, 5, 10, 15):
        f, a = cases[k]
        s = M.assemble(c1_model, f, a)
        xo = {mi: O.pcocr(s, c1_mesh, eps=1e-8, maxiter=mi)[0]
              for mi in range(limits[0] - 1, limits[-1] + 2)}
        for mi in limits:
            x, st = gpu[mi]
            assert st[k]["req"] == -3 and st[k]["iter"] == mi + 1
            d = rel(x[k], xo[mi])
            gap = min(rel(xo[mi - 1], xo[mi]), rel(xo[mi + 1], xo[mi]))
            say(f"case {k} maxiter {mi}: x vs oracle {d:.2e}, one iteration apart {gap:.2e}")
            assert d < 0.01 * gap


def test_graph_replay_matches_direct_launches(O, M, c1_model, c1_mesh, monkeypatch):
    """The FP64 solver's 16-iteration chunks replayed as HIP graphs give the bits of the same
    kernels launched directly (MTCOCR_NO_GRAPH=1): a batch of 16 systems in the grid order
    (stencil SpMV, fused preconditioner, w formed in the update pass, x and p every eighth
    iteration, compaction to 8, 4, 2, 1), solved twice each way on one handle."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import ORDER_GRID, Solver

    freqs = M.frequencies(1e-2, 1e2, 8)
    cases = [(f, a) for f in freqs for a in (1, 0)]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    runs = []
    with Solver(0) as B:
        B.set_mesh_order(c1_mesh, ORDER_GRID)
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(c1_mesh)
        for direct in (False, True, False, True):
            if direct:
                monkeypatch.setenv("MTCOCR_NO_GRAPH", "1")
            else:
                monkeypatch.delenv("MTCOCR_NO_GRAPH", raising=False)
            x, st = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
            runs.append((np.asarray(x), [q["iter"] for q in st], [q["req"] for q in st]))
            say(f"{'direct' if direct else 'graphs'}: iters {runs[-1][1]}")
    x0, it0, rq0 = runs[0]
    assert all(r == 0 for r in rq0)
    for x, it, rq in runs[1:]:
        assert it == it0 and rq == rq0
        assert np.array_equal(x.view(np.uint64), x0.view(np.uint64))


def test_pcocr_2x2_folded_mirror(S, M, c1_system, c1_mesh):
    """The reference-signature entry point gives the same x as the handle API (deterministic)."""
    say("pcocr_2x2_folded mirror")
    from mtcalc_with_dfp_cocr_amd.cocr import pcocr_2x2_folded

    s, ma = c1_system, c1_mesh
    x = np.zeros(2 * s.nb)
    info = {}
    rc = pcocr_2x2_folded(2 * s.nb, s.ig, s.jg, s.idi, s.ijg, s.di, s.gg, s.pr, x, 1e-8, 10000,
                          ma.kpar, s.nb, ma.kuzlov, ma.xyz, ma.nvkat, ma.nver, ma.sig_flags,
                          ma.edges, is_node_bound=ma.is_node_bound, stats=info)
    from mtcalc_with_dfp_cocr_amd.cocr import ORDER_GRID, Solver

    with Solver(0) as H:  # the one-shot call holds the grid order of a box mesh
        H.set_mesh_order(ma, ORDER_GRID)
        H.set_matrix(s)
        H.set_gradient(ma)
        xh, st = H.solve(s.pr, eps=1e-8, precision="fp64")
    assert rc == 0 and info["iter"] == st["iter"] and np.array_equal(x, xh)


EXE = os.path.join(os.path.dirname(__file__), "..", "mtcalc_with_dfp_cocr_amd", "lib", "COCR_FP")


def test_dropin_executable(O, M, c1_model, c1_system, c1_mesh, tmp_path):
    """COCR_FP in a working directory: same files in, v3.dat / kit out, exit 0 on
    non-convergence, exit 2 with --strict."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    say("drop-in executable")
    d = M.write_cocr_fp_dir(tmp_path / "Nu1_X", c1_model, c1_system, mesh=c1_mesh)
    # capitalised names as RegularMeshBuilder writes them
    os.rename(d / "sig3d", d / "Sig3d")
    os.rename(d / "l13d.dat", d / "L13d.dat")
    os.rename(d / "mu3d", d / "mu3D")
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    v3 = np.fromfile(d / "v3.dat", np.float64)
    kit = (d / "kit").read_text()
    assert kit.startswith("r=") and "\teps=1.000000e-08\tkit=" in kit and kit.rstrip().endswith("sec")
    xo, so, _ = O.pcocr(c1_system, c1_mesh, eps=1e-8)
    assert len(v3) == 2 * c1_system.nb and rel(v3, xo) <= 1e-5
    # non-convergence: exit 0 like the reference, 2 with --strict
    (d / "kuslau").write_text(f"{2 * c1_system.nb}\n1.000000e-08\n3\n")
    assert subprocess.run([EXE], cwd=d, capture_output=True, timeout=120).returncode == 0
    assert subprocess.run([EXE, "--strict"], cwd=d, capture_output=True, timeout=120).returncode == 2
    # missing input: error exit
    os.remove(d / "gg")
    assert subprocess.run([EXE], cwd=d, capture_output=True, timeout=120).returncode == 1


def test_dropin_batch_executable(O, M, c1_model, c1_mesh, tmp_path):
    """--batch over three directories of one mesh: every v3.dat against the oracle; each kit
    holds the system's share of the batch time (batch time / 3, printed by the process); a
    directory whose off-diagonal blocks are not K + i omega S of the first, or whose kuslau
    tolerance differs, is refused (exit 1)."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    say("drop-in executable --batch")
    dirs = []
    for f, a in [(0.1, 1), (0.1, 0), (10.0, 1)]:
        s = M.assemble(c1_model, f, a)
        dirs.append(M.write_cocr_fp_dir(tmp_path / f"Nu{f}_{a}", c1_model, s, mesh=c1_mesh))
    r = subprocess.run([EXE, "--batch", *map(str, dirs)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    batch_s = float(r.stdout.split("batch time=")[1].split()[0])
    for d in dirs:
        s = M.read_block_system(d)
        xo, so, _ = O.pcocr(s, c1_mesh, eps=1e-8)
        assert rel(np.fromfile(d / "v3.dat", np.float64), xo) <= 1e-5
        kit_s = float((d / "kit").read_text().split("time=")[1].split()[0])
        assert kit_s == pytest.approx(batch_s / 3, abs=2e-6)
    # one off-diagonal block of the last directory changed: not the first one's K + i omega S
    gg = np.fromfile(dirs[2] / "gg", np.float64)
    gg[len(gg) // 2] *= 1.5
    gg.tofile(dirs[2] / "gg")
    r = subprocess.run([EXE, "--batch", *map(str, dirs)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1 and "K + i*omega*S" in r.stderr
    # another tolerance in kuslau
    (dirs[1] / "kuslau").write_text(f"{2 * s.nb}\n1.000000e-07\n10000\n")
    r = subprocess.run([EXE, "--batch", str(dirs[0]), str(dirs[1])], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 1 and "eps/maxiter" in r.stderr


def test_gradient_and_matrix_sizes_must_agree(M, c1_system, c1_mesh, tiny_model):
    """Call order: a gradient set for one mesh and then a matrix of another size leave the
    handle without a usable gradient (solve and preconditioner apply return MTCOCR_ERR_STATE
    before any launch); a gradient whose edge count differs from the matrix is refused
    (MTCOCR_ERR_ARG); setting the matching gradient again restores the solver."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say("gradient / matrix size checks")
    st = M.assemble(tiny_model, 1.0, 1)
    mt = M.mesh_arrays(tiny_model)
    assert st.nb != c1_system.nb
    with Solver(0) as H:
        H.set_matrix(c1_system)
        H.set_gradient(c1_mesh)
        H.set_matrix(st)
        with pytest.raises(RuntimeError, match=r"\(-13\)"):
            H.solve(st.pr, precision="fp64")
        with pytest.raises(RuntimeError, match=r"\(-13\)"):
            H.apply("precond", st.pr)
        with pytest.raises(RuntimeError, match=r"\(-10\)"):
            H.set_gradient(c1_mesh)
        H.set_gradient(mt)
        x, s1 = H.solve(st.pr, precision="fp64")
        assert s1["req"] == 0


def test_hanging_node_preconditioner(O, M):
    """The folded preconditioner on a mesh with hanging nodes (a 2:1 refined face, T-matrix
    rows of 2 and 4 nodes): rows of G longer than a conforming mesh's 6 entries reach the
    device's D_N set-up and the node and edge kernels. Against the oracle's BuildGMatrix,
    BuildDiagComplex and ApplyPreconditionerComplex on the same arrays, with a complex-
    symmetric matrix that couples every two edges of a node."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    import ctypes as C

    from test_gradient import hanging_mesh

    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say("hanging-node preconditioner")
    m = M.random_model(4, 3, 3, seed=2, h=100.0, n_pad=1, n_air=2, name="g")
    ma = M.mesh_arrays(m)
    edges, nc, ig_t, jg_t, gg_t = hanging_mesh(M, ma)
    ntot = len(ig_t) - 1
    nb = len(edges)
    bound = np.concatenate([ma.is_node_bound, np.zeros(ntot - nc, np.int32)]).astype(np.int32)
    # lower-triangle blocks: every pair of edges that share a node (hanging nodes through
    # their T rows), complex
    rng = np.random.default_rng(8)
    node_edges = {}
    for e, (u, v) in enumerate(edges):
        for w in (u, v):
            ks = [w] if w < nc else list(jg_t[ig_t[w]:ig_t[w + 1]])
            for k in ks:
                node_edges.setdefault(int(k), set()).add(e)
    lower = {i: set() for i in range(nb)}
    for es in node_edges.values():
        for a in es:
            for b in es:
                if b < a:
                    lower[a].add(b)
    ig = np.zeros(nb + 1, np.int32)
    jg = []
    for i in range(nb):
        cols = sorted(lower[i])
        jg += cols
        ig[i + 1] = ig[i] + len(cols)
    jg = np.array(jg, np.int32)
    ijg = np.arange(0, 2 * len(jg) + 1, 2, dtype=np.int32)
    gg = 0.1 * rng.standard_normal(2 * len(jg))
    idi = np.arange(0, 2 * nb + 1, 2, dtype=np.int32)
    di = np.stack([4.0 + rng.random(nb), rng.random(nb)], 1).ravel()
    s = M.BlockSystem(nb, ig, jg, idi, ijg, di, gg, np.zeros(2 * nb))
    x = rng.standard_normal(2 * nb)
    # oracle: BuildGMatrix with the T-matrix, the diagonals, one application
    L = O.lib()
    p = O.OrcDfp()
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    assert L.orc_build_g(C.byref(p), int(ma.kpar), nb, nc, ma.nvkat.astype(np.int32),
                         ma.nver.astype(np.int32), ma.sig_flags.astype(np.float64),
                         edges.ravel().astype(np.int32), vp(ig_t), vp(jg_t), vp(gg_t), bound) == 0
    assert L.orc_build_diag(C.byref(p), ig, jg, idi, ijg, di, gg) == 0
    ref = np.empty(2 * nb)
    L.orc_apply_precond(C.byref(p), x, ref, 1)
    L.orc_dfp_free(C.byref(p))
    with Solver(0) as H:
        H.set_matrix(s)
        rc = H._L.mtcocr_set_gradient_mesh(H._h, int(ma.kpar), nb, nc, ma.nvkat.astype(np.int32),
                                           ma.nver.astype(np.int32),
                                           ma.sig_flags.astype(np.float64),
                                           edges.ravel().astype(np.int32), vp(ig_t), vp(jg_t),
                                           vp(gg_t), bound)
        assert rc == 0
        got = H.apply("precond", x, "fp64")
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


def test_launch_shapes_and_autotune(O, M, c1_model, c1_mesh):
    """Every launch shape gives the same solution to rounding (the caps regroup only the dot
    products) and a deterministic one for a fixed shape; the autotuner keeps a candidate cap."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say("launch shapes + autotune: 3 systems")
    cases = [(0.1, 1), (0.1, 0), (10.0, 1)]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    cands = {512, 1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072}
    with Solver(0) as B:
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(c1_mesh)
        defaults = {"cap_spmv": 8192, "cap_vec": 65536, "cap_node": 2048}
        assert B.get_launch() == dict(defaults, spmv_layout="ell")  # no batch yet: width 1
        x0, s0 = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
        assert B.get_launch() == dict(defaults, spmv_layout="rowblock")  # 3 systems: width 4
        for caps in [(8, 8, 8, "rowpack"), (1024, 512, 16384, "ell"),
                     (16384, 16384, 1024, "rowpack"), (2048, 0, 0, "rowblock")]:
            B.set_launch(*caps)
            x1, s1 = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
            x2, s2 = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
            say(f"caps {caps}: iters {[q['iter'] for q in s1]} vs {[q['iter'] for q in s0]}")
            assert np.array_equal(x1, x2)
            for k in range(len(cases)):
                assert s1[k]["req"] == 0
                assert abs(s1[k]["iter"] - s0[k]["iter"]) <= 2 + 0.02 * s0[k]["iter"]
                assert rel(x1[k], x0[k]) <= 1e-5
        with pytest.raises(RuntimeError):
            B.set_launch(12, 0, 0)  # not a multiple of the XCD count
        B.set_rhs(prs, om)
        for prec in ("fp64", "mixed"):
            t = B.autotune(prec, reps=2)
            say(f"autotune {prec}: {t}")
            # a fixed candidate or a kernel's resident capacity (workgroups per CU x CUs)
            for c in (t["cap_spmv"], t["cap_vec"], t["cap_node"]):
                assert c in cands or (c % 8 == 0 and 8 <= c <= 131072)
            assert t["spmv_ms"] > 0
            assert B.get_launch(prec) == {k: t[k] for k in ("cap_spmv", "cap_vec", "cap_node",
                                                            "spmv_layout")}
        # the resident right-hand sides survive the tuning
        rc, st = B.solve_resident(1e-8, 10000, "fp64")
        xr = B.get_solution()
        for k in range(len(cases)):
            assert st[k]["req"] == 0 and rel(xr[k], x0[k]) <= 1e-5


def test_bench_hooks(M, c1_model, c1_mesh):
    """Kernel timing hooks on a batch of 5 systems (width 8): SpMV, the fused update pass and
    the FP64 x/p pass report positive times and their algorithmic bytes (the FP64 update pass
    moves 8 vectors, the FP32-vector one 14, the x/p pass 12 or 20 at periods of 8 and 16); the
    resident batch is left
    usable."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say("bench hooks: 5 systems")
    cases = [(f, a) for f in (0.1, 1.0, 10.0) for a in (1, 0)][:5]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    with Solver(0) as B:
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(c1_mesh)
        B.set_rhs(prs, om)
        for prec, streams, elem in (("fp64", 8, 16), ("mixed", 14, 8)):
            ms, by = B.bench_update(prec, 3)
            assert ms > 0 and by == streams * elem * p.nb * len(cases)
            ms, by = B.bench_spmv(prec, 3)
            assert ms > 0 and by > 0
        B.set_xp_period(8)
        ms, by = B.bench_flush(3)
        assert B.get_xp_period() == 8 and ms > 0 and by == 12 * 16 * p.nb * len(cases)
        B.set_xp_period(16)
        ms, by = B.bench_flush(3)
        assert B.get_xp_period() == 16 and ms > 0 and by == 20 * 16 * p.nb * len(cases)
        B.set_rhs(prs, om)
        rc, st = B.solve_resident(1e-8, 10000, "fp64")
        assert all(q["req"] == 0 for q in st)


def test_bench_hooks_grid_order(M, c1_model, c1_mesh):
    """The same hooks on a box mesh in the grid order (width 8, fused grid preconditioner): the
    FP64 update pass forms w from the node vector help, so it reports 7 vector streams plus
    help once per interior node and system and the diagonal once per edge; the x/p pass 12
    streams. The batch then solves as usual."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import ORDER_GRID, Solver

    say("bench hooks, grid order: 5 systems")
    cases = [(f, a) for f in (0.1, 1.0, 10.0) for a in (1, 0)][:5]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    nx, ny, nz = c1_model.shape_nodes
    m = (nx - 2) * (ny - 2) * (nz - 2)
    with Solver(0) as B:
        B.set_mesh_order(c1_mesh, ORDER_GRID)
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(c1_mesh)
        B.set_rhs(prs, om)
        ms, by = B.bench_update("fp64", 3)
        assert ms > 0 and by == 7 * 16 * p.nb * len(cases) + 16 * m * len(cases) + 16 * p.nb
        ms, by = B.bench_flush(3)
        xp = B.get_xp_period()
        assert xp in (8, 16) and ms > 0 and by == (xp + 4) * 16 * p.nb * len(cases)
        B.set_rhs(prs, om)
        rc, st = B.solve_resident(1e-8, 10000, "fp64")
        assert all(q["req"] == 0 for q in st)


def test_wide_grid_two_stage_reduction(O, M, c1_model, c1_mesh):
    """16 systems on 45000 edges: at 16384 workgroups the dot-product partials (more than
    16384 doubles) are folded in two stages, at 256 they are reduced in one. Same solution to
    rounding either way, deterministic, and the FP64 oracle's iteration counts on two of the
    systems."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say("two-stage partial reduction: 16 systems")
    freqs = M.frequencies(1e-2, 1e2, 8)
    cases = [(f, a) for f in freqs for a in (1, 0)]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    with Solver(0) as B:
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(c1_mesh)
        B.set_launch(256, 256, 256, "rowpack")
        x1, s1 = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
        B.set_launch(16384, 16384, 16384, "rowpack")
        x2, s2 = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
        x3, s3 = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
    assert np.array_equal(x2, x3)
    for k in range(len(cases)):
        assert s1[k]["req"] == 0 and s2[k]["req"] == 0
        assert abs(s1[k]["iter"] - s2[k]["iter"]) <= 2 + 0.02 * s1[k]["iter"]
        assert rel(x2[k], x1[k]) <= 1e-5
    for k in (0, 15):
        f, a = cases[k]
        xo, so, _ = O.pcocr(M.assemble(c1_model, f, a), c1_mesh, eps=1e-8)
        say(f"case {f:.3g} Hz alpha={a}: wide {s2[k]['iter']} it, oracle {so['iter']} it")
        assert abs(s2[k]["iter"] - so["iter"]) <= 2 + 0.02 * so["iter"]
        assert rel(x2[k], xo) <= 1e-5


@pytest.mark.parametrize("block", [4, 16])
def test_mesh_order_parity(O, M, c1_system, c1_mesh, oracle_c1, block):
    """Matrix, gradient and vectors held in the pencil order: SpMV and preconditioner at the
    same tolerances, and the FP64 solve against the oracle, inputs and outputs in the caller's
    order."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say(f"mesh order block {block}")
    xo, so = oracle_c1
    x = np.random.default_rng(14).standard_normal(2 * c1_system.nb)
    with Solver(0) as B:
        B.set_mesh_order(c1_mesh, block)
        B.set_matrix(c1_system)
        B.set_gradient(c1_mesh)
        ref = O.spmv(c1_system, x)
        got = B.apply("spmv", x, "fp64")
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
        ref = O.Dfp(c1_mesh, c1_system).apply(x)
        got = B.apply("precond", x, "fp64")
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()
        xs, st = B.solve(c1_system.pr, eps=1e-8, maxiter=10000, precision="fp64")
        say(f"ordered solve {st['iter']} it, oracle {so['iter']} it")
        assert st["req"] == 0 and abs(st["iter"] - so["iter"]) <= 2 + 0.02 * so["iter"]
        assert rel(xs, xo) <= 1e-5
        with pytest.raises(RuntimeError):
            B.set_mesh_order(c1_mesh, block)  # after the matrix: refused


def test_adjoint_gradient_gpu(M, small_model):
    """Adjoint gradient with the GPU batched solver (forward and adjoint systems of 2 frequencies
    x 2 polarizations each in one batch) against the same gradient from exact sparse LU
    solves."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from helpers import DirectKSSolver

    from mtcalc_with_dfp_cocr_amd.adjoint import GpuBatchSolver, ImpedanceMisfit

    say("adjoint gradient on the GPU")
    m = small_model
    rng = np.random.default_rng(21)
    rec = np.stack([rng.uniform(m.x[3], m.x[-4], 5), rng.uniform(m.y[3], m.y[-4], 5),
                    np.zeros(5)], 1)
    freqs = [0.1, 3.0]
    ref = ImpedanceMisfit(m, freqs, rec, solver=DirectKSSolver(M, m))
    earth = m.sigma > 1e-6
    sig_true = m.sigma.copy()
    sig_true[earth] *= np.exp(0.3 * rng.standard_normal(int(earth.sum())))
    _, zobs, _ = ref.responses(sig_true)
    ref.z_obs = zobs
    st = ref.evaluate(m.sigma)
    g_ref = ref.gradient(st)
    gpu = GpuBatchSolver(m, eps=1e-12)
    try:
        prob = ImpedanceMisfit(m, freqs, rec, z_obs=zobs, solver=gpu)
        st2 = prob.evaluate(m.sigma)
        g = prob.gradient(st2)
        # the torch.autograd wrapper over the same GPU problem: J forward, the adjoint backward
        import torch

        from mtcalc_with_dfp_cocr_amd.adjoint import torch_misfit

        ls = torch.tensor(np.log(m.sigma), requires_grad=True)
        J = torch_misfit(prob)(ls)
        J.backward()
        gt = ls.grad.numpy()
    finally:
        gpu.close()
    assert float(J.detach()) == pytest.approx(st2.J, rel=1e-8)
    np.testing.assert_allclose(gt, g * m.sigma, rtol=1e-6, atol=1e-8 * np.abs(g * m.sigma).max())
    say(f"J gpu {st2.J:.12e} direct {st.J:.12e}")
    assert st2.J == pytest.approx(st.J, rel=1e-8)
    # in the inversion variable log(sigma): dJ/dlog(sigma) = sigma dJ/dsigma (dJ/dsigma itself
    # is dominated by the 1e-8 S/m air cells, where it is large and irrelevant)
    gl, gl_ref = g * m.sigma, g_ref * m.sigma
    err = np.linalg.norm(gl - gl_ref) / np.linalg.norm(gl_ref)
    say(f"gradient rel err (log sigma) {err:.3e}")
    assert err <= 1e-6


def test_sigma_gradient_contraction(M, small_model):
    """The device contraction sum_s Re(coef_s u_s^T M_c v_s) against the host's per-cell mass
    forms (the element mass matrices of the model layer), on random vectors; the host sums in
    the same order without FMA, so the two agree to rounding."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.adjoint import cell_mass_forms
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say("sigma gradient contraction")
    m = small_model
    rng = np.random.default_rng(5)
    nsys, nb = 3, m.n_edges
    u = rng.standard_normal((nsys, nb)) + 1j * rng.standard_normal((nsys, nb))
    v = rng.standard_normal((nsys, nb)) + 1j * rng.standard_normal((nsys, nb))
    coef = rng.standard_normal(nsys) + 1j * rng.standard_normal(nsys)
    ref = np.zeros(m.n_cells)
    for s in range(nsys):
        ref += np.real(coef[s] * cell_mass_forms(m, u[s], v[s]))
    with Solver(0) as H:  # needs no matrix
        g = H.sigma_gradient(m.shape_nodes, m.x, m.y, m.z, coef, u, v)
        g1 = H.sigma_gradient(m.shape_nodes, m.x, m.y, m.z, coef[:1], u[:1], v[:1])
    ref1 = np.real(coef[0] * cell_mass_forms(m, u[0], v[0]))
    say(f"max |gpu - host| {np.abs(g - ref).max():.3e} of {np.abs(ref).max():.3e}")
    assert np.abs(g - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.abs(g1 - ref1).max() <= 1e-13 * np.abs(ref1).max()


def test_impedance_parity_with_oracle(O, M, c1_model, c1_mesh):
    """Z (and rho_a, phase) at the 25 config-1 receivers from the GPU solutions of both
    polarizations against Z from the oracle's solutions: within 1e-6 relative."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import Solver
    from mtcalc_with_dfp_cocr_amd.receivers import (ReceiverOperator, apparent_resistivity,
                                                    impedance)

    say("impedance parity, config 1, 1 Hz")
    nu = 1.0
    R = ReceiverOperator(c1_model, c1_model.receivers)
    sols = {}
    with Solver(0) as S:
        S.set_mesh_order(c1_mesh, 16)
        for alpha in (1, 0):
            s = M.assemble(c1_model, nu, alpha)
            if alpha == 1:
                S.set_matrix(s)
                S.set_gradient(c1_mesh)
            xg, st = S.solve(s.pr, eps=1e-8, precision="fp64")
            xo, so, _ = O.pcocr(s, c1_mesh, eps=1e-8)
            assert st["req"] == 0 and so["req"] == 0
            sols[alpha] = (M.to_complex(xg), M.to_complex(xo))
    Zs = []
    for k in (0, 1):
        Ex, Hx = R.fields(sols[1][k], nu, 1)
        Ey, Hy = R.fields(sols[0][k], nu, 0)
        Zs.append(impedance(Ex, Hx, Ey, Hy))
    Zg, Zo = Zs
    off = [(0, 1), (1, 0)]
    for i, j in off:
        err = np.abs(Zg[:, i, j] - Zo[:, i, j]) / np.abs(Zo[:, i, j])
        say(f"Z{'xy' if i == 0 else 'yx'} max rel diff {err.max():.2e}")
        assert err.max() <= 1e-6
    rg, pg = apparent_resistivity(Zg, nu)
    ro, po = apparent_resistivity(Zo, nu)
    for i, j in off:
        np.testing.assert_allclose(rg[:, i, j], ro[:, i, j], rtol=2e-6)
        np.testing.assert_allclose(pg[:, i, j], po[:, i, j], atol=1e-4)
    # the tipper from the same solutions; the block is centred under the receiver grid, so by
    # symmetry T vanishes at the centre receiver (index 12 of the 5x5 grid)
    from mtcalc_with_dfp_cocr_amd.receivers import tipper

    Ts = []
    for k in (0, 1):
        _, Hx = R.fields(sols[1][k], nu, 1)
        _, Hy = R.fields(sols[0][k], nu, 0)
        Ts.append(tipper(R.hz(sols[1][k]), Hx, R.hz(sols[0][k]), Hy))
    Tg, To = Ts
    scale = np.abs(To).max()
    say(f"tipper: max |T| {scale:.3e}, GPU vs oracle {np.abs(Tg - To).max() / scale:.2e}")
    assert np.abs(Tg - To).max() <= 1e-6 * scale
    assert np.abs(To[12]).max() <= 1e-4 * scale


def test_launcher_end_to_end(M, c1_model, tmp_path):
    """CalcStarter-equivalent run on config 1 at two frequencies: the working directories, one
    batched COCR_FP launch, the impedance table; the same Z (to 1e-6) as one solver process
    per directory, the reference's way."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd import launcher as L

    say("launcher: config 1, 2 frequencies")
    freqs = [0.3, 3.0]
    rec = np.array([[x, y, 0.0] for x in (-100.0, 0.0, 100.0) for y in (-100.0, 100.0)])
    za = tmp_path / "batched"
    zb = tmp_path / "per_directory"
    assert L.run(L.prepare(za, c1_model, freqs), batch=True, timeout=300) == 0
    Za = L.collect(za, c1_model, freqs, rec)
    assert L.run(L.prepare(zb, c1_model, freqs), batch=False, timeout=300) == 0
    Zb = L.collect(zb, c1_model, freqs, rec)
    lines = (za / "Results" / "impedance.txt").read_text().splitlines()
    assert len(lines) == len(freqs) * len(rec) and len(lines[0].split("\t")) == 15
    for i, j in ((0, 1), (1, 0)):
        d = np.abs(Za[:, :, i, j] - Zb[:, :, i, j]) / np.abs(Zb[:, :, i, j])
        assert d.max() <= 1e-6


def test_full_size_parity(M):
    """The benchmark's mesh at its full size (100x100x60 cells, 1.84 M edges), all 16 systems
    (8 frequencies x 2 polarizations) as one FP64 batch on the benchmark's path (grid order,
    autotuned launch shapes: the LDS-staged stencil SpMV), against the oracle's own run of every
    system to its stop test (tests/golden/config2_oracle.json):
      * the stop test met (MRS residual < 1e-8) and iterations within 2% + 2 of the oracle's;
      * the true relative residual ||b - A x|| / ||b|| (FP64 host product, A = K + i omega S)
        at most FP64_TRUE_BOUND (conftest: the rounding-path spread of the raw iterate at the
        stop point, the oracle's own included);
      * Z at the 5x5 grid within 1e-6 of the oracle's Z (Zxx and Zyy relative to |Zxy|).
    The highest frequency is also solved in mixed precision, whose refinement tests the true
    residual: held to 1e-8."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    import json
    from pathlib import Path

    import scipy.sparse as sp

    from conftest import FP64_TRUE_BOUND
    from mtcalc_with_dfp_cocr_amd.cocr import ORDER_GRID, Solver
    from mtcalc_with_dfp_cocr_amd.receivers import ReceiverOperator, impedance

    gold = json.loads((Path(__file__).parent / "golden" / "config2_oracle.json").read_text())
    say("full size: config-2 mesh, 16 systems against the oracle's fixture")
    m = M.random_model(100, 100, 60, seed=0, h=50.0, name="bench")
    freqs = M.frequencies(1e-3, 1e2, 8)
    cases = [(f, a) for f in freqs for a in (1, 0)]
    p, kd, kg, sd, sg = M.assemble_ks(m)
    ma = M.mesh_arrays(m)
    prs = np.stack([M.assemble_rhs(m, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    with Solver(0) as B:
        B.set_mesh_order(ma, ORDER_GRID)
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(ma)
        B.set_rhs(prs, om)
        tuned = B.autotune("fp64", 2)
        say(f"autotuned: {tuned}")
        assert tuned["spmv_layout"] in ("stencil", "stencil_lds")
        x, st = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
        xm, sm = B.solve_batch(om[-1:], prs[-1:], eps=1e-8, precision="mixed")
    rows = np.repeat(np.arange(p.nb), np.diff(p.ig))

    def sym(di, gg):
        L = sp.csr_matrix((gg, (rows, p.jg)), shape=(p.nb, p.nb))
        return (L + L.T + sp.diags(di)).tocsr()

    K, S = sym(kd, kg), sym(sd, sg)

    def true_rel(k, xv):
        b = M.to_complex(prs[k])
        xc = M.to_complex(xv)
        return np.linalg.norm(b - (K @ xc + 1j * om[k] * (S @ xc))) / np.linalg.norm(b)

    for k, g in enumerate(gold["systems"]):
        assert abs(g["f"] - cases[k][0]) <= 1e-12 * g["f"] and g["pol"] == "XY"[1 - cases[k][1]]
        r = true_rel(k, x[k])
        say(f"{k}: f={g['f']:.3g} {g['pol']}: {st[k]['iter']} it (oracle {g['iter']}), "
            f"true rel res {r:.2e} (oracle {g['true_rel']:.2e})")
        assert st[k]["req"] == 0 and st[k]["residual_rel"] < 1e-8
        assert abs(st[k]["iter"] - g["iter"]) <= 2 + 0.02 * g["iter"]
        assert r <= FP64_TRUE_BOUND
    R = ReceiverOperator(m, np.array(gold["receivers"]), smoothing=None)  # as the fixture
    for zg in gold["Z"]:
        f = zg["f"]
        kx = [k for k, c in enumerate(cases) if c == (f, 1)][0]
        Ex, Hx = R.fields(M.to_complex(x[kx]), f, 1)
        Ey, Hy = R.fields(M.to_complex(x[kx + 1]), f, 0)
        Z = impedance(Ex, Hx, Ey, Hy)
        Zo = np.array(zg["re"]) + 1j * np.array(zg["im"])
        scale = np.abs(Zo[:, 0, 1])
        err = np.abs(Z - Zo) / scale[:, None, None]
        err[:, 1, 0] = np.abs(Z[:, 1, 0] - Zo[:, 1, 0]) / np.abs(Zo[:, 1, 0])
        say(f"Z at {f:.3g} Hz: max rel diff {err.max():.2e}")
        assert err.max() <= 1e-6
    r = true_rel(len(cases) - 1, xm[0])
    say(f"mixed f={cases[-1][0]:.3g} Hz: true rel res {r:.2e}")
    assert sm[0]["req"] == 0 and r <= 1e-8


def test_batch_with_zero_rhs_and_padding(O, M, c1_model, c1_mesh):
    """A batch of 3 systems (one padding slot) whose middle right-hand side is zero: that
    system retires in the prologue (X0 test, x = 0 exactly) and its lanes then run the
    identity path of the kernels for the rest of the solve. Its x stays exactly 0, and the
    other two match a batch of those two alone, FP64 and mixed."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    say("batch with a zero right-hand side and a padding slot")
    freqs = [0.1, 1.0, 10.0]
    p, kd, kg, sd, sg = M.assemble_ks(c1_model)
    prs = np.stack([M.assemble_rhs(c1_model, f, 1) for f in freqs])
    prs[1] = 0.0
    om = np.array([M.omega_of(f) for f in freqs])
    with Solver(0) as B:
        B.set_matrix_ks(p, kd, kg, sd, sg)
        B.set_gradient(c1_mesh)
        x3, s3 = B.solve_batch(om, prs, eps=1e-8, precision="fp64")
        x2, s2 = B.solve_batch(om[[0, 2]], prs[[0, 2]], eps=1e-8, precision="fp64")
        xm, sm = B.solve_batch(om, prs, eps=1e-8, precision="mixed")
    say(f"iters {[q['iter'] for q in s3]} vs {[q['iter'] for q in s2]}")
    assert s3[1]["req"] == -1 and not x3[1].any()
    assert s3[0]["req"] == 0 and s3[2]["req"] == 0
    for k, j in ((0, 0), (2, 1)):
        assert abs(s3[k]["iter"] - s2[j]["iter"]) <= 2 + 0.02 * s2[j]["iter"]
        assert rel(x3[k], x2[j]) <= 1e-6
    assert sm[1]["req"] == -1 and not xm[1].any()
    for k in (0, 2):
        s = M.assemble(c1_model, freqs[k], 1)
        res = M.to_complex(s.pr) - M.sparse_matrix(s) @ M.to_complex(xm[k])
        assert sm[k]["req"] == 0
        assert np.linalg.norm(res) / np.linalg.norm(M.to_complex(s.pr)) < 1e-8
