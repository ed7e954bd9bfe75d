"""Adjoint gradient of the impedance misfit (host logic, exact sparse LU solves): central finite
differences in log(sigma), the normal-field DOFs against the right-hand side they build, the
cell mass forms against the assembled S, and the torch.autograd wrapper."""
import numpy as np
import pytest

from helpers import DirectKSSolver


@pytest.fixture(scope="module")
def problem(M, small_model):
    from mtcalc_with_dfp_cocr_amd.adjoint import ImpedanceMisfit

    m = small_model
    rng = np.random.default_rng(21)
    xs, ys = m.x, m.y
    rec = np.stack([rng.uniform(xs[3], xs[-4], 5), rng.uniform(ys[3], ys[-4], 5), np.zeros(5)], 1)
    freqs = [0.1, 3.0]
    prob = ImpedanceMisfit(m, freqs, rec, solver=DirectKSSolver(M, m))
    earth = m.sigma > 1e-6
    sig_true = m.sigma.copy()
    sig_true[earth] *= np.exp(0.3 * rng.standard_normal(int(earth.sum())))
    _, zobs, _ = prob.responses(sig_true)
    prob.z_obs = zobs
    return prob


def test_normal_dofs_build_the_rhs(M, small_model):
    """b = -i omega (S(sigma) - S(sigma0)) a0 on interior rows: the RHS from the model layer
    equals the one rebuilt from the normal-field DOFs and the K/S split."""
    import scipy.sparse as sp

    from mtcalc_with_dfp_cocr_amd.adjoint import normal_dofs, with_sigma

    m = small_model
    nu = 2.0
    w = M.omega_of(nu)
    for alpha in (1, 0):
        a0 = normal_dofs(m, nu, alpha)
        b = M.assemble_rhs(m, nu, alpha)
        b = b[0::2] + 1j * b[1::2]
        ds = m.sigma - m.sigma0
        p, kd, kg, sd, sg = M.assemble_ks(with_sigma(m, ds))
        rows = np.repeat(np.arange(p.nb), np.diff(p.ig))
        S = (sp.csr_matrix((sg, (rows, p.jg)), shape=(p.nb, p.nb))
             + sp.csr_matrix((sg, (p.jg, rows)), shape=(p.nb, p.nb)) + sp.diags(sd))
        ref = -1j * w * (S @ a0)
        interior = kd != 1.0  # Dirichlet rows hold K = 1, S = 0 and b = 0
        assert np.abs(b[~interior]).max() == 0
        # the K/S split also zeroes the Dirichlet columns, which the RHS keeps (a0 is not 0
        # there): compare the rows that touch no Dirichlet edge
        touch = np.zeros(p.nb, bool)
        bnd = ~interior
        np.logical_or.at(touch, rows, bnd[p.jg])
        np.logical_or.at(touch, p.jg, bnd[rows])
        keep = interior & ~touch
        assert keep.sum() > 0.3 * p.nb
        np.testing.assert_allclose(b[keep], ref[keep], rtol=1e-9,
                                   atol=1e-12 * np.abs(ref).max())


def test_boundary_edges_are_the_dirichlet_rows(M, small_model):
    """The adjoint's Dirichlet edges are the rows the assembly holds at identity: K = 1 and
    S = 0 on the diagonal and no off-diagonal entry in their rows or columns."""
    from mtcalc_with_dfp_cocr_amd.adjoint import boundary_edges

    p, kd, kg, sd, sg = M.assemble_ks(small_model)
    bnd = boundary_edges(small_model)
    rows = np.repeat(np.arange(p.nb), np.diff(p.ig))
    touched = np.zeros(p.nb, bool)
    nzk = (kg != 0) | (sg != 0)
    touched[rows[nzk]] = True
    touched[p.jg[nzk]] = True
    assert bnd.sum() > 0
    assert np.all(kd[bnd] == 1.0) and np.all(sd[bnd] == 0.0) and not touched[bnd].any()
    assert np.all(touched[~bnd] | (sd[~bnd] != 0))


def test_cell_mass_forms_sum_to_s(M, small_model):
    """sum_c sigma_c u_c^T M_c v_c = u^T S v (interior rows and columns)."""
    from mtcalc_with_dfp_cocr_amd.adjoint import cell_mass_forms

    m = small_model
    rng = np.random.default_rng(3)
    p, kd, kg, sd, sg = M.assemble_ks(m)
    interior = kd != 1.0
    u = (rng.standard_normal(m.n_edges) + 1j * rng.standard_normal(m.n_edges)) * interior
    v = (rng.standard_normal(m.n_edges) + 1j * rng.standard_normal(m.n_edges)) * interior
    rows = np.repeat(np.arange(p.nb), np.diff(p.ig))
    Sv = sd * v
    np.add.at(Sv, rows, sg * v[p.jg])
    np.add.at(Sv, p.jg, sg * v[rows])
    lhs = np.sum(m.sigma * cell_mass_forms(m, u, v))
    np.testing.assert_allclose(lhs, u @ Sv, rtol=1e-12)


def test_gradient_matches_finite_differences(problem):
    sigma = problem.m.sigma.copy()
    st = problem.evaluate(sigma)
    grad = problem.gradient(st) * sigma
    xc, yc, zc = problem.m.cell_centres()
    earth = np.flatnonzero(zc < 0)
    cells = np.concatenate([earth[np.argsort(np.abs(grad[earth]))[-4:]],
                            np.random.default_rng(5).choice(earth, 4, replace=False)])
    h = 1e-5
    for c in cells:
        s1, s2 = sigma.copy(), sigma.copy()
        s1[c] *= np.exp(h)
        s2[c] *= np.exp(-h)
        fd = (problem.misfit(s1) - problem.misfit(s2)) / (2 * h)
        assert grad[c] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(grad).max())


def test_defect_correction_reproduces_the_solution(problem):
    """responses(sigma', base) = x0 + A'^-1 [(b' - b) - i omega (S' - S) x0] solves
    A' x = b' - r0 (r0 the base's residual, at LU's level here): with exact LU solves, at a
    perturbed cell and along a whole direction, its true residual (from the solver's own K and
    S) is at LU's level and its impedances and J equal those of the direct solution. (x
    itself is not compared: in the air, sigma = 1e-8 leaves a near-null space of gradient
    fields that round-off moves without changing any field at the surface.)"""
    from mtcalc_with_dfp_cocr_amd import adjoint as AD

    sigma = problem.m.sigma.copy()
    st = problem.evaluate(sigma)
    xc, yc, zc = problem.m.cell_centres()
    earth = np.flatnonzero(zc < 0)
    s1 = sigma.copy()
    s1[earth[len(earth) // 2]] *= np.exp(1e-3)
    s2 = sigma * np.exp(0.05 * np.random.default_rng(2).standard_normal(len(sigma)))
    om = np.repeat(problem.omegas, 2)

    def true_rel(s, x):
        m = AD.with_sigma(problem.m, s)
        problem.solver.set_sigma(s)
        K, S = problem.solver.K, problem.solver.S
        out = []
        for k, f in enumerate(problem.freqs):
            for p, alpha in enumerate((1, 0)):
                pr = problem.solver.M.assemble_rhs(m, f, alpha)
                b = pr[0::2] + 1j * pr[1::2]
                xs = x[k, p]
                out.append(np.linalg.norm(b - K @ xs - 1j * om[2 * k] * (S @ xs))
                           / np.linalg.norm(b))
        return max(out)

    for s in (s1, s2):
        x_direct, Z_direct, _ = problem.responses(s)
        x_dc, Z_dc, _ = problem.responses(s, st)
        assert true_rel(s, x_dc) <= max(1e-11, 2 * true_rel(s, x_direct))
        np.testing.assert_allclose(Z_dc, Z_direct, rtol=1e-9, atol=1e-9 * np.abs(Z_direct).max())
        assert problem.misfit(s, st) == pytest.approx(problem.misfit(s), rel=1e-10)
    # the right-hand side without the (S' - S) x0 term leaves A' x = b' + (A' - A) x0
    good = AD.s_matrix
    try:
        problem._base_key = None
        AD.s_matrix = lambda m: 0.0 * good(m)
        x_bad, _, _ = problem.responses(s2, st)
    finally:
        AD.s_matrix = good
        problem._base_key = None
    assert true_rel(s2, x_bad) > 1e-3


def test_torch_wrapper(problem):
    torch = pytest.importorskip("torch")
    from mtcalc_with_dfp_cocr_amd.adjoint import torch_misfit

    sigma = problem.m.sigma.copy()
    ls = torch.tensor(np.log(sigma), requires_grad=True)
    J = torch_misfit(problem)(ls)
    J.backward()
    st = problem.evaluate(sigma)
    assert float(J.detach()) == pytest.approx(st.J, rel=1e-12)
    ref = problem.gradient(st) * sigma
    # the wrapper evaluates at exp(log(sigma)), one rounding away from sigma
    np.testing.assert_allclose(ls.grad.numpy(), ref, rtol=1e-8, atol=1e-9 * np.abs(ref).max())


def test_directional_derivative(problem):
    """A random +-1 direction in log(sigma) over the earth cells: grad . d against the central
    difference of J along d."""
    from mtcalc_with_dfp_cocr_amd.adjoint import _fd_direction

    sigma = problem.m.sigma.copy()
    grad = problem.gradient(problem.evaluate(sigma)) * sigma
    xc, yc, zc = problem.m.cell_centres()
    earth = zc < 0
    d = np.where(earth, np.random.default_rng(9).choice([-1.0, 1.0], size=len(sigma)), 0.0)
    fd = _fd_direction(problem, sigma, d, 1e-5)
    assert grad @ d == pytest.approx(fd, rel=1e-6)
