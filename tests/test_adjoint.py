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
