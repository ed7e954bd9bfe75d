"""Receiver fields and impedance (FP64): exactness on fields the edge space contains, and the
analytic half-space impedance."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def op(M, small_model):
    from mtcalc_with_dfp_cocr_amd.receivers import ReceiverOperator

    rng = np.random.default_rng(4)
    xs = small_model.x
    ys = small_model.y
    rec = np.stack([rng.uniform(xs[2], xs[-3], 7), rng.uniform(ys[2], ys[-3], 7), np.zeros(7)], 1)
    return ReceiverOperator(small_model, rec), rec


def _interp(M, m, field):
    """Edge DOFs of a vector field: tangential component at the edge midpoint times h/2."""
    ma = M.mesh_arrays(m)
    e = ma.edges.reshape(-1, 2)
    p0, p1 = ma.xyz[e[:, 0]], ma.xyz[e[:, 1]]
    mid = 0.5 * (p0 + p1)
    t = p1 - p0
    F = field(mid)
    return np.einsum("ij,ij->i", F, t) / 2.0  # (F . t_hat) * |t| / 2


def test_linear_fields_exact(M, small_model, op):
    """A = (a + b y + c z, d + e x + f z, 0) lies in the edge space: the operator returns A and
    curl A exactly at any surface point."""
    R, rec = op
    a, b, c, d, e, f = 0.3, -1.2, 0.7, 2.0, 0.4, -0.9
    field = lambda P: np.stack([a + b * P[:, 1] + c * P[:, 2], d + e * P[:, 0] + f * P[:, 2],
                                np.zeros(len(P))], 1)
    x = _interp(M, small_model, field)
    A = (R.RE @ x).reshape(-1, 2)
    H = (R.RH @ x).reshape(-1, 2) * M.MU_0
    np.testing.assert_allclose(A[:, 0], a + b * rec[:, 1], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(A[:, 1], d + e * rec[:, 0], rtol=1e-12, atol=1e-12)
    # curl = (dAz/dy - dAy/dz, dAx/dz - dAz/dx, .) = (-f, c, .)
    np.testing.assert_allclose(H[:, 0], -f, rtol=1e-12)
    np.testing.assert_allclose(H[:, 1], c, rtol=1e-12)


@pytest.mark.parametrize("nu", [1e-2, 1.0, 100.0])
def test_halfspace_impedance(M, nu):
    """sigma = sigma0 (no anomaly, x = 0): Z_xy = -Z_yx = -i w mu / k, rho_a = 1/sigma,
    phase 45 degrees (within the 1-D discretisation error)."""
    from mtcalc_with_dfp_cocr_amd.receivers import (ReceiverOperator, apparent_resistivity,
                                                    impedance)

    sig = 0.01
    x = M.padded_axis(100.0, 4, 2)
    z = M.z_axis(100.0, 2, 4, 2)
    ncell = (len(x) - 1) ** 2 * (len(z) - 1)
    m = M.Model(x, x.copy(), z, np.full(ncell, sig), np.full(ncell, sig), np.array([0.0]),
                np.array([sig]))
    R = ReceiverOperator(m, np.array([[0.0, 0.0, 0.0], [50.0, -30.0, 0.0]]))
    zero = np.zeros(m.n_edges, complex)
    Ex, Hx = R.fields(zero, nu, 1)
    Ey, Hy = R.fields(zero, nu, 0)
    Z = impedance(Ex, Hx, Ey, Hy)
    rho, ph = apparent_resistivity(Z, nu)
    np.testing.assert_allclose(rho[:, 0, 1], 1 / sig, rtol=3e-3)
    np.testing.assert_allclose(rho[:, 1, 0], 1 / sig, rtol=3e-3)
    np.testing.assert_allclose(Z[:, 0, 1], -Z[:, 1, 0], rtol=1e-12)
    assert np.abs(Z[:, 0, 0]).max() == 0 and np.abs(Z[:, 1, 1]).max() == 0
    w = M.omega_of(nu)
    k = np.sqrt(1j * w * 4e-7 * np.pi * sig)
    np.testing.assert_allclose(Z[:, 0, 1], -1j * w * 4e-7 * np.pi / k, rtol=3e-3)
