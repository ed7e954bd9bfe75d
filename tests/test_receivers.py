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


@pytest.fixture(scope="module")
def layered(M):
    """Three homogeneous layers under air on a graded mesh, receivers on and between nodes:
    one earth material at the surface, so the E subdomain has rings."""
    x = M.padded_axis(100.0, 5, 2)
    y = M.padded_axis(80.0, 4, 2)
    z = M.z_axis(100.0, 3, 6, 2)
    tops = M.snap_to_nodes(np.array([-500.0, -200.0, 0.0]), z)
    sig = np.array([1e-3, 1e-1, 1e-2])
    m0 = M.Model(x, y, z, np.zeros(1), np.zeros(1), tops, sig)
    xc, yc, zc = m0.cell_centres()
    s0 = M._layered(zc, tops, sig)
    s0[zc > 0] = M.AIR_SIGMA
    m = M._finish(M.Model(x, y, z, s0, s0.copy(), tops, sig))
    rec = np.array([[0.0, 0.0, 0.0], [100.0, -80.0, 0.0], [37.0, 51.0, 0.0],
                    [-160.0, 20.0, 0.0]])
    return m, rec


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
    """A = (a + b y + c z, d + e x + f z, 0) lies in the edge space: the point evaluation
    returns A and curl A exactly at any surface point."""
    from mtcalc_with_dfp_cocr_amd.receivers import ReceiverOperator

    _, rec = op
    R = ReceiverOperator(small_model, rec, smoothing=None)
    a, b, c, d, e, f = 0.3, -1.2, 0.7, 2.0, 0.4, -0.9
    field = lambda P: np.stack([a + b * P[:, 1] + c * P[:, 2], d + e * P[:, 0] + f * P[:, 2],
                                np.zeros(len(P))], 1)
    x = _interp(M, small_model, field)
    A = (R.RE @ x).reshape(-1, 2)
    H = (R.RH @ x).reshape(-1, 2) * M.MU_0
    np.testing.assert_allclose(A[:, 0], a + b * rec[:, 1], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(A[:, 1], d + e * rec[:, 0], rtol=1e-12, atol=1e-12)
    # curl = (dAz/dy - dAy/dz, dAx/dz - dAz/dx, dAy/dx - dAx/dy) = (-f, c, e - b)
    np.testing.assert_allclose(H[:, 0], -f, rtol=1e-12)
    np.testing.assert_allclose(H[:, 1], c, rtol=1e-12)
    np.testing.assert_allclose(R.hz(x) * M.MU_0, e - b, rtol=1e-12)


def test_smoothed_exact_on_constant_fields(M, small_model, layered, op):
    """The L2 smoothing reproduces what the trilinear space holds: a constant A gives E = A, and
    a linear A (constant curl) gives H = curl A, exactly, on a model whose every cell is its
    own material (E from one cell) and on a layered one (E over rings of the top layer)."""
    from mtcalc_with_dfp_cocr_amd.receivers import ReceiverOperator

    for m, rec in ((small_model, op[1]), layered):
        R = ReceiverOperator(m, rec, smoothing=5)
        a, b, c, d, e, f = 0.3, -1.2, 0.7, 2.0, 0.4, -0.9
        const = lambda P: np.stack([np.full(len(P), a), np.full(len(P), d), np.zeros(len(P))], 1)
        A = (R.RE @ _interp(M, m, const)).reshape(-1, 2)
        np.testing.assert_allclose(A[:, 0], a, rtol=1e-11)
        np.testing.assert_allclose(A[:, 1], d, rtol=1e-11)
        lin = lambda P: np.stack([a + b * P[:, 1] + c * P[:, 2], d + e * P[:, 0] + f * P[:, 2],
                                  np.zeros(len(P))], 1)
        x = _interp(M, m, lin)
        H = (R.RH @ x).reshape(-1, 2) * M.MU_0
        np.testing.assert_allclose(H[:, 0], -f, rtol=1e-11)
        np.testing.assert_allclose(H[:, 1], c, rtol=1e-11)
        np.testing.assert_allclose(R.hz(x) * M.MU_0, e - b, rtol=1e-11)


def _brute_smoothing(M, m, rec, x, n_neib):
    """The smoothing restated independently for H on a regular mesh: the subdomain as the cells
    within n_neib cells (Chebyshev distance in cell indices) of the receiver cells, the first
    cell holding each receiver by a scan in element order, centre curls from the edge DOFs by
    Stokes on the cell's mid-planes, the nodal mass matrix and load by 2x2x2 Gauss quadrature,
    a dense solve, trilinear interpolation."""
    ma = M.mesh_arrays(m)
    e2 = ma.edges.reshape(-1, 2)
    ncx, ncy, ncz = m.shape_cells
    nxn, nyn = ncx + 1, ncy + 1
    # the value of each edge's tangential field: DOF * 2 / length (the DOF is F.t |t| / 2)
    L = np.linalg.norm(ma.xyz[e2[:, 1]] - ma.xyz[e2[:, 0]], axis=1)
    tang = x * 2.0 / L
    eid = {}
    for k, (a, b) in enumerate(e2):
        eid[(min(a, b), max(a, b))] = k
    nid = lambda i, j, k: (k * nyn + j) * nxn + i

    def edge(n0, n1):
        return eid[(min(n0, n1), max(n0, n1))]

    def curl_centre(i, j, k):
        hx, hy, hz = m.x[i + 1] - m.x[i], m.y[j + 1] - m.y[j], m.z[k + 1] - m.z[k]
        # mean tangential component over a cell face = mean of its two parallel edges
        Ax = lambda jj, kk: 0.5 * (tang[edge(nid(i, jj, kk), nid(i + 1, jj, kk))])
        Ay = lambda ii, kk: 0.5 * (tang[edge(nid(ii, j, kk), nid(ii, j + 1, kk))])
        Az = lambda ii, jj: 0.5 * (tang[edge(nid(ii, jj, k), nid(ii, jj, k + 1))])
        Axf = lambda side, ax: (Ax(j + side, k) + Ax(j + side, k + 1)) if ax == "y" else \
            (Ax(j, k + side) + Ax(j + 1, k + side))
        Ayf = lambda side, ax: (Ay(i + side, k) + Ay(i + side, k + 1)) if ax == "x" else \
            (Ay(i, k + side) + Ay(i + 1, k + side))
        Azf = lambda side, ax: (Az(i + side, j) + Az(i + side, j + 1)) if ax == "x" else \
            (Az(i, j + side) + Az(i + 1, j + side))
        cx = (Azf(1, "y") - Azf(0, "y")) / hy - (Ayf(1, "z") - Ayf(0, "z")) / hz
        cy = (Axf(1, "z") - Axf(0, "z")) / hz - (Azf(1, "x") - Azf(0, "x")) / hx
        cz = (Ayf(1, "x") - Ayf(0, "x")) / hx - (Axf(1, "y") - Axf(0, "y")) / hy
        return np.array([cx, cy, cz])

    def first_cell(p):
        for k in range(ncz):
            for j in range(ncy):
                for i in range(ncx):
                    if (m.x[i] - 1e-6 <= p[0] <= m.x[i + 1] + 1e-6 and
                            m.y[j] - 1e-6 <= p[1] <= m.y[j + 1] + 1e-6 and
                            m.z[k] - 1e-6 <= p[2] <= m.z[k + 1] + 1e-6):
                        return i, j, k

    rc = [first_cell(p) for p in rec]
    cells = [(i, j, k) for k in range(ncz) for j in range(ncy) for i in range(ncx)
             if any(max(abs(i - a), abs(j - b), abs(k - c)) <= n_neib for a, b, c in rc)]
    nodes = sorted({nid(i + bx, j + by, k + bz) for i, j, k in cells
                    for bx in (0, 1) for by in (0, 1) for bz in (0, 1)})
    loc = {n: q for q, n in enumerate(nodes)}
    g = np.array([-1.0, 1.0]) / np.sqrt(3.0)
    Mm = np.zeros((len(nodes), len(nodes)))
    F = np.zeros((len(nodes), 3))
    for i, j, k in cells:
        h = (m.x[i + 1] - m.x[i], m.y[j + 1] - m.y[j], m.z[k + 1] - m.z[k])
        v = curl_centre(i, j, k)
        ids = [loc[nid(i + bx, j + by, k + bz)] for bz in (0, 1) for by in (0, 1) for bx in (0, 1)]
        for gx in g:
            for gy in g:
                for gz in g:
                    t = (0.5 * (1 + gx), 0.5 * (1 + gy), 0.5 * (1 + gz))
                    phi = np.array([(t[0] if bx else 1 - t[0]) * (t[1] if by else 1 - t[1]) *
                                    (t[2] if bz else 1 - t[2])
                                    for bz in (0, 1) for by in (0, 1) for bx in (0, 1)])
                    w = h[0] * h[1] * h[2] / 8.0
                    Mm[np.ix_(ids, ids)] += w * np.outer(phi, phi)
                    F[ids] += w * np.outer(phi, v)
    U = np.linalg.solve(Mm, F)
    out = []
    for p, (i, j, k) in zip(rec, rc):
        t = ((p[0] - m.x[i]) / (m.x[i + 1] - m.x[i]), (p[1] - m.y[j]) / (m.y[j + 1] - m.y[j]),
             (p[2] - m.z[k]) / (m.z[k + 1] - m.z[k]))
        val = np.zeros(3)
        for bz in (0, 1):
            for by in (0, 1):
                for bx in (0, 1):
                    wgt = (t[0] if bx else 1 - t[0]) * (t[1] if by else 1 - t[1]) * \
                        (t[2] if bz else 1 - t[2])
                    val += wgt * U[loc[nid(i + bx, j + by, k + bz)]]
        out.append(val)
    return np.array(out)


@pytest.mark.parametrize("n_neib", [0, 2])
def test_smoothed_h_matches_brute_force(M, layered, n_neib):
    """H from the smoothing operator against the independent restatement above, for a random
    edge vector (so the centre curls vary from cell to cell)."""
    from mtcalc_with_dfp_cocr_amd.receivers import ReceiverOperator

    m, rec = layered
    x = np.random.default_rng(11).standard_normal(m.n_edges)
    R = ReceiverOperator(m, rec, smoothing=n_neib)
    ref = _brute_smoothing(M, m, rec, x, n_neib)
    H = (R.RH @ x).reshape(-1, 2) * M.MU_0
    np.testing.assert_allclose(H[:, 0], ref[:, 0], rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_allclose(H[:, 1], ref[:, 1], rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_allclose(R.hz(x) * M.MU_0, ref[:, 2], rtol=1e-9,
                               atol=1e-12 * np.abs(ref).max())


def test_smoothed_e_stays_in_the_material(M, layered):
    """E is smoothed within the receivers' material only: a field on the edges outside the top
    layer (deeper layers, air) gives E = 0, one on the top layer's edges does not; H spans
    both sides of the surface. The transposed maps are the transposes (random probes)."""
    from mtcalc_with_dfp_cocr_amd.receivers import ReceiverOperator

    m, rec = layered
    R = ReceiverOperator(m, rec, smoothing=5)
    ma = M.mesh_arrays(m)
    e2 = ma.edges.reshape(-1, 2)
    zlo = np.minimum(ma.xyz[e2[:, 0], 2], ma.xyz[e2[:, 1], 2])
    zhi = np.maximum(ma.xyz[e2[:, 0], 2], ma.xyz[e2[:, 1], 2])
    top = m.layer_top[-2]          # base of the top layer
    in_top = (zlo >= top - 1e-9) & (zhi <= 1e-9)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(m.n_edges)
    assert np.abs(R.RE @ np.where(in_top, 0.0, x)).max() == 0.0
    assert np.abs(R.RE @ np.where(in_top, x, 0.0)).min() > 0.0
    assert np.abs(R.RH @ np.where(zlo >= 0, x, 0.0)).min() > 0.0
    for F in (R.RE, R.RH, R.RHz):
        y = rng.standard_normal(F.shape[0]) + 1j * rng.standard_normal(F.shape[0])
        xc = x + 1j * rng.standard_normal(m.n_edges)
        assert np.vdot(y.conj(), F @ xc) == pytest.approx(np.vdot((F.T @ y).conj(), xc),
                                                          rel=1e-12)


def test_smoothed_fields_keep_the_models_symmetry(O, M, c1_model, c1_mesh):
    """Config 1 is mirror-symmetric about x = 0 and y = 0 (block and mesh), and its receivers
    sit on mesh nodes. The smoothed fields keep that symmetry: Zxy(x, y) = Zxy(-x, y) and
    Zyx(x, y) = Zyx(x, -y) to 1e-4, and the tipper vanishes at the centre receiver (1e-4 of
    its largest value). The point evaluation, which takes H from the cell on one side of a
    node, misses both by more than 1e-2."""
    from mtcalc_with_dfp_cocr_amd.receivers import ReceiverOperator, impedance, tipper

    nu = 1.0
    sols = {}
    for alpha in (1, 0):
        x, so, _ = O.pcocr(M.assemble(c1_model, nu, alpha), c1_mesh, eps=1e-10, nthreads=4)
        assert so["req"] == 0
        sols[alpha] = M.to_complex(x)
    out = {}
    for smoothing in (5, None):
        R = ReceiverOperator(c1_model, c1_model.receivers, smoothing=smoothing)
        Ex, Hx = R.fields(sols[1], nu, 1)
        Ey, Hy = R.fields(sols[0], nu, 0)
        Z = impedance(Ex, Hx, Ey, Hy).reshape(5, 5, 2, 2)     # receiver (ix, iy)
        T = tipper(R.hz(sols[1]), Hx, R.hz(sols[0]), Hy)
        mirror = max(np.abs(Z[:, :, 0, 1] - Z[::-1, :, 0, 1]).max() / np.abs(Z[:, :, 0, 1]).max(),
                     np.abs(Z[:, :, 1, 0] - Z[:, ::-1, 1, 0]).max() / np.abs(Z[:, :, 1, 0]).max())
        out[smoothing] = (mirror, np.abs(T[12]).max() / np.abs(T).max())
    assert out[5][0] <= 1e-4 and out[5][1] <= 1e-4, out
    assert out[None][0] > 1e-2 and out[None][1] > 1e-2, out


def test_tipper_solves_hz_equals_t_h():
    """T from two polarizations reproduces Hz = Tzx Hx + Tzy Hy for both of them, and a field
    with no Hz (the 1-D normal field alone) has T = 0."""
    from mtcalc_with_dfp_cocr_amd.receivers import tipper

    rng = np.random.default_rng(9)
    R = 6
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    T = c(R, 2)
    Hx, Hy = c(R, 2), c(R, 2)   # H of the X and Y polarizations
    hzx = T[:, 0] * Hx[:, 0] + T[:, 1] * Hx[:, 1]
    hzy = T[:, 0] * Hy[:, 0] + T[:, 1] * Hy[:, 1]
    np.testing.assert_allclose(tipper(hzx, Hx, hzy, Hy), T, rtol=1e-12)
    assert not tipper(np.zeros(R), Hx, np.zeros(R), Hy).any()


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
    # no anomaly, no vertical field: the tipper vanishes
    from mtcalc_with_dfp_cocr_amd.receivers import tipper

    assert not tipper(R.hz(zero), Hx, R.hz(zero), Hy).any()
