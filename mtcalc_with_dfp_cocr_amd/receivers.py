"""Receiver fields and the impedance tensor in FP64.

Restates what follows the solver in the reference, for the quantities its results are compared by:
  fields   E = -i*omega*(A + A0), H = curl(A + A0)/mu0 at the receivers
           (Code/OutputNu/give_out_vec_mt.cpp:101-272: E from the vector potential, H from its
           curl, the normal field of Harm1D added back; Harm1D writes the normal H as mu*alpha /
           mu*(alpha-1) at z >= 0, Code/Harm1D/divgrad_1d.cpp:283-297)
  impedance  Z = E H^-1 from the two polarizations, rho_a = |Z|^2/(omega*mu0), phase = -arg Z
           (Code/calcimpedance/calcimpedance.for:123-255 solves the same 2x2 complex system in
           its 4x4 real form, single precision; here FP64)
  tipper   the Wiese-Parkinson vector T = (Tzx, Tzy) with Hz = Tzx Hx + Tzy Hy, from the same
           system with Hz in place of E (calcimpedance.for:294-324); the 1-D normal field has no
           Hz
Two evaluations of A and curl A at a receiver, as the reference has (the vfem_direct_output
switch, Code/OutputNu/give_out_vec_mt.cpp:109, default off, GeoPrepDocSettings.h:207):
  smoothed (smoothing = n_neib, the reference's default, n_neib = 5,
           OutputResultant3d.cpp:46): each component's element-centre values are L2-projected
           onto the trilinear nodal functions of a subdomain (the receivers' elements and
           n_neib rings of node-neighbours), and the projection is interpolated at the
           receiver. H (curl A, no discontinuity) uses one subdomain over every material; E (A,
           with discontinuity) one per material of the receivers' elements, air excluded, the
           rings kept within the material (OutputResultant3d.cpp:82-197, 265-445;
           Subdomain.cpp:130-271, 484-623). Each receiver belongs to the first element, in
           element order, whose box holds it within 1e-6; the projection is continuous, so
           the choice matters only through the subdomains. The reference factorises the
           subdomain mass matrix with its own sparse Cholesky; here scipy's sparse LU.
  point    (smoothing = None, the direct-output counterpart, OutputArbitrary.cpp:202-262):
           the edge basis evaluated at the receiver: E from the cell below the surface
           (tangential E is continuous), H as the mean of the curls of the cells below and
           above.
Both are linear in the solution vector (sparse maps RE, RH, RHz from the edge DOFs).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .model import MU_0, Model, harm1d, mesh_arrays, omega_of

# local edge positions on [-1,1]^3 (x-edges at (y,z), y-edges at (x,z), z-edges at (x,y))
_POS = {0: [(-1, -1), (1, -1), (-1, 1), (1, 1)],
        1: [(-1, -1), (-1, 1), (1, -1), (1, 1)],
        2: [(-1, -1), (1, -1), (-1, 1), (1, 1)]}
_OTHER = {0: (1, 2), 1: (0, 2), 2: (0, 1)}


def _basis_at(h, xi, zeta_face):
    """Values and curls of the 12 local basis functions at reference point (xi0, xi1, zeta_face).
    Returns phi[12,3], curl[12,3] (physical components)."""
    X = np.array([xi[0], xi[1], zeta_face], dtype=np.float64)
    phi = np.zeros((12, 3))
    curl = np.zeros((12, 3))
    k = 0
    for d in range(3):
        k1, k2 = _OTHER[d]
        for s1, s2 in _POS[d]:
            c = (2.0 / h[d]) * 0.25
            phi[k, d] = c * (1 + s1 * X[k1]) * (1 + s2 * X[k2])
            grad = np.zeros(3)
            grad[k1] = c * s1 * (2.0 / h[k1]) * (1 + s2 * X[k2])
            grad[k2] = c * s2 * (2.0 / h[k2]) * (1 + s1 * X[k1])
            e = np.zeros(3)
            e[d] = 1.0
            curl[k] = np.cross(grad, e)
            k += 1
    return phi, curl


def _trilinear_mass(h):
    """8x8 mass matrix of the trilinear nodal functions on an h[0] x h[1] x h[2] box, local node
    l = bx + 2 by + 4 bz (bit d set: the node at the upper end of axis d)."""
    one = lambda a: a / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]])
    return np.kron(np.kron(one(h[2]), one(h[1])), one(h[0]))


class _Grid:
    """Cell and node indexing of a regular mesh: cell (k, j, i) = (k*ncy + j)*ncx + i, node
    (k, j, i) = (k*nyn + j)*nxn + i, the element and node orders of the mesh files."""

    def __init__(self, m: Model):
        self.m = m
        self.ncx, self.ncy, self.ncz = m.shape_cells
        self.nxn, self.nyn, self.nzn = m.shape_nodes

    def cell_nodes(self, cells):
        k, r = np.divmod(cells, self.ncx * self.ncy)
        j, i = np.divmod(r, self.ncx)
        out = np.empty((len(cells), 8), np.int64)
        for l in range(8):
            bx, by, bz = l & 1, (l >> 1) & 1, (l >> 2) & 1
            out[:, l] = ((k + bz) * self.nyn + j + by) * self.nxn + i + bx
        return out

    def cell_sizes(self, cells):
        k, r = np.divmod(cells, self.ncx * self.ncy)
        j, i = np.divmod(r, self.ncx)
        m = self.m
        return np.stack([m.x[i + 1] - m.x[i], m.y[j + 1] - m.y[j], m.z[k + 1] - m.z[k]], 1)

    def first_cell_holding(self, p, allowed, tol=1e-6):
        """The first cell in element order whose box, widened by tol, holds point p and whose
        flag in `allowed` is set (OutputResultant3d.cpp:265-410), or -1."""
        m = self.m
        rng = []
        for ax, c in zip((m.x, m.y, m.z), p):
            lo = np.flatnonzero((ax[:-1] - tol <= c) & (c <= ax[1:] + tol))
            rng.append(lo)
        ii, jj, kk = rng
        for k in kk:
            for j in jj:
                cells = (k * self.ncy + j) * self.ncx + ii
                ok = cells[allowed[cells]]
                if len(ok):
                    return int(ok[0])
        return -1

    def grow(self, start, n_levels, allowed):
        """The cells of `start` and n_levels rings of cells that share a node with the set so
        far, each ring added at once and only cells with `allowed` set (Subdomain.cpp:
        140-227)."""
        inside = np.zeros(self.ncz * self.ncy * self.ncx, bool)
        inside[start] = True
        shape_c = (self.ncz, self.ncy, self.ncx)
        for _ in range(n_levels):
            node = np.zeros((self.nzn, self.nyn, self.nxn), bool)
            c = inside.reshape(shape_c)
            for bz in (0, 1):
                for by in (0, 1):
                    for bx in (0, 1):
                        node[bz:bz + self.ncz, by:by + self.ncy, bx:bx + self.ncx] |= c
            touch = np.zeros(shape_c, bool)
            for bz in (0, 1):
                for by in (0, 1):
                    for bx in (0, 1):
                        touch |= node[bz:bz + self.ncz, by:by + self.ncy, bx:bx + self.ncx]
            inside |= touch.ravel() & allowed
        return np.flatnonzero(inside)


class _Projection:
    """One smoothing subdomain: its cells, nodal mass matrix M (factorised), load matrix
    B[i, e] = integral of phi_i over cell e (the row sums of the local mass matrix,
    Subdomain.cpp:500-506, 595-623), and the trilinear weights W (nodes x receivers) of its
    receivers in the cells that hold them."""

    def __init__(self, grid: "_Grid", rec, recv_cells, cells):
        import scipy.sparse.linalg as spl

        self.cells = cells
        nodes = grid.cell_nodes(cells)                       # (nc, 8) global
        _, loc = np.unique(nodes, return_inverse=True)
        loc = loc.reshape(nodes.shape)
        nn = int(loc.max()) + 1
        hs = grid.cell_sizes(cells)
        Me = np.stack([_trilinear_mass(h) for h in hs])      # (nc, 8, 8)
        M = sp.csc_matrix((Me.ravel(), (np.repeat(loc, 8, axis=1).ravel(),
                                        np.tile(loc, (1, 8)).ravel())), shape=(nn, nn))
        self.lu = spl.splu(M)
        vol8 = np.prod(hs, axis=1) / 8.0
        self.B = sp.csr_matrix((np.repeat(vol8, 8), (loc.ravel(),
                                                     np.repeat(np.arange(len(cells)), 8))),
                               shape=(nn, len(cells)))
        pos = np.full(grid.ncx * grid.ncy * grid.ncz, -1, np.int64)
        pos[cells] = np.arange(len(cells))
        m = grid.m
        W = np.zeros((nn, len(rec)))
        for r, (p, c) in enumerate(zip(rec, recv_cells)):
            q = pos[int(c)]
            k, rr = divmod(int(c), grid.ncx * grid.ncy)
            j, i = divmod(rr, grid.ncx)
            t = np.clip([(p[0] - m.x[i]) / hs[q, 0], (p[1] - m.y[j]) / hs[q, 1],
                         (p[2] - m.z[k]) / hs[q, 2]], 0.0, 1.0)
            for l in range(8):
                b = [(l >> d) & 1 for d in range(3)]
                W[loc[q, l], r] += np.prod([t[d] if b[d] else 1.0 - t[d] for d in range(3)])
        self.W = W

    def solve(self, f):
        """M^-1 f for real or complex f (M is real)."""
        if np.iscomplexobj(f):
            return self.lu.solve(np.ascontiguousarray(f.real)) + \
                1j * self.lu.solve(np.ascontiguousarray(f.imag))
        return self.lu.solve(np.ascontiguousarray(f))


class _SmoothedMap:
    """x -> the smoothed values at the receivers, rows[p][:, comp] for piece p: a sum over
    subdomain pieces (projection, centre-value matrices C_comp (cells x edges)). Applied as
    W^T M^-1 B C x per piece and component, the reference's order of operations (project,
    then interpolate, OutputResultant3d.cpp:149-197), never as explicit rows (M^-1 is dense
    over a subdomain). `T` is the transposed map, for adjoint right-hand sides."""

    def __init__(self, n_out, nb, pieces, scale=1.0):
        self.shape = (n_out, nb)
        self.pieces = pieces
        self.scale = scale

    def __matmul__(self, x):
        x = np.asarray(x)
        out = np.zeros((self.shape[0],) + x.shape[1:], np.result_type(x, np.float64))
        for proj, Cs, rows in self.pieces:
            for comp, C in enumerate(Cs):
                out[rows[:, comp]] = proj.W.T @ proj.solve(proj.B @ (C @ x))
        return out * self.scale

    def __mul__(self, s):
        return _SmoothedMap(self.shape[0], self.shape[1], self.pieces, self.scale * s)

    def __truediv__(self, s):
        return self * (1.0 / s)

    @property
    def T(self):
        return _SmoothedMapT(self)


class _SmoothedMapT:
    def __init__(self, f: _SmoothedMap):
        self.f = f
        self.shape = f.shape[::-1]

    def __matmul__(self, y):
        y = np.asarray(y)
        out = np.zeros((self.shape[0],) + y.shape[1:], np.result_type(y, np.float64))
        for proj, Cs, rows in self.f.pieces:
            for comp, C in enumerate(Cs):
                out += C.T @ (proj.B.T @ proj.solve(proj.W @ y[rows[:, comp]]))
        return out * self.f.scale


class ReceiverOperator:
    """Sparse linear maps from the anomalous edge solution to E_x, E_y, H_x, H_y at receivers
    on the surface z = 0 (rows: receiver r, component order x, y), and H_z (one row per
    receiver). smoothing = n_neib rings of the reference's L2 smoothing (default 5), or None
    for the point evaluation."""

    def __init__(self, m: Model, receivers: np.ndarray, smoothing: int | None = 5):
        self.m = m
        rec = np.atleast_2d(np.asarray(receivers, dtype=np.float64))
        self.receivers = rec
        self.smoothing = smoothing
        ma = mesh_arrays(m)
        ed = ma.ed.reshape(ma.kpar, 25)[:, :12]
        ncx, ncy, ncz = m.shape_cells
        iz0 = int(np.argmin(np.abs(m.z)))
        if m.z[iz0] != 0.0 or iz0 == 0 or iz0 == len(m.z) - 1:
            raise ValueError("the mesh needs a node plane at z = 0 with cells above and below")
        if np.any(rec[:, 2] != 0.0):
            raise ValueError("receivers must lie on the surface z = 0")
        self.R = len(rec)
        if smoothing is None:
            self._point(m, rec, ed, iz0)
        else:
            self._smoothed(m, rec, ed, ma.nvkat, int(smoothing))

    def _smoothed(self, m, rec, ed, nvkat, n_neib):
        grid = _Grid(m)
        R = len(rec)
        nb = m.n_edges
        nc = grid.ncx * grid.ncy * grid.ncz

        def centre(cells, kind):
            """(cells x edges) matrices of the element-centre values of A (kind "A") or curl A,
            one per component (give_out_vec_mt.cpp:645-689, local point (0, 0, 0))."""
            hs = grid.cell_sizes(cells)
            sizes, which = np.unique(hs, axis=0, return_inverse=True)
            per = np.stack([(lambda pc: (pc[0] if kind == "A" else pc[1]).T)(
                _basis_at(h, (0.0, 0.0), 0.0)) for h in sizes])          # (nsizes, 3, 12)
            vals = np.transpose(per[which.ravel()], (1, 0, 2))             # (3, nc, 12)
            rows = np.repeat(np.arange(len(cells)), 12)
            return [sp.csr_matrix((vals[c].ravel(), (rows, ed[cells].ravel())),
                                  shape=(len(cells), nb)) for c in range(3)]

        # H: one subdomain over every material (vtWithoutDiscontinuity)
        every = np.ones(nc, bool)
        hc = np.array([grid.first_cell_holding(p, every) for p in rec])
        if np.any(hc < 0):
            raise ValueError(f"no element holds receiver(s) {np.flatnonzero(hc < 0).tolist()}")
        sub = grid.grow(hc, n_neib, every)
        proj = _Projection(grid, rec, hc, sub)
        Cs = centre(sub, "curl")
        r = np.arange(R)
        self.RH = _SmoothedMap(2 * R, nb, [(proj, Cs[:2], np.stack([2 * r, 2 * r + 1], 1))],
                               1.0 / MU_0)
        self.RHz = _SmoothedMap(R, nb, [(proj, Cs[2:], r[:, None])], 1.0 / MU_0)
        # E: one subdomain per material of the receivers' elements, air (material 0) excluded
        earth = nvkat != 0
        ec = np.array([grid.first_cell_holding(p, earth) for p in rec])
        if np.any(ec < 0):
            raise ValueError(f"no earth element holds receiver(s) "
                             f"{np.flatnonzero(ec < 0).tolist()}")
        pieces = []
        for mat in np.unique(nvkat[ec]):
            sel = np.flatnonzero(nvkat[ec] == mat)
            sub = grid.grow(ec[sel], n_neib, nvkat == mat)
            pieces.append((_Projection(grid, rec[sel], ec[sel], sub), centre(sub, "A")[:2],
                           np.stack([2 * sel, 2 * sel + 1], 1)))
        self.RE = _SmoothedMap(2 * R, nb, pieces)

    def _point(self, m, rec, ed, iz0):
        ncx, ncy, ncz = m.shape_cells
        nb = m.n_edges
        R = len(rec)
        Ee = [[], [], []]
        He = [[], [], []]
        Hz = [[], [], []]
        for r, (xr, yr, _) in enumerate(rec):
            cx = int(np.clip(np.searchsorted(m.x, xr, side="right") - 1, 0, ncx - 1))
            cy = int(np.clip(np.searchsorted(m.y, yr, side="right") - 1, 0, ncy - 1))
            h = np.array([m.x[cx + 1] - m.x[cx], m.y[cy + 1] - m.y[cy], 0.0])
            xi = (2 * (xr - m.x[cx]) / h[0] - 1, 2 * (yr - m.y[cy]) / h[1] - 1)
            for cz, face, wgt, what in [(iz0 - 1, 1.0, 1.0, "E"), (iz0 - 1, 1.0, 0.5, "H"),
                                        (iz0, -1.0, 0.5, "H")]:
                h[2] = m.z[cz + 1] - m.z[cz]
                phi, curl = _basis_at(h, xi, face)
                cell = (cz * ncy + cy) * ncx + cx
                edges = ed[cell]
                for comp in (0, 1):
                    vals = (phi if what == "E" else curl)[:, comp] * wgt
                    nz = vals != 0
                    tgt = Ee if what == "E" else He
                    tgt[0].append(np.full(nz.sum(), 2 * r + comp))
                    tgt[1].append(edges[nz])
                    tgt[2].append(vals[nz])
                if what == "H":  # H_z, continuous across the surface: the same mean
                    vals = curl[:, 2] * wgt
                    nz = vals != 0
                    Hz[0].append(np.full(nz.sum(), r))
                    Hz[1].append(edges[nz])
                    Hz[2].append(vals[nz])
        mk = lambda t: sp.csr_matrix((np.concatenate(t[2]), (np.concatenate(t[0]),
                                                             np.concatenate(t[1]))),
                                     shape=(2 * R, nb))
        self.RE = mk(Ee)            # E-shape (the vector potential, before -i*omega)
        self.RH = mk(He) / MU_0     # curl / mu0
        self.RHz = sp.csr_matrix((np.concatenate(Hz[2]), (np.concatenate(Hz[0]),
                                                          np.concatenate(Hz[1]))),
                                 shape=(R, nb)) / MU_0

    def fields(self, x: np.ndarray, nu: float, alpha: int):
        """E (R,2) and H (R,2) complex for the anomalous solution x (complex, nb) of one
        polarization (alpha = 1: X, 0: Y), normal field included."""
        w = omega_of(nu)
        z1, us, uc = harm1d(self.m, nu)
        a0 = us[-1] + 1j * uc[-1]            # A0 at z = 0
        A = (self.RE @ x).reshape(self.R, 2)
        Hc = (self.RH @ x).reshape(self.R, 2)
        A = A.copy()
        if alpha == 1:
            A[:, 0] += a0
            Hc[:, 1] += 1.0                  # normal H = (0, 1): (1/mu) A0'(0) = 1
        else:
            A[:, 1] += a0
            Hc[:, 0] += -1.0                 # normal H = (-1, 0)
        return -1j * w * A, Hc


    def hz(self, x: np.ndarray) -> np.ndarray:
        """H_z (R,) complex at the receivers for the anomalous solution x of one polarization
        (the normal field contributes none)."""
        return self.RHz @ x


def tipper(Hz_x_pol, H_x_pol, Hz_y_pol, H_y_pol):
    """T (R, 2) = (Tzx, Tzy) from the two polarizations: Hz = T H (columns = polarizations),
    the system calcimpedance.for:294-324 solves with the inverse it formed for Z."""
    Hz = np.stack([Hz_x_pol, Hz_y_pol], axis=1)[:, None, :]   # (R, 1, polarization)
    H = np.stack([H_x_pol, H_y_pol], axis=2)
    return (Hz @ np.linalg.inv(H))[:, 0, :]


def impedance(E_x_pol, H_x_pol, E_y_pol, H_y_pol):
    """Z (R,2,2) from the fields of the X and Y polarizations: E = Z H (columns = polarizations)."""
    E = np.stack([E_x_pol, E_y_pol], axis=2)   # (R, component, polarization)
    H = np.stack([H_x_pol, H_y_pol], axis=2)
    return E @ np.linalg.inv(H)


def apparent_resistivity(Z, nu):
    """rho_a = |Z|^2 / (omega mu0) and phase = -arg Z in degrees (calcimpedance.for:231-239)."""
    w = omega_of(nu)
    return np.abs(Z) ** 2 / (w * MU_0), -np.angle(Z, deg=True)
