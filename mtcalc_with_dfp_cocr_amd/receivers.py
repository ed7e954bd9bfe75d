"""Receiver fields and the impedance tensor in FP64.

Restates what follows the solver in the reference, for the quantities its results are compared by:
  fields   E = -i*omega*(A + A0), H = curl(A + A0)/mu0 at the receivers
           (Code/OutputNu/give_out_vec_mt.cpp:101-272: E from the vector potential, H from its
           curl, the normal field of Harm1D added back; Harm1D writes the normal H as mu*alpha /
           mu*(alpha-1) at z >= 0, Code/Harm1D/divgrad_1d.cpp:283-297)
  impedance  Z = E H^-1 from the two polarizations, rho_a = |Z|^2/(omega*mu0), phase = -arg Z
           (Code/calcimpedance/calcimpedance.for:123-255 solves the same 2x2 complex system in
           its 4x4 real form, single precision; here FP64)
The reference evaluates the fields after an L2 smoothing of A and curl A onto trilinear nodes
around each receiver (MKL PARDISO, Code/OutputNu/OutputResultant3d.cpp:82-197). This module
evaluates the edge basis at the receiver point instead: E from the cell below the surface
(tangential E is continuous), H as the mean of the curls of the cells below and above. The
operator is linear in the solution vector, so Z comparisons between solvers measure solver
parity; absolute Z differs from the reference's smoothed values at the discretisation level.
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


class ReceiverOperator:
    """Sparse linear maps from the anomalous edge solution to E_x, E_y, H_x, H_y at receivers
    on the surface z = 0 (rows: receiver r, component order x, y)."""

    def __init__(self, m: Model, receivers: np.ndarray):
        self.m = m
        rec = np.atleast_2d(np.asarray(receivers, dtype=np.float64))
        self.receivers = rec
        ma = mesh_arrays(m)
        ed = ma.ed.reshape(ma.kpar, 25)[:, :12]
        ncx, ncy, ncz = m.shape_cells
        iz0 = int(np.argmin(np.abs(m.z)))
        if m.z[iz0] != 0.0 or iz0 == 0 or iz0 == len(m.z) - 1:
            raise ValueError("the mesh needs a node plane at z = 0 with cells above and below")
        if np.any(rec[:, 2] != 0.0):
            raise ValueError("receivers must lie on the surface z = 0")
        nb = m.n_edges
        R = len(rec)
        rows_e, rows_h = ([], [], []), ([], [], [])
        Ee = [[], [], []]
        He = [[], [], []]
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
        mk = lambda t: sp.csr_matrix((np.concatenate(t[2]), (np.concatenate(t[0]),
                                                             np.concatenate(t[1]))),
                                     shape=(2 * R, nb))
        self.RE = mk(Ee)            # E-shape (the vector potential, before -i*omega)
        self.RH = mk(He) / MU_0     # curl / mu0
        self.R = R

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


def impedance(E_x_pol, H_x_pol, E_y_pol, H_y_pol):
    """Z (R,2,2) from the fields of the X and Y polarizations: E = Z H (columns = polarizations)."""
    E = np.stack([E_x_pol, E_y_pol], axis=2)   # (R, component, polarization)
    H = np.stack([H_x_pol, H_y_pol], axis=2)
    return E @ np.linalg.inv(H)


def apparent_resistivity(Z, nu):
    """rho_a = |Z|^2 / (omega mu0) and phase = -arg Z in degrees (calcimpedance.for:231-239)."""
    w = omega_of(nu)
    return np.abs(Z) ** 2 / (w * MU_0), -np.angle(Z, deg=True)
