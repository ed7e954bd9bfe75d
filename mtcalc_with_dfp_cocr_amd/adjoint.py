"""Adjoint gradient of an impedance misfit with respect to the per-cell log-conductivity.

The reference models only the forward problem. This module adds its gradient, with the GPU
solver doing both the forward and the adjoint solves (all frequencies and polarizations in one
batch each).

Problem. For each frequency f (omega = 2 pi f) and polarization p in {X, Y}:
    A(sigma, omega) x_p = b_p(sigma),   A = K/mu + i omega S(sigma),  S = sum_c sigma_c M_c,
    b_p = -i omega sum_c (sigma_c - sigma0_c) M_c a0_p      (interior rows; Dirichlet rows 0)
with M_c the cell mass matrices and a0_p the edge DOFs of the 1-D normal field (model.cpp:
mtmodel_assemble_ks / mtmodel_assemble_rhs / mtmodel_normal_dofs). Receiver fields are linear
in x_p (receivers.ReceiverOperator): E = -i omega (RE x_p + E0), H = RH x_p + H0, and
Z = E H^-1 (columns = polarizations). The misfit over receivers r and the chosen components
ij (Zxy, Zyx by default) is
    J = sum_f sum_r sum_ij |Z_ij - O_ij|^2 / |O_ij|^2.

Gradient. With W_ij = conj(Z_ij - O_ij) / |O_ij|^2 on the chosen components and
Q_r = H_r^-1 W_r^T, dJ = 2 Re sum_p g_p . dx_p with
    g_p = sum_r sum_c [ -i omega Q_r[p, c] RE[2r+c, :] - (Q_r Z_r)[p, c] RH[2r+c, :] ].
A is complex symmetric (A^T = A), so the adjoint state solves A lambda_p = g_p with the same
solver, and
    dJ/dsigma_c = 2 Re sum_f sum_p [ -i omega lambda_p^T M_c (x_p + a0_p) ]_c,
    dJ/dlog(sigma_c) = sigma_c dJ/dsigma_c.
lambda vanishes on Dirichlet edges (identity rows, zero right-hand side there).
"""
from __future__ import annotations

import copy
import time
from dataclasses import dataclass

import numpy as np

from . import model as M
from ._lib import check, lib
from .receivers import ReceiverOperator, impedance

COMPONENTS = ((0, 1), (1, 0))  # Zxy, Zyx


def with_sigma(m: M.Model, sigma: np.ndarray) -> M.Model:
    """The model with per-cell conductivity `sigma` taken as is (no rounding)."""
    m2 = copy.copy(m)
    m2.sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    return m2


def normal_dofs(m: M.Model, nu: float, alpha: int) -> np.ndarray:
    """Edge DOFs a0 (complex, nb) of the normal field the right-hand side is built from."""
    z1, us, uc = M.harm1d(m, nu)
    out = np.empty(2 * m.n_edges)
    nxn, nyn, nzn = m.shape_nodes
    check(lib().mtmodel_normal_dofs(nxn, nyn, nzn, m.x, m.y, m.z, int(alpha), len(z1), z1, us,
                                    uc, out), "mtmodel_normal_dofs")
    return out[0::2] + 1j * out[1::2]


def boundary_edges(m: M.Model) -> np.ndarray:
    """Dirichlet edges: both end nodes on the outer boundary (model.cpp is_boundary_edge,
    Code/BuildMatrix/bound_cond_vec_harm.cpp:81-108)."""
    ma = M.mesh_arrays(m)
    e = ma.edges.reshape(-1, 2)
    b = ma.is_node_bound.astype(bool)
    return b[e[:, 0]] & b[e[:, 1]]


def _ks_matrices(m: M.Model):
    """K and S (full symmetric, host FP64) from the model's K/S arrays (mtmodel_assemble_ks:
    lower triangle + diagonal, Dirichlet rows at identity in K)."""
    import scipy.sparse as sp

    p, kd, kg, sd, sg = M.assemble_ks(m)
    rows = np.repeat(np.arange(p.nb), np.diff(p.ig))

    def sym(di, gg):
        L = sp.csr_matrix((gg, (rows, p.jg)), shape=(p.nb, p.nb))
        return (L + L.T + sp.diags(di)).tocsr()

    return sym(kd, kg), sym(sd, sg)


def s_matrix(m: M.Model):
    """S(sigma), the sigma-weighted edge mass matrix as the solver sees it (host FP64)."""
    return _ks_matrices(m)[1]


def host_operator(m: M.Model):
    """(omegas, X) -> rows A(omega_s) x_s = K x_s + i omega_s S x_s on the host in FP64."""
    K, S = _ks_matrices(m)

    def apply(omegas, X):
        X = np.atleast_2d(X)
        return (K @ X.T).T + 1j * np.asarray(omegas)[:, None] * (S @ X.T).T

    return apply


def cell_mass_forms(m: M.Model, u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """u_c^T M_c v_c per cell (complex, unconjugated), u and v complex edge vectors."""
    out = np.empty(2 * m.n_cells)
    nxn, nyn, nzn = m.shape_nodes
    check(lib().mtmodel_cell_mass_forms(nxn, nyn, nzn, m.x, m.y, m.z, M.to_interleaved(u),
                                        M.to_interleaved(v), out), "mtmodel_cell_mass_forms")
    return out[0::2] + 1j * out[1::2]


class GpuBatchSolver:
    """Solves A(sigma, omega_s) x_s = b_s for a batch of systems on the GPU (K/S form), keeping
    the handle, the mesh order and the gradient mesh across conductivity updates."""

    def __init__(self, m: M.Model, eps: float = 1e-10, maxiter: int = 40000, device: int = 0,
                 order_block: int = 16):
        from .cocr import Solver

        self.m = m
        self.eps = eps
        self.maxiter = maxiter
        self.ma = M.mesh_arrays(m)
        self.S = Solver(device)
        if order_block:
            self.S.set_mesh_order(self.ma, order_block)
        self._sigma = None
        self.stats = []

    def set_sigma(self, sigma: np.ndarray):
        if self._sigma is not None and np.array_equal(sigma, self._sigma):
            return
        p, kd, kg, sd, sg = M.assemble_ks(with_sigma(self.m, sigma))
        self.S.set_matrix_ks(p, kd, kg, sd, sg)
        if self._sigma is None:
            self.S.set_gradient(self.ma)
        self._sigma = np.array(sigma, copy=True)

    def solve(self, omegas: np.ndarray, rhs: np.ndarray, eps: float | None = None) -> np.ndarray:
        """rhs: (nsys, nb) complex. Returns (nsys, nb) complex, solved to eps (default: the
        solver's). Each right-hand side is solved at unit norm and the solution scaled back:
        the stop test is relative, but the COCR breakdown test on (a, w) is absolute (1e-40,
        rci.cpp:38) and would fire near a tight tolerance on a small right-hand side."""
        eps = self.eps if eps is None else eps
        out = np.empty_like(rhs)
        scale = np.linalg.norm(rhs, axis=1)
        scale[scale == 0] = 1.0
        rhs = rhs / scale[:, None]
        for k0 in range(0, len(omegas), 64):
            k1 = min(len(omegas), k0 + 64)
            prs = np.stack([M.to_interleaved(b) for b in rhs[k0:k1]])
            x, st = self.S.solve_batch(np.asarray(omegas[k0:k1]), prs, eps=eps,
                                       maxiter=self.maxiter, precision="fp64")
            self.stats.extend(st)
            # a breakdown of the (a, w) recurrence once the residual is within 10x the
            # tolerance is convergence at the attainable FP64 level
            bad = [q for q in st if q["req"] != 0 and
                   not (q["req"] == -2 and q["residual_rel"] < 10 * eps)]
            if bad:
                raise RuntimeError(f"batched solve did not converge: {bad[0]}")
            out[k0:k1] = x[:, 0::2] + 1j * x[:, 1::2]
        return out * scale[:, None]

    def sigma_gradient(self, m: M.Model, coef: np.ndarray, u: np.ndarray,
                       v: np.ndarray) -> np.ndarray:
        """sum_s Re(coef_s * u_s^T M_c v_s) per cell, contracted on the device."""
        return self.S.sigma_gradient(m.shape_nodes, m.x, m.y, m.z, coef, u, v)

    def close(self):
        self.S.close()


@dataclass
class State:
    sigma: np.ndarray
    x: np.ndarray          # (nfreq, 2, nb) complex, polarizations X, Y
    a0: np.ndarray         # (nfreq, 2, nb)
    Z: np.ndarray          # (nfreq, R, 2, 2)
    H: np.ndarray          # (nfreq, R, 2, 2)
    J: float


class ImpedanceMisfit:
    """J(sigma) and its adjoint gradient for a model, frequencies, surface receivers and
    observed impedances. `solver` has set_sigma(sigma) and solve(omegas, rhs); the default is
    the GPU batched solver."""

    def __init__(self, m: M.Model, freqs, receivers, z_obs=None, solver=None,
                 components=COMPONENTS, eps: float = 1e-10):
        self.m = m
        self.freqs = np.asarray(freqs, dtype=np.float64)
        self.omegas = np.array([M.omega_of(f) for f in self.freqs])
        self.R = ReceiverOperator(m, receivers)
        self.components = tuple(components)
        self.solver = solver if solver is not None else GpuBatchSolver(m, eps=eps)
        self.a0 = np.stack([np.stack([normal_dofs(m, f, 1), normal_dofs(m, f, 0)])
                            for f in self.freqs])
        self.interior = ~boundary_edges(m)
        self.refine_eps = 1e-10
        self.z_obs = z_obs
        self.timings = {}

    # ---- forward ---------------------------------------------------------------------------
    def responses(self, sigma: np.ndarray, base: "State | None" = None):
        """Solutions (nfreq, 2, nb), impedances (nfreq, R, 2, 2) and H (nfreq, R, 2, 2).

        With base (the state at a nearby sigma_b), each solution is refined out of base.x:
        x = base.x + d with A(sigma) d = (b(sigma) - b(sigma_b)) - i omega (S(sigma) -
        S(sigma_b)) base.x, solved to eps relative to that right-hand side. The right-hand
        side is exactly local to the cells where sigma differs (assembled entries elsewhere are
        bitwise equal), and x solves A(sigma) x = b(sigma) - r_b, r_b the base's own residual,
        which two nearby sigma share. A finite difference of J at a step of 1e-4 needs this:
        solved from scratch, each x holds ~eps |b| of independent Krylov noise, J ~1e-11 J of
        it, and (J+ - J-) / 2h a floor near 1e-7 whatever the cell's sensitivity. d is solved
        to refine_eps (1e-10): its relative error multiplies J(sigma) - J(sigma_b), of order
        |sigma - sigma_b| |dJ|, so a central difference carries ~1e-10 of the cell's own
        derivative from it."""
        t0 = time.perf_counter()
        m = with_sigma(self.m, sigma)
        rhs = self._rhs(m)
        t1 = time.perf_counter()
        nf = len(self.freqs)
        self.solver.set_sigma(sigma)
        om = np.repeat(self.omegas, 2)
        if base is None:
            x = self.solver.solve(om, rhs).reshape(nf, 2, -1)
        else:
            x0 = np.asarray(base.x).reshape(nf * 2, -1)
            rb, Sb = self._base_parts(base)
            dS = s_matrix(m) - Sb
            defect = (rhs - rb) - 1j * om[:, None] * (dS @ x0.T).T
            x = (x0 + self.solver.solve(om, defect, eps=self.refine_eps)).reshape(nf, 2, -1)
        t2 = time.perf_counter()
        Z = np.empty((nf, self.R.R, 2, 2), complex)
        H = np.empty_like(Z)
        for k, f in enumerate(self.freqs):
            Ex, Hx = self.R.fields(x[k, 0], f, 1)
            Ey, Hy = self.R.fields(x[k, 1], f, 0)
            Z[k] = impedance(Ex, Hx, Ey, Hy)
            H[k] = np.stack([Hx, Hy], axis=2)
        self.timings.update(forward_assembly_s=t1 - t0, forward_solve_s=t2 - t1)
        return x, Z, H

    def _rhs(self, m: M.Model) -> np.ndarray:
        nf = len(self.freqs)
        rhs = np.empty((nf * 2, self.m.n_edges), complex)
        for k, f in enumerate(self.freqs):
            for p, alpha in enumerate((1, 0)):
                pr = M.assemble_rhs(m, f, alpha)
                rhs[2 * k + p] = pr[0::2] + 1j * pr[1::2]
        return rhs

    def _base_parts(self, base: "State"):
        """b(sigma_b) and S(sigma_b) of a base state, kept for the last base used."""
        key = getattr(self, "_base_key", None)
        if key is None or key[0] is not base:
            mb = with_sigma(self.m, base.sigma)
            self._base_key = (base, self._rhs(mb), s_matrix(mb))
        return self._base_key[1], self._base_key[2]

    def weights(self):
        O = self.z_obs
        W = np.zeros(O.shape)
        for i, j in self.components:
            W[..., i, j] = 1.0 / np.abs(O[..., i, j]) ** 2
        return W

    def evaluate(self, sigma: np.ndarray, base: State | None = None) -> State:
        x, Z, H = self.responses(sigma, base)
        W = self.weights()
        J = float(np.sum(W * np.abs(Z - self.z_obs) ** 2))
        return State(np.array(sigma, copy=True), x, self.a0, Z, H, J)

    # ---- adjoint ---------------------------------------------------------------------------
    def adjoint_rhs(self, st: State) -> np.ndarray:
        """g_p (nfreq, 2, nb): dJ = 2 Re sum g_p . dx_p."""
        nf, R = len(self.freqs), self.R.R
        W = self.weights() * np.conj(st.Z - self.z_obs)          # (nf, R, 2, 2)
        g = np.zeros((nf, 2, self.m.n_edges), complex)
        RE, RH = self.R.RE, self.R.RH
        for k in range(nf):
            w = self.omegas[k]
            Hinv = np.linalg.inv(st.H[k])                          # (R, 2, 2)
            Q = Hinv @ np.transpose(W[k], (0, 2, 1))               # Q[r] = H^-1 W^T
            QZ = Q @ st.Z[k]
            for p in range(2):
                # coefficients on the rows 2r + c of RE and RH
                ce = (-1j * w) * Q[:, p, :].reshape(-1)            # (2R,)
                ch = -QZ[:, p, :].reshape(-1)
                g[k, p] = RE.T @ ce + RH.T @ ch
        # x is held at 0 on the Dirichlet edges, so J varies only with the interior ones: the
        # receiver maps can read boundary edges (a smoothing subdomain that reaches the outer
        # boundary), and their part of g would enter lambda through the identity rows
        return g * self.interior

    def gradient(self, st: State) -> np.ndarray:
        """dJ/dsigma per cell."""
        t0 = time.perf_counter()
        g = self.adjoint_rhs(st)
        nf = len(self.freqs)
        self.solver.set_sigma(st.sigma)
        lam = self.solver.solve(np.repeat(self.omegas, 2), g.reshape(nf * 2, -1)).reshape(
            nf, 2, -1)
        t1 = time.perf_counter()
        # dJ/dsigma_c = sum over systems of 2 Re[-i omega lambda^T M_c (x + a0)]: on the device
        # with the GPU solver, else (exact reference solvers of the tests) on the host
        coef = np.repeat(-2j * self.omegas, 2)
        contract = getattr(self.solver, "sigma_gradient", None)
        if contract is not None:
            grad = contract(self.m, coef, lam.reshape(nf * 2, -1),
                            (st.x + st.a0).reshape(nf * 2, -1))
        else:
            grad = np.zeros(self.m.n_cells)
            for k in range(nf):
                for p in range(2):
                    form = cell_mass_forms(self.m, lam[k, p], st.x[k, p] + st.a0[k, p])
                    grad += np.real(coef[2 * k + p] * form)
        t2 = time.perf_counter()
        self.timings.update(adjoint_solve_s=t1 - t0, gradient_assembly_s=t2 - t1)
        return grad

    def misfit(self, sigma: np.ndarray, base: State | None = None) -> float:
        return self.evaluate(sigma, base).J


def torch_misfit(problem: ImpedanceMisfit):
    """J(log sigma) as a torch function with the adjoint gradient as its backward."""
    import torch

    class _Misfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, log_sigma):
            sigma = np.exp(log_sigma.detach().cpu().numpy().astype(np.float64))
            st = problem.evaluate(sigma)
            ctx.state = st
            return torch.tensor(st.J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            st = ctx.state
            g = problem.gradient(st) * st.sigma               # d/dlog(sigma)
            return gJ * torch.from_numpy(g)

    return _Misfit.apply


# ------------------------------------------------------------------------------------------
# config 4: 64x64x40 seeded smoothed log-sigma, Zxy / Zyx on a 15x15 surface grid,
# 6 frequencies logspace(-2, 2), observations from the model perturbed by 0.1 N(0,1) (smoothed)

def config4(seed: int = 4, ncx: int = 64, ncy: int = 64, ncz: int = 40, nrec: int = 15,
            nfreq: int = 6):
    m = M.random_model(ncx, ncy, ncz, seed=seed, h=50.0, name=f"c4_{ncx}x{ncy}x{ncz}")
    xs, ys = m.x, m.y
    core = lambda a: (a[8], a[-9])                  # inside the padding
    gx = np.linspace(*core(xs), nrec + 2)[1:-1]
    gy = np.linspace(*core(ys), nrec + 2)[1:-1]
    rx, ry = np.meshgrid(gx, gy, indexing="ij")
    rec = np.stack([rx.ravel(), ry.ravel(), np.zeros(rx.size)], axis=1)
    freqs = M.frequencies(1e-2, 1e2, nfreq)
    # observed data: log10(rho) + 0.1 N(0, 1), smoothed, on the earth cells
    rng = np.random.default_rng(seed + 1000)
    earth = m.sigma > M.AIR_SIGMA * 10
    logrho = -np.log10(m.sigma)
    pert = np.zeros(m.n_cells)
    pert[earth] = 0.1 * rng.standard_normal(int(earth.sum()))
    pert = M._smooth3(pert.reshape(ncz, ncy, ncx)).ravel()
    sigma_true = np.where(earth, 10.0 ** -(logrho + pert), m.sigma)
    return m, rec, freqs, sigma_true


def fd_cells(m: M.Model, rec: np.ndarray, n: int = 20, seed: int = 4) -> np.ndarray:
    """n cells drawn with the seed among the earth cells under the receiver grid within 600 m
    of the surface (the cells the surface impedances see)."""
    area = fd_area(m, rec)
    return np.random.default_rng(seed).choice(area, size=min(n, len(area)), replace=False)


def fd_area(m: M.Model, rec: np.ndarray) -> np.ndarray:
    xc, yc, zc = m.cell_centres()
    return np.flatnonzero((zc < 0) & (zc > -600) & (np.abs(xc) < np.abs(rec[:, 0]).max())
                          & (np.abs(yc) < np.abs(rec[:, 1]).max()))


def _fd_check(problem: ImpedanceMisfit, sigma: np.ndarray, cells, h: float,
              base: State | None = None):
    """Central differences of J in log(sigma_c) for the given cells. With base (the state at
    sigma), J(sigma +- h e_c) comes from solutions refined out of base.x
    (ImpedanceMisfit.responses)."""
    out = []
    for c in cells:
        s1 = sigma.copy()
        s2 = sigma.copy()
        s1[c] *= np.exp(h)
        s2[c] *= np.exp(-h)
        out.append((problem.misfit(s1, base) - problem.misfit(s2, base)) / (2 * h))
    return np.array(out)


def _fd_direction(problem: ImpedanceMisfit, sigma: np.ndarray, d: np.ndarray, h: float,
                  base: State | None = None):
    """Central difference of J along the log(sigma) direction d (per cell)."""
    return (problem.misfit(sigma * np.exp(h * d), base)
            - problem.misfit(sigma * np.exp(-h * d), base)) / (2 * h)


def main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="adjoint gradient of the impedance misfit (config 4)"
                                 " checked against central finite differences")
    ap.add_argument("--ncx", type=int, default=64)
    ap.add_argument("--ncy", type=int, default=64)
    ap.add_argument("--ncz", type=int, default=40)
    ap.add_argument("--nrec", type=int, default=15)
    ap.add_argument("--nfreq", type=int, default=6)
    ap.add_argument("--fd-cells", type=int, default=20)
    ap.add_argument("--h", type=float, default=1e-4)
    ap.add_argument("--eps", type=float, default=1e-10)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--directions", type=int, default=2,
                    help="random log(sigma) directions checked as a whole")
    a = ap.parse_args(argv)
    m, rec, freqs, sigma_true = config4(a.seed, a.ncx, a.ncy, a.ncz, a.nrec, a.nfreq)
    t0 = time.perf_counter()
    prob = ImpedanceMisfit(m, freqs, rec, eps=a.eps)
    _, Zobs, _ = prob.responses(sigma_true)
    prob.z_obs = Zobs
    sigma = m.sigma.copy()
    t1 = time.perf_counter()
    st = prob.evaluate(sigma)
    t2 = time.perf_counter()
    grad = prob.gradient(st) * sigma
    t3 = time.perf_counter()
    # FD cells drawn with the seed among the earth cells under the receiver area within 600 m
    # of the surface, each J(sigma +- h e_c) refined out of the solutions at sigma by defect
    # correction; then whole random +-1 directions over all those cells, whose derivative
    # gathers every cell's sensitivity.
    area = fd_area(m, rec)
    cells = fd_cells(m, rec, a.fd_cells, a.seed)
    fd = _fd_check(prob, sigma, cells, a.h, st)
    rng = np.random.default_rng(a.seed + 7)
    dirs = []
    for _ in range(a.directions):
        d = np.zeros_like(sigma)
        d[area] = rng.choice([-1.0, 1.0], size=len(area))
        fdd = _fd_direction(prob, sigma, d, a.h, st)
        add = float(grad @ d)
        dirs.append({"adjoint": add, "fd": fdd, "rel_err": abs(add - fdd) / abs(fdd)})
    t4 = time.perf_counter()
    ad = grad[cells]
    rel = np.abs(ad - fd) / np.maximum(np.abs(fd), 1e-300)
    out = {"workload": f"{a.ncx}x{a.ncy}x{a.ncz} cells, nb={m.n_edges}, {a.nfreq} freqs x 2 pol,"
                       f" {len(rec)} receivers, Zxy+Zyx",
           "J": st.J, "eps": a.eps, "h": a.h, "cells": cells.tolist(),
           "adjoint": ad.tolist(), "fd": fd.tolist(), "rel_err": rel.tolist(),
           "rel_err_max": float(rel.max()),
           "rel_err_norm": float(np.linalg.norm(ad - fd) / np.linalg.norm(fd)),
           "directions": dirs,
           "seconds": {"setup_and_observed": t1 - t0, "forward": t2 - t1, "gradient": t3 - t2,
                       "fd_total": t4 - t3}, "timings": prob.timings,
           "iterations": [q["iter"] for q in prob.solver.stats[:4 * a.nfreq]]}
    print(json.dumps(out))
    prob.solver.close()


if __name__ == "__main__":
    main()
