"""Independent numpy restatements used to pin the oracle and the GPU path."""
import numpy as np
import scipy.sparse as sp


def brute_force_g(ma, nb):
    """G (interior nodes x edges): -1/2 at an edge's first node, +1/2 at its second, interior
    nodes (not on the outer boundary) numbered in ascending order."""
    interior = np.flatnonzero(ma.is_node_bound == 0)
    renum = -np.ones(ma.kuzlov, np.int64)
    renum[interior] = np.arange(len(interior))
    e = ma.edges.reshape(-1, 2)
    rows, cols, vals = [], [], []
    for k, sgn in [(0, -0.5), (1, 0.5)]:
        r = renum[e[:, k]]
        ok = r >= 0
        rows.append(r[ok])
        cols.append(np.flatnonzero(ok))
        vals.append(np.full(ok.sum(), sgn))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(len(interior), nb))


def dense_preconditioner(A, G):
    """M^-1 = diag(A)^-1 + G^T diag(G A G^T)^-1 G (closed form of the folded preconditioner)."""
    da = 1.0 / A.diagonal()
    GAGt = (G @ A @ G.T)
    dn = 1.0 / GAGt.diagonal()
    Gc = G.tocsr()

    def apply(v):
        return da * v + Gc.T @ (dn * (Gc @ v))

    return apply, da, dn


def textbook_pcocr(A, minv, b, iters, residuals=False):
    """Preconditioned COCR for complex-symmetric A (unconjugated bilinear forms):
    u = M^-1 r, z = A u, p = u, q = z;  alpha = (u,z)/(q, M^-1 q);  x += alpha p;
    r -= alpha q;  u -= alpha M^-1 q;  beta = (u,Au)_new/(u,Au)_old;  p = u + beta p;
    q = Au + beta q.  Returns the iterates x_1..x_iters (and, with residuals=True, also the
    recurrence residuals r_0..r_iters)."""
    x = np.zeros_like(b)
    r = b.copy()
    u = minv(r)
    z = A @ u
    p, q = u.copy(), z.copy()
    uz = u @ z
    xs = []
    rs = [r.copy()]
    for _ in range(iters):
        mq = minv(q)
        alpha = uz / (q @ mq)
        x = x + alpha * p
        r = r - alpha * q
        u = u - alpha * mq
        z = A @ u
        uz_new = u @ z
        beta = uz_new / uz
        uz = uz_new
        p = u + beta * p
        q = z + beta * q
        xs.append(x.copy())
        rs.append(r.copy())
    return (xs, rs) if residuals else xs


def mrs_residuals(rs, clamp=True, copy=True):
    """Minimal residual smoothing of a residual sequence r_0, r_1, ... taken as real vectors
    (re and im parts side by side): s_0 = r_0; for k >= 1, u = r_k - s, w = -(s.u)/(u.u)
    (0 when u.u < 1e-100); w < 0 leaves s unchanged (clamp), w > 1 restarts s = r_k (copy),
    otherwise s += w u. Returns the norms ||s_k|| and the unclamped weights w_k. The flags
    switch the two clamps off, to show that a check against this sequence sees them."""
    real = lambda v: np.concatenate([v.real, v.imag])
    s = real(rs[0])
    norms, ws = [np.linalg.norm(s)], []
    for rk in rs[1:]:
        r = real(rk)
        u = r - s
        uu = u @ u
        w = 0.0 if abs(uu) < 1e-100 else -(s @ u) / uu
        ws.append(w)
        if clamp and w < 0:
            pass
        elif copy and w > 1:
            s = r.copy()
        else:
            s = s + w * u
        norms.append(np.linalg.norm(s))
    return np.array(norms), np.array(ws)


def identity_system(M):
    """A 1-cell mesh (12 edges, every node on the boundary, so G is empty) with A = I and
    b = (1, i, 0, ...): the first preconditioned bilinear form (a, M^-1 a) = 1 + i^2 = 0."""
    m = M.Model(np.array([0.0, 1.0]), np.array([0.0, 1.0]), np.array([0.0, 1.0]),
                np.array([1.0]), np.array([1.0]), np.array([0.0]), np.array([1.0]))
    ma = M.mesh_arrays(m)
    nb = 12
    di = np.tile([1.0, 0.0], nb)
    pr = np.zeros(2 * nb)
    pr[0] = 1.0
    pr[3] = 1.0
    s = M.BlockSystem(nb, np.zeros(nb + 1, np.int32), np.zeros(0, np.int32),
                      np.arange(0, 2 * nb + 1, 2, dtype=np.int32), np.zeros(1, np.int32), di,
                      np.zeros(0), pr)
    return m, ma, s


class DirectKSSolver:
    """Sparse LU solves of A(omega) = K + i*omega*S assembled for a conductivity (the exact
    reference for gradient checks; scipy's SuperLU, no iteration)."""

    def __init__(self, M, m):
        self.M, self.m = M, m
        self._sigma = None

    def set_sigma(self, sigma):
        import scipy.sparse as sp

        from mtcalc_with_dfp_cocr_amd.adjoint import with_sigma

        if self._sigma is not None and np.array_equal(sigma, self._sigma):
            return
        p, kd, kg, sd, sg = self.M.assemble_ks(with_sigma(self.m, sigma))
        rows = np.repeat(np.arange(p.nb), np.diff(p.ig))
        full = lambda d, g: (sp.csr_matrix((g, (rows, p.jg)), shape=(p.nb, p.nb))
                             + sp.csr_matrix((g, (p.jg, rows)), shape=(p.nb, p.nb))
                             + sp.diags(d))
        self.K, self.S = full(kd, kg), full(sd, sg)
        self._lu = {}
        self._sigma = np.array(sigma, copy=True)

    def solve(self, omegas, rhs, eps=None):
        import scipy.sparse.linalg as spl

        out = np.empty_like(rhs)
        for k, w in enumerate(omegas):
            if w not in self._lu:
                self._lu[w] = spl.splu((self.K + 1j * w * self.S).tocsc())
            out[k] = self._lu[w].solve(rhs[k])
        return out
