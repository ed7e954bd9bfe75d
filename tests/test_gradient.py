"""BuildGMatrix on the host (no GPU): the library's discrete gradient against a brute-force
construction and against the oracle, on a conforming mesh and on one with hanging nodes.

FoldedPreconditioner::BuildGMatrix (Code/COCR_FP/FoldedPreconditioner.cpp:68-259) keeps the
nodes of conductive cells that are not on the boundary, numbers them in ascending order, and
gives edge e = (v0, v1) the gradient row -1/2 at v0 and +1/2 at v1. A hanging node h (number
>= n_nodes_c) of a non-conforming mesh is spread onto the conforming nodes of its T-matrix row
(HarmVect3d.cpp:255-301): -1/2 * T[h, k] at each k. The brute force forms G^T = (edge-node
incidence) @ T densely and restricts it to the kept nodes.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import brute_force_g


def hanging_mesh(M, ma):
    """A 2:1 refinement pattern on the top face of the mesh's centre cell: its 4 edge midpoints
    (T rows of two nodes at 1/2) and its face centre (four nodes at 1/4) become hanging nodes,
    joined to each other and to the face corners by the edges of the refined neighbour above.
    Returns the new edges array, n_nodes_c and the T-matrix (ig_t over all nodes, jg_t, gg_t)."""
    nver = ma.nver.reshape(-1, 14)
    c = len(nver) // 2
    corner = nver[c, 4:8]  # the cell's top face (vertices 4..7: z = 1 in the local order)
    a, b, cc, d = (int(v) for v in corner)  # (0,0), (1,0), (0,1), (1,1) in x-y
    nc = int(ma.kuzlov)
    mids = [(a, b), (cc, d), (a, cc), (b, d)]
    rows = [[(u, 0.5), (v, 0.5)] for u, v in mids] + [[(a, .25), (b, .25), (cc, .25), (d, .25)]]
    h = [nc + k for k in range(len(rows))]  # hanging node numbers
    ctr = h[4]
    new = [(a, h[0]), (h[0], b), (cc, h[1]), (h[1], d), (a, h[2]), (h[2], cc), (b, h[3]),
           (h[3], d), (h[0], ctr), (ctr, h[1]), (h[2], ctr), (ctr, h[3])]
    edges = np.concatenate([ma.edges.reshape(-1, 2), np.array(new)]).astype(np.int32)
    ig_t = np.zeros(nc + len(rows) + 1, np.int32)
    jg_t, gg_t = [], []
    for k, r in enumerate(rows):
        ig_t[nc + k + 1] = ig_t[nc + k] + len(r)
        for v, w in r:
            jg_t.append(v)
            gg_t.append(w)
    return edges, nc, ig_t, np.array(jg_t, np.int32), np.array(gg_t)


def brute_force_gt(ma, edges, nc, ig_t, jg_t, gg_t):
    ntot = len(ig_t) - 1
    T = np.zeros((ntot, nc))
    T[np.arange(nc), np.arange(nc)] = 1.0
    for v in range(nc, ntot):
        for q in range(ig_t[v], ig_t[v + 1]):
            T[v, jg_t[q]] += gg_t[q]
    inc = np.zeros((len(edges), ntot))
    inc[np.arange(len(edges)), edges[:, 0]] -= 0.5
    inc[np.arange(len(edges)), edges[:, 1]] += 0.5
    keep = np.zeros(nc, bool)
    nver = ma.nver.reshape(-1, 14)[:, :8]
    for cell, mat in enumerate(ma.nvkat):
        if ma.sig_flags[mat] > 0:
            for v in nver[cell]:
                if v < nc and not ma.is_node_bound[v]:
                    keep[v] = True
    return (inc @ T)[:, keep]


def csr_dense(ptr, col, val, shape):
    D = np.zeros(shape)
    for r in range(shape[0]):
        for q in range(ptr[r], ptr[r + 1]):
            D[r, col[q]] += val[q]
    return D


def oracle_g(O, ma, edges, nc, ig_t, jg_t, gg_t):
    p = O.OrcDfp()
    vp = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    rc = O.lib().orc_build_g(C.byref(p), int(ma.kpar), len(edges), nc,
                             np.ascontiguousarray(ma.nvkat, np.int32),
                             np.ascontiguousarray(ma.nver, np.int32),
                             np.ascontiguousarray(ma.sig_flags, np.float64),
                             np.ascontiguousarray(edges.ravel(), np.int32), vp(ig_t), vp(jg_t),
                             vp(gg_t), np.ascontiguousarray(ma.is_node_bound, np.int32))
    assert rc == 0
    n, m = len(edges), p.m
    it = np.ctypeslib.as_array(p.iptr_gt, (n + 1,)).copy()
    jt = np.ctypeslib.as_array(p.jptr_gt, (max(it[-1], 1),))[:it[-1]].copy()
    gt = np.ctypeslib.as_array(p.gg_gt, (max(it[-1], 1),))[:it[-1]].copy()
    ig = np.ctypeslib.as_array(p.iptr_g, (m + 1,)).copy()
    jg = np.ctypeslib.as_array(p.jptr_g, (max(ig[-1], 1),))[:ig[-1]].copy()
    gg = np.ctypeslib.as_array(p.gg_g, (max(ig[-1], 1),))[:ig[-1]].copy()
    O.lib().orc_dfp_free(C.byref(p))
    return (it, jt, gt), (ig, jg, gg)


@pytest.fixture(scope="module")
def conforming(M):
    m = M.random_model(4, 3, 3, seed=2, h=100.0, n_pad=1, n_air=2, name="g")
    return M.mesh_arrays(m)


def test_conforming_gradient(M, conforming):
    from mtcalc_with_dfp_cocr_amd.cocr import build_gmatrix

    ma = conforming
    (it, jt, gt), (ip, jp, gp) = build_gmatrix(ma)
    nb = len(ma.edges) // 2
    G = brute_force_g(ma, nb).toarray()
    np.testing.assert_array_equal(csr_dense(ip, jp, gp, G.shape), G)
    np.testing.assert_array_equal(csr_dense(it, jt, gt, G.T.shape), G.T)
    assert np.all(np.diff(ip) <= 6)  # a conforming hexahedral node has at most 6 edges


def test_hanging_node_gradient(O, M, conforming):
    """The T-matrix branch of BuildGMatrix: the same G^T and G as the brute force (dense, with
    the reference's duplicate entries summed) and the same CSR arrays, entry for entry, as the
    oracle's restatement; nodes next to the refined face carry more than 6 entries."""
    from mtcalc_with_dfp_cocr_amd.cocr import build_gmatrix

    ma = conforming
    edges, nc, ig_t, jg_t, gg_t = hanging_mesh(M, ma)

    class Mesh:  # the mesh arrays with the refined neighbour's edges
        pass

    mh = Mesh()
    for k in ("kpar", "nvkat", "nver", "sig_flags"):
        setattr(mh, k, getattr(ma, k))
    mh.edges = edges.ravel()
    mh.kuzlov = len(ig_t) - 1
    mh.is_node_bound = np.concatenate([ma.is_node_bound, np.zeros(len(ig_t) - 1 - nc, np.int32)])
    (it, jt, gt), (ip, jp, gp) = build_gmatrix(mh, ig_t, jg_t, gg_t, n_nodes_c=nc)
    Gt = brute_force_gt(mh, edges, nc, ig_t, jg_t, gg_t)
    np.testing.assert_allclose(csr_dense(it, jt, gt, Gt.shape), Gt, rtol=0, atol=1e-15)
    np.testing.assert_allclose(csr_dense(ip, jp, gp, Gt.T.shape), Gt.T, rtol=0, atol=1e-15)
    (oit, ojt, ogt), (oip, ojp, ogp) = oracle_g(O, mh, edges, nc, ig_t, jg_t, gg_t)
    for a, b in ((it, oit), (jt, ojt), (gt, ogt), (ip, oip), (jp, ojp), (gp, ogp)):
        np.testing.assert_array_equal(a, b)
    assert np.diff(ip).max() > 6 and np.diff(ip).max() <= 32
    # a T-matrix row naming a node outside the conforming range is refused
    bad = jg_t.copy()
    bad[0] = len(ig_t)
    with pytest.raises(RuntimeError):
        build_gmatrix(mh, ig_t, bad, gg_t, n_nodes_c=nc)
