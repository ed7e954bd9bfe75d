"""Device locality order of the edges (host side, no GPU): a permutation; the identity when one
pencil covers the mesh (the reference's numbering is lexicographic in (z, y, x, direction));
pencils in order otherwise; and a refusal for cells that are not structured hexahedra."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def mesh(M, small_model):
    return M.mesh_arrays(small_model), small_model


def test_one_pencil_is_the_reference_numbering(mesh):
    from mtcalc_with_dfp_cocr_amd.cocr import mesh_order

    ma, m = mesh
    o = mesh_order(ma, block=1 << 20)
    np.testing.assert_array_equal(o, np.arange(len(ma.edges) // 2))


@pytest.mark.parametrize("block", [1, 2, 3, 16])
def test_pencil_order(mesh, block):
    from mtcalc_with_dfp_cocr_amd.cocr import mesh_order

    ma, m = mesh
    o = mesh_order(ma, block=block)
    n = len(ma.edges) // 2
    assert np.array_equal(np.sort(o), np.arange(n))
    # integer node coordinates of the lexicographic node numbering
    nxn, nyn = len(m.x), len(m.y)
    e = ma.edges.reshape(-1, 2)[o]
    lo = np.minimum(e[:, 0], e[:, 1])
    x, y, z = lo % nxn, (lo // nxn) % nyn, lo // (nxn * nyn)
    d = np.where(np.abs(e[:, 1] - e[:, 0]) == 1, 0, np.where(np.abs(e[:, 1] - e[:, 0]) == nxn, 1, 2))
    key = np.stack([x // block, y // block, z, y, x, d], axis=1)
    # rows are in lexicographic key order
    diff = np.diff(key, axis=0)
    first_nonzero = np.argmax(diff != 0, axis=1)
    assert np.all(diff[np.arange(len(diff)), first_nonzero] > 0)


def test_not_hexahedral_refused(mesh):
    from mtcalc_with_dfp_cocr_amd._lib import lib

    ma, m = mesh
    nver = ma.nver.copy()
    nver[[1, 2]] = nver[[2, 1]]  # swap two vertices of cell 0
    order = np.empty(len(ma.edges) // 2, np.int32)
    rc = lib().mtcocr_mesh_order(int(ma.kpar), int(ma.kuzlov), nver.astype(np.int32),
                                 len(order), ma.edges.astype(np.int32), 16, order)
    assert rc == -10


def test_grid_order(mesh):
    """The grid order of a box mesh: the x-edges, then the y- and the z-edges, each set in
    (z, y, x) lexicographic order of its lower node, a permutation of the edges."""
    from mtcalc_with_dfp_cocr_amd.cocr import ORDER_GRID, mesh_order

    ma, m = mesh
    o = mesh_order(ma, block=ORDER_GRID)
    n = len(ma.edges) // 2
    assert np.array_equal(np.sort(o), np.arange(n))
    nxn, nyn, nzn = len(m.x), len(m.y), len(m.z)
    e = ma.edges.reshape(-1, 2)[o]
    lo = np.minimum(e[:, 0], e[:, 1])
    x, y, z = lo % nxn, (lo // nxn) % nyn, lo // (nxn * nyn)
    d = np.where(np.abs(e[:, 1] - e[:, 0]) == 1, 0, np.where(np.abs(e[:, 1] - e[:, 0]) == nxn, 1, 2))
    counts = [(nxn - 1) * nyn * nzn, nxn * (nyn - 1) * nzn, nxn * nyn * (nzn - 1)]
    assert [int((d == k).sum()) for k in range(3)] == counts
    key = np.stack([d, z, y, x], axis=1)
    diff = np.diff(key, axis=0)
    first_nonzero = np.argmax(diff != 0, axis=1)
    assert np.all(diff[np.arange(len(diff)), first_nonzero] > 0)
    # a mesh that is not one full box (an edge dropped) is refused
    from mtcalc_with_dfp_cocr_amd._lib import lib

    order = np.empty(n - 1, np.int32)
    rc = lib().mtcocr_mesh_order(int(ma.kpar), int(ma.kuzlov), ma.nver.astype(np.int32), n - 1,
                                 ma.edges[:-2].astype(np.int32), ORDER_GRID, order)
    assert rc == -10
