"""Regular-mesh MT models, their edge-element systems and the reference's file formats.

Models follow the benchmark configurations of this project (regular hexahedral meshes with
geometrically growing padding and air cells, layered or seeded random conductivity). The heavy
parts (portrait, element assembly, 1-D normal field) run in the native library
(include/mtmodel.h); this module builds the per-cell conductivity arrays and reads / writes the
working-directory files that the reference's COCR_FP process consumes
(Code/COCR_FP/HarmVect3d.cpp:36-336, Code/RegularMeshBuilder/main.cpp:1301-2141).

Material numbering of the written files follows the reference's mesh builder
(Code/RegularMeshBuilder/main.cpp:769-878): material 1 is air, materials 2..L+1 are the 1-D
layers from the deepest to the top one (the order Harm1D reads from sig3d, Code/Harm1D/
mesh_1d_mtz.cpp:55-80), further materials are the distinct (sigma, sigma0) pairs of the other
cells. Conductivities are rounded through the "%.7e" text of sig3d so that a model solved in
memory and the same model read back from its files are the same system.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from ._lib import check, lib, ptr

MU_0 = 1.25663706143591729e-6  # Code/COCR_FP/stdafx.h:31 (mu3d is read in units of MU_0)
PI = 3.1415926535897932384626433832795
AIR_SIGMA = 1e-8


def round_sig(a: np.ndarray) -> np.ndarray:
    """Round through the "%.7e" format of sig3d (RegularMeshBuilder/main.cpp:1555)."""
    a = np.asarray(a, dtype=np.float64)
    flat = np.char.mod("%.7e", a.ravel()).astype(np.float64)
    return flat.reshape(a.shape)


def padded_axis(h: float, n_core: int, n_pad: int, growth: float = 1.4) -> np.ndarray:
    """Node coordinates of a symmetric axis: n_core cells of size h centred on 0, n_pad cells
    on each side growing by `growth` (sizes h*growth**k, k = 1..n_pad)."""
    half = n_core * h / 2.0
    core = -half + h * np.arange(n_core + 1)
    steps = h * growth ** np.arange(1, n_pad + 1)
    right = half + np.cumsum(steps)
    left = -half - np.cumsum(steps)
    return np.concatenate([left[::-1], core, right])


def z_axis(h: float, n_air: int, n_core: int, n_pad: int, growth: float = 1.4) -> np.ndarray:
    """Vertical nodes (ascending): n_pad padding cells below the core, n_core earth cells of
    size h below z = 0, n_air air cells above z = 0, padding and air growing by `growth`."""
    core = -n_core * h + h * np.arange(n_core + 1)
    below = -n_core * h - np.cumsum(h * growth ** np.arange(1, n_pad + 1))
    air = np.cumsum(h * growth ** np.arange(1, n_air + 1))
    z = np.concatenate([below[::-1], core, air])
    z[n_pad + n_core] = 0.0
    return z


@dataclass
class Model:
    """Regular mesh + per-cell conductivities (cells ordered x fastest, then y, then z)."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    sigma: np.ndarray       # per cell
    sigma0: np.ndarray      # per cell, the layered normal (background) conductivity
    layer_top: np.ndarray   # 1-D layers, deepest first; the last top is 0
    layer_sigma: np.ndarray
    receivers: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    name: str = "model"

    @property
    def shape_nodes(self) -> tuple[int, int, int]:
        return len(self.x), len(self.y), len(self.z)

    @property
    def shape_cells(self) -> tuple[int, int, int]:
        return len(self.x) - 1, len(self.y) - 1, len(self.z) - 1

    @property
    def n_cells(self) -> int:
        a, b, c = self.shape_cells
        return a * b * c

    @property
    def n_edges(self) -> int:
        return int(lib().mtmodel_num_edges(*self.shape_nodes))

    @property
    def mu(self) -> np.ndarray:
        return np.full(self.sigma.shape, 1.0 * MU_0)

    def cell_centres(self):
        xc = 0.5 * (self.x[1:] + self.x[:-1])
        yc = 0.5 * (self.y[1:] + self.y[:-1])
        zc = 0.5 * (self.z[1:] + self.z[:-1])
        Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
        return X.ravel(), Y.ravel(), Z.ravel()


def _layered(zc: np.ndarray, tops_bottom_first: np.ndarray, sig_bottom_first: np.ndarray):
    """Conductivity of the layer containing each depth; layer i spans (top[i-1], top[i]] and
    the deepest layer extends down. Cells above the top layer get NaN (air is set by the
    caller)."""
    out = np.full_like(zc, np.nan)
    for t, s in zip(tops_bottom_first[::-1], sig_bottom_first[::-1]):
        out[zc < t] = s  # deeper layers overwrite the cells below their top
    return out


def snap_to_nodes(tops: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Move layer interfaces to the nearest mesh node, so that the 1-D normal-field model and
    the cell-centre assignment of sigma0 describe the same layered medium."""
    return np.array([z[np.argmin(np.abs(z - t))] for t in tops])


def config1_model() -> Model:
    """24x24x24 cells, 100 m core, 8 padding cells per side (x1.4), 6 air cells; layers
    100/10/1000 Ohm*m (0-500 m, 500-1500 m, half-space) with a 1 Ohm*m block of 600x600x400 m
    whose top is 300 m deep; 5x5 receivers at 100 m spacing on the surface."""
    h = 100.0
    x = padded_axis(h, 8, 8)
    y = padded_axis(h, 8, 8)
    z = z_axis(h, 6, 10, 8)
    tops = snap_to_nodes(np.array([-1500.0, -500.0, 0.0]), z)
    sig_layers = np.array([1.0 / 1000.0, 1.0 / 10.0, 1.0 / 100.0])
    xc, yc, zc = Model(x, y, z, np.zeros(1), np.zeros(1), tops, sig_layers).cell_centres()
    sigma0 = _layered(zc, tops, sig_layers)
    sigma0[zc > 0] = AIR_SIGMA
    sigma = sigma0.copy()
    block = (np.abs(xc) < 300) & (np.abs(yc) < 300) & (zc < -300) & (zc > -700)
    sigma[block] = 1.0
    g = np.arange(-200.0, 201.0, 100.0)
    rx, ry = np.meshgrid(g, g, indexing="ij")
    rec = np.stack([rx.ravel(), ry.ravel(), np.zeros(rx.size)], axis=1)
    return _finish(Model(x, y, z, sigma, sigma0, tops, sig_layers, rec, name="c1_24x24x24"))


def _smooth3(a: np.ndarray) -> np.ndarray:
    """Separable Gaussian smoothing with a 3-cell kernel (std 1 cell), edges replicated."""
    w = np.exp(-0.5 * np.array([1.0, 0.0, 1.0]))
    w /= w.sum()
    for ax in range(a.ndim):
        p = np.pad(a, [(1, 1) if k == ax else (0, 0) for k in range(a.ndim)], mode="edge")
        sl = [slice(None)] * a.ndim
        acc = 0.0
        for k in range(3):
            sl[ax] = slice(k, k + a.shape[ax])
            acc = acc + w[k] * p[tuple(sl)]
        a = acc
    return a


def random_model(ncx: int, ncy: int, ncz: int, seed: int, h: float = 50.0, n_pad: int = 8,
                 n_air: int = 6, name: str | None = None) -> Model:
    """Seeded random model: log10(rho) ~ U[0, 3] i.i.d. per earth cell, smoothed with a 3-cell
    Gaussian, clipped to [0, 3]. The normal conductivity sigma0 of every earth cell is the
    geometric mean of its z-layer (each earth cell layer is one 1-D layer)."""
    x = padded_axis(h, ncx - 2 * n_pad, n_pad)
    y = padded_axis(h, ncy - 2 * n_pad, n_pad)
    n_core_z = ncz - n_air - n_pad
    z = z_axis(h, n_air, n_core_z, n_pad)
    n_earth = n_pad + n_core_z
    rng = np.random.default_rng(seed)
    logrho = rng.uniform(0.0, 3.0, size=(n_earth, ncy, ncx))
    logrho = np.clip(_smooth3(logrho), 0.0, 3.0)
    sig_e = 10.0 ** (-logrho)
    layer_sig = 10.0 ** (-logrho.reshape(n_earth, -1).mean(axis=1))
    sigma = np.full((ncz, ncy, ncx), AIR_SIGMA)
    sigma0 = np.full((ncz, ncy, ncx), AIR_SIGMA)
    sigma[:n_earth] = sig_e
    sigma0[:n_earth] = layer_sig[:, None, None]
    tops = z[1:n_earth + 1].copy()  # top of every earth cell layer, deepest first
    tops[-1] = 0.0
    return _finish(Model(x, y, z, sigma.ravel(), sigma0.ravel(), tops, layer_sig,
                         name=name or f"rand{ncx}x{ncy}x{ncz}_s{seed}"))


def ellipsoid_model(ncx: int, ncy: int, ncz: int, seed: int, n_ellipsoids: int = 10,
                    h: float = 50.0, n_pad: int = 8, n_air: int = 6) -> Model:
    """Seeded layered background (5 layers, 10-1000 Ohm*m) with random conductive / resistive
    ellipsoids of 1-1000 Ohm*m."""
    x = padded_axis(h, ncx - 2 * n_pad, n_pad)
    y = padded_axis(h, ncy - 2 * n_pad, n_pad)
    n_core_z = ncz - n_air - n_pad
    z = z_axis(h, n_air, n_core_z, n_pad)
    rng = np.random.default_rng(seed)
    depth = n_core_z * h
    tops = np.sort(-rng.uniform(0.1, 0.9, size=4) * depth)
    tops = snap_to_nodes(np.concatenate([tops, [0.0]]), z)
    layer_sig = 1.0 / (10.0 ** rng.uniform(1.0, 3.0, size=tops.size))
    xc, yc, zc = Model(x, y, z, np.zeros(1), np.zeros(1), tops, layer_sig).cell_centres()
    sigma0 = _layered(zc, tops, layer_sig)
    sigma0[zc > 0] = AIR_SIGMA
    sigma = sigma0.copy()
    hx = (ncx - 2 * n_pad) * h / 2
    hy = (ncy - 2 * n_pad) * h / 2
    for _ in range(n_ellipsoids):
        c = np.array([rng.uniform(-0.7, 0.7) * hx, rng.uniform(-0.7, 0.7) * hy,
                      -rng.uniform(0.1, 0.8) * depth])
        r = rng.uniform(0.05, 0.25, size=3) * np.array([hx, hy, depth])
        inside = ((xc - c[0]) / r[0]) ** 2 + ((yc - c[1]) / r[1]) ** 2 + ((zc - c[2]) / r[2]) ** 2 < 1
        sigma[inside & (zc < 0)] = 1.0 / 10.0 ** rng.uniform(0.0, 3.0)
    return _finish(Model(x, y, z, sigma, sigma0, tops, layer_sig,
                         name=f"ellip{ncx}x{ncy}x{ncz}_s{seed}"))


# ------------------------------------------------------------------------------------------
# the regular mesh builder's axes and box objects (the reference's example input format)


@dataclass
class MeshSettings:
    """The first four lines of the mesh builder's settings.cfg (RegularMeshBuilder/main.cpp:
    175-190, 1253-1264): start step, sparse (growth) factor, gap between the receivers' box and
    the sparse part, and distance from the receivers' box to the domain boundary, each for XY
    (both horizontal directions) and Z; eps decides when an inserted line replaces a near one."""

    step_xy: float
    step_z: float
    sparse_xy: float
    sparse_z: float
    gap_xy: float
    gap_z: float
    far_xy: float
    far_z: float
    eps: float = 0.1

    def per_axis(self):
        return [(self.step_xy, self.sparse_xy, self.gap_xy, self.far_xy)] * 2 + \
               [(self.step_z, self.sparse_z, self.gap_z, self.far_z)]


def read_settings(path) -> MeshSettings:
    """Four pairs of numbers at the top of settings.cfg (main.cpp:1253-1264); the rest of the
    file (SLAE eps, iteration cap, notes) is not part of the mesh."""
    v = [float(s) for s in Path(path).read_text().split()[:8]]
    return MeshSettings(*v)


def read_objects(path) -> np.ndarray:
    """objects: a count, then x0 x1 y0 y1 z0 z1 sigma per box (main.cpp:1116-1145)."""
    tok = Path(path).read_text().split()
    n = int(tok[0])
    return np.array([float(s) for s in tok[1:1 + 7 * n]]).reshape(n, 7)


def read_layers(path) -> np.ndarray:
    """z_sig_2d: a count, then (z, sigma) per layer, deepest first (main.cpp:1169-1202)."""
    tok = Path(path).read_text().split()
    n = int(tok[0])
    return np.array([float(s) for s in tok[1:1 + 2 * n]]).reshape(n, 2)


def read_receivers(path) -> np.ndarray:
    """xyzVectorB: a count, then x y z per receiver (main.cpp:1219-1240)."""
    tok = Path(path).read_text().split()
    n = int(tok[0])
    return np.array([float(s) for s in tok[1:1 + 3 * n]]).reshape(n, 3)


def _nearest_line(axis: list, value: float) -> int:
    """Index of the line nearest to value by bisection, the upper one on a tie; the end lines
    for values outside (BinarySearchInVectorSorted, main.cpp:241-265)."""
    if value < axis[0]:
        return 0
    if value > axis[-1]:
        return len(axis) - 1
    lo, hi = 0, len(axis) - 1
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if value > axis[mid]:
            lo = mid
        else:
            hi = mid
    return lo if abs(axis[lo] - value) < abs(axis[hi] - value) else hi


def _insert_line(axis: list, value: float, eps: float) -> None:
    """A line at `value` (InsertCoordinateIntoSortedArray, main.cpp:487-512): the nearest line
    moves there when it lies within eps of the local step, otherwise a line is added; values
    outside the axis are ignored."""
    if value < axis[0] or value > axis[-1]:
        return
    k = _nearest_line(axis, value)
    nbr = k - 1 if (axis[k] > value and k != 0) or k == len(axis) - 1 else k + 1
    step = abs(axis[k] - axis[nbr])
    if abs(axis[k] - value) / step > eps:
        axis.insert(k if axis[k] > value else k + 1, value)
    else:
        axis[k] = value


def _uniform(step: float, lo: float, hi: float) -> list:
    """Build1DUniformMesh (main.cpp:268-284): the step shrunk to divide [lo, hi] evenly."""
    n = int((hi - lo) / step + 0.5)
    if abs(step * n - (hi - lo)) > 1e-5:
        n += 1
    h = (hi - lo) / n
    return [lo + i * h for i in range(n + 1)]


def _sparse(step: float, growth: float, lo: float, hi: float, outward_up: bool) -> list:
    """Build1DSparsedMesh (main.cpp:296-340): steps growing by `growth` away from the uniform
    part, until the boundary is passed."""
    out = []
    if outward_up:
        c = lo
        out.append(c)
        while True:
            c += step
            step *= growth
            out.append(c)
            if not c < hi:
                break
        return out
    c = hi
    out.append(c)
    while True:
        c -= step
        step *= growth
        out.append(c)
        if not c > lo:
            break
    return out[::-1]


def _round_down(value: float, dec: float) -> float:
    """RoundDouble_ (main.cpp:572-576): the multiple of dec at or below a positive value, one
    step below a negative one."""
    n = int(abs(value / dec))
    return -(n + 1) * dec if value < 0 else n * dec


def builder_axis(r_lo: float, r_hi: float, step: float, growth: float, gap: float,
                 far: float) -> list:
    """One axis of the regular mesh builder (CalculateModelBounds, BuildBaseMesh1D and
    AlignMeshBounds, main.cpp:344-376,441-462,579-683): uniform steps over the receivers' extent
    widened by `gap` (rounded to whole steps), growing steps out to `far` beyond the receivers'
    extent, the outer lines placed on the boundary and a last step under 0.3 of its neighbour
    merged into it."""
    lo, hi = r_lo - gap, r_hi + gap
    n = int((hi - lo) / step + 0.5)
    if abs(n * step - (hi - lo)) > 1e-7:
        n += 1
        lo = _round_down(lo, step)
        hi = lo + n * step
    b = [r_lo - far, lo, hi, r_hi + far]
    mid = _uniform(step, b[1], b[2])
    left = _sparse(step, growth, b[0], b[1], False)
    right = _sparse(step, growth, b[2], b[3], True)
    ax = left[:-1] + mid[:-1] + right
    ax[0] = b[0]
    if ax[1] - ax[0] < (ax[2] - ax[1]) * 0.3:
        del ax[1]
    ax[-1] = b[3]
    if ax[-1] - ax[-2] < (ax[-2] - ax[-3]) * 0.3:
        del ax[-2]
    return ax


def box_model(receivers: np.ndarray, settings: MeshSettings, objects: np.ndarray,
              layers: np.ndarray, name: str = "boxes") -> Model:
    """A model in the reference's example input format (BuildMesh, main.cpp:1086-1113): the
    builder's axes around the receivers, each box's six faces and each layer's top inserted as
    lines, a layered earth (z_sig_2d, deepest first) under air at 1e-8 S/m, boxes painted over
    it in file order. sigma0 is the layered earth (SetNormalMaterials, main.cpp:750-757).

    Layer lines: z_sig_2d's layer tops. (The reference derives each layer's lower line from a
    layer top it has not read yet, main.cpp:1193; for a file whose layers share one
    conductivity, as the example's do, only the mesh lines depend on that.)"""
    rec = np.asarray(receivers, dtype=np.float64)
    axes = []
    for d, (step, growth, gap, far) in enumerate(settings.per_axis()):
        axes.append(builder_axis(rec[:, d].min(), rec[:, d].max(), step, growth, gap, far))
    for o in objects:
        for f in range(6):
            _insert_line(axes[f // 2], float(o[f]), settings.eps)
    lay = np.asarray(layers, dtype=np.float64)
    for ztop in lay[:, 0]:
        _insert_line(axes[2], float(ztop), settings.eps)
    x, y, z = (np.array(a) for a in axes)
    tops = lay[:, 0].copy()
    tops[-1] = 0.0
    sig_layers = lay[:, 1].copy()
    xc, yc, zc = Model(x, y, z, np.zeros(1), np.zeros(1), tops, sig_layers).cell_centres()
    sigma0 = _layered(zc, tops, sig_layers)
    sigma0[zc > 0] = AIR_SIGMA
    sigma = sigma0.copy().reshape(len(z) - 1, len(y) - 1, len(x) - 1)
    for o in objects:
        ix = [_nearest_line(axes[0], o[0]), _nearest_line(axes[0], o[1])]
        iy = [_nearest_line(axes[1], o[2]), _nearest_line(axes[1], o[3])]
        iz = [_nearest_line(axes[2], o[4]), _nearest_line(axes[2], o[5])]
        sigma[iz[0]:iz[1], iy[0]:iy[1], ix[0]:ix[1]] = o[6]
    return _finish(Model(x, y, z, sigma.ravel(), sigma0, tops, sig_layers, rec, name=name))


def _finish(m: Model) -> Model:
    m.sigma = round_sig(m.sigma)
    m.sigma0 = round_sig(m.sigma0)
    m.layer_sigma = round_sig(m.layer_sigma)
    return m


# ------------------------------------------------------------------------------------------
# systems


@dataclass
class BlockSystem:
    """One (frequency, polarization) system in the reference's 0-based block layout."""

    nb: int
    ig: np.ndarray
    jg: np.ndarray
    idi: np.ndarray
    ijg: np.ndarray
    di: np.ndarray
    gg: np.ndarray
    pr: np.ndarray
    nu: float = 0.0
    alpha: int = 1


@dataclass
class Portrait:
    nb: int
    ig: np.ndarray
    jg: np.ndarray
    idi: np.ndarray
    ijg: np.ndarray


_portrait_cache: dict = {}


def portrait(m: Model) -> Portrait:
    key = m.shape_nodes
    if key in _portrait_cache:
        return _portrait_cache[key]
    L = lib()
    nxn, nyn, nzn = key
    nb = int(L.mtmodel_num_edges(nxn, nyn, nzn))
    nnz = C.c_int64(0)
    check(L.mtmodel_portrait_size(nxn, nyn, nzn, C.byref(nnz)), "mtmodel_portrait_size")
    ig = np.empty(nb + 1, np.int32)
    jg = np.empty(nnz.value, np.int32)
    idi = np.empty(nb + 1, np.int32)
    ijg = np.empty(nnz.value + 1, np.int32)
    check(L.mtmodel_portrait(nxn, nyn, nzn, ig, jg, idi, ijg), "mtmodel_portrait")
    p = Portrait(nb, ig, jg, idi, ijg)
    if len(_portrait_cache) > 4:
        _portrait_cache.clear()
    _portrait_cache[key] = p
    return p


def harm1d(m: Model, nu: float):
    """1-D normal field (z ascending, Re A0, Im A0) of the model's layers at frequency nu."""
    L = lib()
    tops = np.ascontiguousarray(m.layer_top, dtype=np.float64)
    sig = np.ascontiguousarray(m.layer_sigma, dtype=np.float64)
    n = C.c_int(0)
    L.mtmodel_harm1d(len(tops), tops, sig, float(nu), 0, C.byref(n), None, None, None)
    z1 = np.empty(n.value)
    us = np.empty(n.value)
    uc = np.empty(n.value)
    check(L.mtmodel_harm1d(len(tops), tops, sig, float(nu), n.value, C.byref(n), ptr(z1),
                           ptr(us), ptr(uc)), "mtmodel_harm1d")
    return z1, us, uc


def assemble(m: Model, nu: float, alpha: int = 1) -> BlockSystem:
    """System of frequency nu (Hz) and polarization alpha (1 = X, 0 = Y) with Dirichlet BCs."""
    L = lib()
    p = portrait(m)
    z1, us, uc = harm1d(m, nu)
    di = np.empty(int(p.idi[-1]))
    gg = np.empty(int(p.ijg[-1]))
    pr = np.empty(2 * p.nb)
    nxn, nyn, nzn = m.shape_nodes
    check(L.mtmodel_assemble(nxn, nyn, nzn, m.x, m.y, m.z, m.sigma, m.sigma0, m.mu, float(nu),
                             int(alpha), len(z1), z1, us, uc, p.ig, p.jg, p.idi, p.ijg, di, gg,
                             pr), "mtmodel_assemble")
    return BlockSystem(p.nb, p.ig, p.jg, p.idi, p.ijg, di, gg, pr, float(nu), int(alpha))


def assemble_rhs(m: Model, nu: float, alpha: int = 1) -> np.ndarray:
    """Right-hand side of (nu, alpha) alone (interleaved, 2*nb)."""
    z1, us, uc = harm1d(m, nu)
    pr = np.empty(2 * m.n_edges)
    nxn, nyn, nzn = m.shape_nodes
    check(lib().mtmodel_assemble_rhs(nxn, nyn, nzn, m.x, m.y, m.z, m.sigma, m.sigma0, float(nu),
                                     int(alpha), len(z1), z1, us, uc, pr), "mtmodel_assemble_rhs")
    return pr


def assemble_ks(m: Model):
    """Frequency-independent K (curl-curl/mu) and S (sigma-mass), one real per block."""
    L = lib()
    p = portrait(m)
    nnz = len(p.jg)
    kd, sd = np.empty(p.nb), np.empty(p.nb)
    kg, sg = np.empty(nnz), np.empty(nnz)
    nxn, nyn, nzn = m.shape_nodes
    check(L.mtmodel_assemble_ks(nxn, nyn, nzn, m.x, m.y, m.z, m.sigma, m.mu, p.ig, p.jg, kd, kg,
                                sd, sg), "mtmodel_assemble_ks")
    return p, kd, kg, sd, sg


def element_matrices(hx: float, hy: float, hz: float):
    b = np.empty(144)
    c = np.empty(144)
    check(lib().mtmodel_element_matrices(hx, hy, hz, b, c), "mtmodel_element_matrices")
    return b.reshape(12, 12), c.reshape(12, 12)


@dataclass
class MeshArrays:
    """0-based mesh arrays in the form COCR_FP passes to BuildGMatrix."""

    kpar: int
    kuzlov: int
    nver: np.ndarray          # kpar*14 (8 nodes then 6 unused)
    ed: np.ndarray            # kpar*25 (12 edges, 12 zeros, 1), 0-based like inputvec
    edges: np.ndarray         # nb*2
    l13d: np.ndarray
    is_node_bound: np.ndarray
    nvkat: np.ndarray
    sig_flags: np.ndarray     # per material: 0 if sigma == 0 else 1
    materials: np.ndarray     # per material (sigma, sigma0)
    xyz: np.ndarray


def mesh_arrays(m: Model) -> MeshArrays:
    L = lib()
    nxn, nyn, nzn = m.shape_nodes
    kpar = (nxn - 1) * (nyn - 1) * (nzn - 1)
    nb = int(L.mtmodel_num_edges(nxn, nyn, nzn))
    kt1 = C.c_int64(0)
    L.mtmodel_topology(nxn, nyn, nzn, None, None, None, None, C.byref(kt1))
    nver = np.empty(kpar * 14, np.int32)
    ed = np.empty(kpar * 25, np.int32)
    edges = np.empty(nb * 2, np.int32)
    l13d = np.empty(kt1.value, np.int32)
    check(L.mtmodel_topology(nxn, nyn, nzn, ptr(nver), ptr(ed), ptr(edges), ptr(l13d),
                             C.byref(kt1)), "mtmodel_topology")
    kuzlov = nxn * nyn * nzn
    # 1-based file arrays -> 0-based (the reference decrements after reading; unused slots -1)
    nver0 = nver - 1
    ed0 = ed - 1
    edges0 = edges - 1
    l130 = l13d - 1
    inb = np.zeros(kuzlov, np.int32)
    inb[l130] = 1
    mats, nvkat = material_table(m)
    flags = np.where(mats[:, 0] == 0.0, 0.0, 1.0)
    X, Y, Z = np.meshgrid(m.x, m.y, m.z, indexing="ij")
    xyz = np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(),
                    Z.transpose(2, 1, 0).ravel()], axis=1)
    return MeshArrays(kpar, kuzlov, nver0, ed0, edges0, l130, inb, nvkat, flags, mats,
                      np.ascontiguousarray(xyz))


def material_table(m: Model):
    """(materials[(sigma, sigma0)], nvkat 0-based): air, layers deepest first, then the
    remaining distinct pairs in order of first appearance."""
    base = np.array([(AIR_SIGMA, AIR_SIGMA)] + [(s, s) for s in m.layer_sigma], dtype=np.float64)
    pairs = np.stack([m.sigma, m.sigma0], axis=1)
    uniq, first, inv = np.unique(pairs, axis=0, return_index=True, return_inverse=True)
    inv = inv.ravel()
    ids = np.full(len(uniq), -1, np.int64)
    for i in range(len(base) - 1, -1, -1):  # the first matching base material wins
        ids[(uniq[:, 0] == base[i, 0]) & (uniq[:, 1] == base[i, 1])] = i
    extra = np.flatnonzero(ids < 0)
    extra = extra[np.argsort(first[extra], kind="stable")]  # order of first appearance
    ids[extra] = len(base) + np.arange(len(extra))
    mats = np.concatenate([base, uniq[extra]], axis=0)
    return mats, ids[inv].astype(np.int32)


# ------------------------------------------------------------------------------------------
# reference working-directory files


def _bin(path: Path, a, dtype):
    np.ascontiguousarray(a, dtype=dtype).tofile(path)


def write_cocr_fp_dir(path, m: Model, sys: BlockSystem, eps: float = 1e-8, maxiter: int = 10000,
                      nthreads: int = 1, mesh: MeshArrays | None = None) -> Path:
    """Write every file the COCR_FP process reads (1-based binaries, text control files)."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    ma = mesh or mesh_arrays(m)
    nb = sys.nb
    with open(path / "sig3d", "w") as f:
        for i, (s, s0) in enumerate(ma.materials):
            f.write(f"{i + 1}\t{s:.7e}\t{s0:.7e}\n")
    for name in ("mu3d", "dpr3d"):
        with open(path / name, "w") as f:
            v = 1.0 if name == "mu3d" else 0.0
            for i in range(len(ma.materials)):
                f.write(f"{i + 1}\t{v:.7e}\t{v:.7e}\n")
    with open(path / "inftry.dat", "w") as f:
        f.write(f" ISLAU= 0 INDKU1= 0 INDFPO= 0\nKUZLOV= {ma.kuzlov}   KPAR= {ma.kpar}"
                f"    KT1= {len(ma.l13d)}   KTR2= 0   KTR3= 0\n")
    _bin(path / "nver.dat", np.where(ma.nver >= 0, ma.nver + 1, 0), np.int32)
    _bin(path / "nvkat.dat", ma.nvkat + 1, np.int32)
    _bin(path / "xyz.dat", ma.xyz, np.float64)
    _bin(path / "l13d.dat", ma.l13d + 1, np.int32)
    _bin(path / "nodesforedges.dat", ma.edges + 1, np.int32)
    _bin(path / "edges.dat", ma.ed + 1, np.int32)
    (path / "tsize3d.dat").write_text("0\n")
    (path / "tsize3d_.dat").write_text(f"0\n{nb}\n")
    (path / "nthreads.txt").write_text(f"{nthreads}\n")
    (path / "kuslau").write_text(f"{2 * nb}\n{eps:e}\n{maxiter}\n")
    (path / "nu").write_text(f"{sys.nu:.13e}\n3")
    (path / "alfa").write_text(f"{sys.alpha}\n")
    with open(path / "sreda1d.ay", "w") as f:
        f.write(f"{len(m.layer_top)}\n")
        for t, s in zip(m.layer_top, m.layer_sigma):
            f.write(f"{t:.13e}\t{s:.13e}\t1\n")
    _bin(path / "pr", sys.pr, np.float64)
    _bin(path / "di", sys.di, np.float64)
    _bin(path / "gg", sys.gg, np.float64)
    _bin(path / "ig", sys.ig + 1, np.int32)
    _bin(path / "jg", sys.jg + 1, np.int32)
    _bin(path / "idi", sys.idi + 1, np.int32)
    _bin(path / "ijg", sys.ijg + 1, np.int32)
    return path


def read_block_system(path) -> BlockSystem:
    """Read the SLAE files of a working directory (0-based)."""
    path = Path(path)
    n, eps, maxiter = (path / "kuslau").read_text().split()[:3]
    nb = int(n) // 2
    ig = np.fromfile(path / "ig", np.int32) - 1
    idi = np.fromfile(path / "idi", np.int32) - 1
    jg = np.fromfile(path / "jg", np.int32) - 1
    ijg = np.fromfile(path / "ijg", np.int32) - 1
    di = np.fromfile(path / "di", np.float64)
    gg = np.fromfile(path / "gg", np.float64)
    pr = np.fromfile(path / "pr", np.float64)
    assert len(ig) == nb + 1 and len(pr) == 2 * nb
    nu = 0.0
    if (path / "nu").exists():
        nu = float((path / "nu").read_text().split()[0])
    return BlockSystem(nb, ig, jg, idi, ijg, di, gg, pr, nu)


def dense_matrix(sys: BlockSystem) -> np.ndarray:
    """Dense complex A = L + L^T + D (small systems only, for checks)."""
    nb = sys.nb
    A = np.zeros((nb, nb), np.complex128)
    for i in range(nb):
        b, sz = sys.idi[i], sys.idi[i + 1] - sys.idi[i]
        A[i, i] = sys.di[b] + (1j * sys.di[b + 1] if sz == 2 else 0)
        for k in range(sys.ig[i], sys.ig[i + 1]):
            j = sys.jg[k]
            b, sz = sys.ijg[k], sys.ijg[k + 1] - sys.ijg[k]
            v = sys.gg[b] + (1j * sys.gg[b + 1] if sz == 2 else 0)
            A[i, j] = v
            A[j, i] = v
    return A


def sparse_matrix(sys: BlockSystem):
    """scipy.sparse complex A = L + L^T + D."""
    import scipy.sparse as sp

    nb = sys.nb
    rows = np.repeat(np.arange(nb), np.diff(sys.ig))
    sz = np.diff(sys.ijg)
    re = sys.gg[sys.ijg[:-1]]
    im = np.where(sz == 2, sys.gg[np.minimum(sys.ijg[:-1] + 1, len(sys.gg) - 1)], 0.0)
    v = re + 1j * im
    dsz = np.diff(sys.idi)
    dv = sys.di[sys.idi[:-1]] + 1j * np.where(dsz == 2, sys.di[np.minimum(sys.idi[:-1] + 1, len(sys.di) - 1)], 0.0)
    L = sp.csr_matrix((v, (rows, sys.jg)), shape=(nb, nb))
    return (L + L.T + sp.diags(dv)).tocsr()


def to_complex(v: np.ndarray) -> np.ndarray:
    return v[0::2] + 1j * v[1::2]


def to_interleaved(z: np.ndarray) -> np.ndarray:
    out = np.empty(2 * len(z))
    out[0::2] = z.real
    out[1::2] = z.imag
    return out


def frequencies(lo: float, hi: float, n: int) -> np.ndarray:
    return np.logspace(np.log10(lo), np.log10(hi), n)


def omega_of(nu: float) -> float:
    return 2.0 * PI * nu
