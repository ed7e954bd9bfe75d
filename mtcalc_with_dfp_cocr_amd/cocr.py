"""Host-side mirror of the reference's COCR_FP interface over the HIP C-ABI.

`pcocr_2x2_folded` keeps the argument list, argument meaning and return codes of
PCOCR::PCOCR_2x2_Folded (Code/COCR_FP/pcocr.h:27-31, pcocr.cpp:40-126) minus the per-thread
scatter buffer y_omp: x is updated in place and the reference's request code is returned
(0 REQ_OK, -1 REQ_X0_NULL, -2 REQ_DIV_BY_ZERO, -3 REQ_MAXITER, rci.h:23,40-42).

`Solver` is the handle-based form (one matrix + gradient resident on the GPU, many right-hand
sides, batched frequencies sharing K and S, mixed precision).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import MtcocrStats, check, lib, ptr
from .model import BlockSystem, MeshArrays, Portrait

REQ_OK, REQ_X0_NULL, REQ_DIV_BY_ZERO, REQ_MAXITER = 0, -1, -2, -3
PRECISION = {"fp64": 0, "mixed": 1}


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


MAX_BATCH = 64  # systems of one batch (the C side rejects more)
ORDER_GRID = -1  # set_mesh_order / mesh_order: the grid order of a box mesh


def _vec(a, n: int, what: str) -> np.ndarray:
    """A float64 vector of exactly n entries (the C side reads and writes n doubles)."""
    v = _f64(a)
    if v.ndim != 1 or v.size != n:
        raise ValueError(f"{what} must be a vector of {n} doubles (2 * nb), got shape {v.shape}")
    return v


def _batch(a, nb: int, what: str, nsys: int | None = None) -> np.ndarray:
    """nsys x 2*nb float64 right-hand sides or solutions (1 <= nsys <= 64)."""
    v = _f64(np.atleast_2d(a))
    if v.ndim != 2 or v.shape[1] != 2 * nb or not 1 <= v.shape[0] <= MAX_BATCH:
        raise ValueError(f"{what} must have shape (nsys, {2 * nb}) with 1 <= nsys <= "
                         f"{MAX_BATCH}, got {v.shape}")
    if nsys is not None and v.shape[0] != nsys:
        raise ValueError(f"{what} must hold {nsys} systems, got {v.shape[0]}")
    return v


class Solver:
    """One GPU-resident matrix (single system or K/S pair) plus the folded preconditioner's
    gradient, solved for one or many right-hand sides."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        check(self._L.mtcocr_create(int(device), C.byref(h)), "mtcocr_create")
        self._h = h
        self.nb = 0
        self.nsys = 0

    def close(self):
        if self._h:
            self._L.mtcocr_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up -------------------------------------------------------------------------
    def set_mesh_order(self, ma: MeshArrays, block: int = 16) -> None:
        """Hold the matrix, gradient and vectors in the mesh's pencil order on the device
        (inputs and outputs keep the caller's order), or with block = ORDER_GRID in the grid
        order of a box mesh, which also enables the stencil SpMV. Call before set_matrix /
        set_gradient."""
        check(self._L.mtcocr_set_mesh_order(self._h, int(ma.kpar), int(ma.kuzlov),
                                            _i32(ma.nver), int(len(ma.edges) // 2),
                                            _i32(ma.edges), int(block)), "mtcocr_set_mesh_order")

    def set_matrix(self, s: BlockSystem):
        self._keep = (_i32(s.ig), _i32(s.jg), _i32(s.idi), _i32(s.ijg), _f64(s.di), _f64(s.gg))
        check(self._L.mtcocr_set_matrix_block2x2(self._h, int(s.nb), *self._keep),
              "mtcocr_set_matrix_block2x2")
        self.nb = int(s.nb)

    def set_matrix_ks(self, p: Portrait, kd, kg, sd, sg):
        args = (_i32(p.ig), _i32(p.jg), _f64(kd), _f64(kg), _f64(sd), _f64(sg))
        check(self._L.mtcocr_set_matrix_ks(self._h, int(p.nb), *args), "mtcocr_set_matrix_ks")
        self.nb = int(p.nb)

    def set_gradient(self, ma: MeshArrays):
        check(self._L.mtcocr_set_gradient_mesh(
            self._h, int(ma.kpar), int(len(ma.edges) // 2), int(ma.kuzlov), _i32(ma.nvkat),
            _i32(ma.nver), _f64(ma.sig_flags), _i32(ma.edges), None, None, None,
            _i32(ma.is_node_bound)), "mtcocr_set_gradient_mesh")

    # ---- solves -------------------------------------------------------------------------
    def solve(self, pr, x0=None, eps: float = 1e-8, maxiter: int = 10000,
              precision: str = "fp64"):
        """One system set by set_matrix. Returns (x, stats dict)."""
        n = 2 * self.nb
        pr = _vec(pr, n, "pr")
        x = (np.zeros(n) if x0 is None else _vec(x0, n, "x0")).copy()
        st = MtcocrStats()
        rc = self._L.mtcocr_solve(self._h, pr, x, float(eps), int(maxiter),
                                  PRECISION[precision], C.byref(st))
        check(rc, "mtcocr_solve")
        return x, st.as_dict()

    def solve_batch(self, omega, prs, x0=None, eps: float = 1e-8, maxiter: int = 10000,
                    precision: str = "fp64"):
        """nsys systems A(omega_s) x_s = b_s on the K/S matrix. prs: (nsys, 2*nb)."""
        prs = _batch(prs, self.nb, "prs")
        nsys = prs.shape[0]
        x = (np.zeros_like(prs) if x0 is None else _batch(x0, self.nb, "x0", nsys)).copy()
        om = None if omega is None else _vec(omega, nsys, "omega")
        sts = (MtcocrStats * nsys)()
        rc = self._L.mtcocr_solve_batch(self._h, nsys, ptr(om), prs.ravel(), x.ravel(),
                                        float(eps), int(maxiter), PRECISION[precision], sts)
        check(rc, "mtcocr_solve_batch")
        return x.reshape(prs.shape), [s.as_dict() for s in sts]

    def set_rhs(self, prs, omega=None, x0=None):
        prs = _batch(prs, self.nb, "prs")
        nsys = prs.shape[0]
        om = None if omega is None else _vec(omega, nsys, "omega")
        x0 = None if x0 is None else _batch(x0, self.nb, "x0", nsys)
        check(self._L.mtcocr_set_rhs(self._h, nsys, ptr(om), prs.ravel(), ptr(x0)),
              "mtcocr_set_rhs")
        self.nsys = nsys

    def solve_resident(self, eps: float = 1e-8, maxiter: int = 10000, precision: str = "fp64"):
        sts = (MtcocrStats * self.nsys)()
        rc = self._L.mtcocr_solve_resident(self._h, float(eps), int(maxiter),
                                           PRECISION[precision], sts)
        check(rc, "mtcocr_solve_resident")
        return rc, [s.as_dict() for s in sts]

    def get_solution(self):
        x = np.empty(self.nsys * 2 * self.nb)
        check(self._L.mtcocr_get_solution(self._h, x), "mtcocr_get_solution")
        return x.reshape(self.nsys, 2 * self.nb)

    # ---- measurement ----------------------------------------------------------------------
    def bench_spmv(self, precision: str = "fp64", reps: int = 20):
        ms, by = C.c_double(), C.c_double()
        check(self._L.mtcocr_bench_spmv(self._h, PRECISION[precision], self.nsys, int(reps),
                                        C.byref(ms), C.byref(by)), "mtcocr_bench_spmv")
        return ms.value, by.value

    def bench_update(self, precision: str = "fp64", reps: int = 20):
        """Average ms and algorithmic bytes of the fused per-iteration update pass."""
        ms, by = C.c_double(), C.c_double()
        check(self._L.mtcocr_bench_update(self._h, PRECISION[precision], int(reps),
                                          C.byref(ms), C.byref(by)), "mtcocr_bench_update")
        return ms.value, by.value

    def bench_flush(self, reps: int = 20):
        """Average ms and algorithmic bytes of the FP64 solver's x/p pass (every 8th iteration)."""
        ms, by = C.c_double(), C.c_double()
        check(self._L.mtcocr_bench_flush(self._h, int(reps), C.byref(ms), C.byref(by)),
              "mtcocr_bench_flush")
        return ms.value, by.value

    def bench_copy(self, nbytes: int, reps: int = 20) -> float:
        """Average ms of the plain copy kernel over nbytes (HBM counter calibration)."""
        ms = C.c_double()
        check(self._L.mtcocr_bench_copy(self._h, int(nbytes), int(reps), C.byref(ms)),
              "mtcocr_bench_copy")
        return ms.value

    LAYOUTS = {"ell": 0, "rowpack": 1, "rowblock": 2, "stencil": 3, "stencil_lds": 4}
    LAYOUT_NAMES = ("ell", "rowpack", "rowblock", "stencil", "stencil_lds")

    def set_launch(self, cap_spmv: int = 0, cap_vec: int = 0, cap_node: int = 0,
                   spmv_layout: str | None = None) -> None:
        """Workgroup caps of the SpMV / edge-streaming / node kernels (0 keeps a value) and the
        SpMV layout (a key of LAYOUTS; None keeps it)."""
        lay = -1 if spmv_layout is None else self.LAYOUTS[spmv_layout]
        check(self._L.mtcocr_set_launch(self._h, int(cap_spmv), int(cap_vec), int(cap_node),
                                        lay), "mtcocr_set_launch")

    def get_launch(self, precision: str = "fp64") -> dict:
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(self._L.mtcocr_get_launch(self._h, PRECISION[precision], C.byref(a), C.byref(b),
                                        C.byref(c), C.byref(d)), "mtcocr_get_launch")
        return {"cap_spmv": a.value, "cap_vec": b.value, "cap_node": c.value,
                "spmv_layout": self.LAYOUT_NAMES[d.value]}

    def set_stencil_sweep(self, kc: int) -> None:
        """Rows a task of the stencil SpMVs sweeps along z (every width and precision)."""
        check(self._L.mtcocr_set_stencil_sweep(self._h, int(kc)), "mtcocr_set_stencil_sweep")

    def set_xp_period(self, period: int) -> None:
        """x/p period of the FP64 solver: 8 or 16 iterations, 0 to choose by free memory."""
        check(self._L.mtcocr_set_xp_period(self._h, int(period)), "mtcocr_set_xp_period")

    def get_xp_period(self) -> int:
        v = C.c_int()
        check(self._L.mtcocr_get_xp_period(self._h, C.byref(v)), "mtcocr_get_xp_period")
        return v.value

    def set_stencil_band(self, band: int) -> None:
        """Tile columns per band of the LDS-staged stencil SpMV's task order (0: one band)."""
        check(self._L.mtcocr_set_stencil_band(self._h, int(band)), "mtcocr_set_stencil_band")

    def get_stencil_band(self, precision: str = "fp64") -> int:
        b = C.c_int()
        check(self._L.mtcocr_get_stencil_band(self._h, PRECISION[precision], C.byref(b)),
              "mtcocr_get_stencil_band")
        return b.value

    def get_stencil_sweep(self, precision: str = "fp64") -> int:
        kc = C.c_int()
        check(self._L.mtcocr_get_stencil_sweep(self._h, PRECISION[precision], C.byref(kc)),
              "mtcocr_get_stencil_sweep")
        return kc.value

    def autotune(self, precision: str = "fp64", reps: int = 10) -> dict:
        """Time the tunable kernels of the current batch at each candidate shape and keep the
        fastest (needs a batch on the handle: set_rhs or a previous solve_batch)."""
        a, b, c, d, ms = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(self._L.mtcocr_autotune(self._h, PRECISION[precision], int(reps), C.byref(a),
                                      C.byref(b), C.byref(c), C.byref(d), C.byref(ms)),
              "mtcocr_autotune")
        out = {"cap_spmv": a.value, "cap_vec": b.value, "cap_node": c.value,
               "spmv_layout": self.LAYOUT_NAMES[d.value], "spmv_ms": ms.value}
        if out["spmv_layout"] in ("stencil", "stencil_lds"):
            out["stencil_sweep"] = self.get_stencil_sweep(precision)
        if out["spmv_layout"] == "stencil_lds":
            out["stencil_band"] = self.get_stencil_band(precision)
        return out

    def apply(self, op: str, x, precision: str = "fp64"):
        """Kernel-level hook: op "spmv" (A x) or "precond" (folded preconditioner)."""
        y = np.empty(2 * self.nb)
        check(self._L.mtcocr_apply(self._h, {"spmv": 0, "precond": 1}[op], PRECISION[precision],
                                   _vec(x, 2 * self.nb, "x"), y), "mtcocr_apply")
        return y

    def sigma_gradient(self, shape_nodes, x, y, z, coef, u, v) -> np.ndarray:
        """grad[c] = sum_s Re(coef_s * u_s^T M_c v_s) per cell of the regular grid, on the device
        (M_c: the cell's edge mass matrix; u, v: (nsys, n_edges) complex; coef: (nsys,))."""
        nxn, nyn, nzn = (int(n) for n in shape_nodes)
        u = np.ascontiguousarray(u, dtype=np.complex128)
        v = np.ascontiguousarray(v, dtype=np.complex128)
        nsys = u.shape[0]
        if u.ndim != 2 or v.shape != u.shape or len(coef) != nsys:
            raise ValueError("u, v and coef must have one row / entry per system")
        ci = np.ascontiguousarray(coef, dtype=np.complex128).view(np.float64)
        out = np.empty((nxn - 1) * (nyn - 1) * (nzn - 1))
        check(self._L.mtcocr_sigma_gradient(self._h, nxn, nyn, nzn, _f64(x), _f64(y), _f64(z),
                                            nsys, ci, u.view(np.float64).reshape(-1),
                                            v.view(np.float64).reshape(-1), out),
              "mtcocr_sigma_gradient")
        return out

    def info(self) -> dict:
        a, b, c, d = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        check(self._L.mtcocr_matrix_info(self._h, C.byref(a), C.byref(b), C.byref(c),
                                         C.byref(d)), "mtcocr_matrix_info")
        return {"nnz_complex": a.value, "nnz_real": b.value, "m": c.value,
                "device_bytes": d.value}


def build_gmatrix(ma: MeshArrays, ig_t=None, jg_t=None, gg_t=None, n_nodes_c: int | None = None):
    """FoldedPreconditioner::BuildGMatrix (FoldedPreconditioner.cpp:68-259) on the host: the
    CSR arrays (iptr_gt, jptr_gt, gg_gt) of G^T and (iptr_g, jptr_g, gg_g) of G, 0-based, in
    the reference's node numbering. n_nodes_c: conforming nodes (default: all of the mesh's);
    nodes above it are hanging and spread through the T-matrix ig_t/jg_t/gg_t (indexed by the
    hanging node's global number)."""
    L = lib()
    n = len(ma.edges) // 2
    nc = int(ma.kuzlov if n_nodes_c is None else n_nodes_c)
    t = (ptr(None if ig_t is None else _i32(ig_t)), ptr(None if jg_t is None else _i32(jg_t)),
         ptr(None if gg_t is None else _f64(gg_t)))
    args = (int(ma.kpar), n, nc, _i32(ma.nvkat), _i32(ma.nver), _f64(ma.sig_flags),
            _i32(ma.edges), *t, _i32(ma.is_node_bound))
    m, nnz = C.c_int(), C.c_int64()
    check(L.mtcocr_build_gmatrix(*args, C.byref(m), C.byref(nnz), None, None, None, None, None,
                                 None), "mtcocr_build_gmatrix")
    it, jt, gt = np.empty(n + 1, np.int32), np.empty(nnz.value, np.int32), np.empty(nnz.value)
    ip, jp, gp = np.empty(m.value + 1, np.int32), np.empty(nnz.value, np.int32), np.empty(nnz.value)
    keep = [a.ctypes.data_as(C.c_void_p) for a in (it, jt, gt, ip, jp, gp)]
    check(L.mtcocr_build_gmatrix(*args, C.byref(m), C.byref(nnz), *keep),
          "mtcocr_build_gmatrix")
    return (it, jt, gt), (ip, jp, gp)


def mesh_order(ma: MeshArrays, block: int = 16) -> np.ndarray:
    """The device's locality order of a structured hexahedral mesh: order[i] = edge index of
    device row i (pencils of block x block nodes swept along z)."""
    n = len(ma.edges) // 2
    order = np.empty(n, np.int32)
    check(lib().mtcocr_mesh_order(int(ma.kpar), int(ma.kuzlov), _i32(ma.nver), n,
                                  _i32(ma.edges), int(block), order), "mtcocr_mesh_order")
    return order


def pcocr_2x2_folded(n, ig, jg, idi, ijg, di, gg, pr, x, eps, maxiter, kpar, n_edges_c,
                     n_nodes_c, xyz, nvkat, nver, sigma3d, edges, ig_t=None, jg_t=None,
                     gg_t=None, is_node_bound=None, stats: dict | None = None) -> int:
    """PCOCR::PCOCR_2x2_Folded on the GPU. x (float64, length n) is overwritten with the
    solution; returns the reference's request code."""
    if x.dtype != np.float64 or not x.flags.c_contiguous:
        raise TypeError("x must be a C-contiguous float64 array (it is updated in place)")
    if int(n) <= 0 or x.size != int(n):
        raise ValueError(f"x must hold n = {int(n)} doubles, got {x.size}")
    pr = _vec(pr, int(n), "pr")
    st = MtcocrStats()
    xyz_a = None if xyz is None else _f64(xyz)
    rc = lib().mtcocr_pcocr_2x2_folded(
        int(n), _i32(ig), _i32(jg), _i32(idi), _i32(ijg), _f64(di), _f64(gg), _f64(pr), x,
        float(eps), int(maxiter), int(kpar), int(n_edges_c), int(n_nodes_c), ptr(xyz_a),
        _i32(nvkat), _i32(nver), _f64(sigma3d), _i32(edges),
        ptr(None if ig_t is None else _i32(ig_t)), ptr(None if jg_t is None else _i32(jg_t)),
        ptr(None if gg_t is None else _f64(gg_t)), _i32(is_node_bound), C.byref(st))
    check(rc, "mtcocr_pcocr_2x2_folded")
    if stats is not None:
        stats.update(st.as_dict())
    return rc
