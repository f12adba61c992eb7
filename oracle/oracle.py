"""ctypes wrapper of the oracle (liboracle.so, built from pn_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

OK, ERR_ARG, ERR_OUT_OF_BOUNDS, ERR_CELL_FULL, ERR_QUERY_BOUNDS = 0, 1, 2, 3, 4

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i64, dbl, i32, ci = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int32, ctypes.c_int
        L.orc_floor_to_int_f64.restype = i64
        L.orc_floor_to_int_f64.argtypes = [dbl]
        L.orc_floor_to_int_f32.restype = i64
        L.orc_floor_to_int_f32.argtypes = [ctypes.c_float]
        L.orc_fullgrid_geometry.restype = ci
        L.orc_fullgrid_geometry.argtypes = [ci, ci, ci, vp, vp, dbl, vp, vp, vp]
        L.orc_cell_coords.restype = None
        L.orc_cell_coords.argtypes = [ci, ci, ci, vp, vp, dbl, vp]
        L.orc_candidates_dict.restype = i64
        L.orc_candidates_dict.argtypes = [ci, vp, i64, vp, dbl, vp]
        L.orc_grid_create.restype = vp
        L.orc_grid_create.argtypes = [ci, ci, ci, ci, vp, vp, dbl, i64]
        L.orc_grid_destroy.restype = None
        L.orc_grid_destroy.argtypes = [vp]
        L.orc_grid_build.restype = ci
        L.orc_grid_build.argtypes = [vp, vp, vp, i64, vp, vp]
        L.orc_grid_lengths.restype = None
        L.orc_grid_lengths.argtypes = [vp, vp]
        L.orc_grid_cell.restype = i32
        L.orc_grid_cell.argtypes = [vp, i64, vp]
        L.orc_traverse.restype = ci
        L.orc_traverse.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp, vp, vp]
        L.orc_trivial_count.restype = ci
        L.orc_trivial_count.argtypes = [ci, ci, ci, vp, vp, i64, i64, dbl, vp, vp, vp]
        L.orc_trivial_count_periodic.restype = ci
        L.orc_trivial_count_periodic.argtypes = [ci, ci, ci, vp, vp, i64, i64, dbl, vp, vp, vp, vp]
        L.orc_grid_set_periodic.restype = None
        L.orc_grid_set_periodic.argtypes = [vp, vp, vp, vp]
        L.orc_grid_cell_coords.restype = None
        L.orc_grid_cell_coords.argtypes = [vp, vp, vp]
        L.orc_periodic_sizing.restype = None
        L.orc_periodic_sizing.argtypes = [ci, ci, vp, dbl, vp, vp]
        L.orc_precomputed_lists.restype = ci
        L.orc_precomputed_lists.argtypes = [vp, vp, vp, i64, i64, vp, vp, vp]
        L.orc_threaded_count_f32.restype = ci
        L.orc_threaded_count_f32.argtypes = [vp, vp, i64, ci, vp, vp, vp]
        L.orc_threaded_wcsph_f32.restype = ci
        L.orc_threaded_wcsph_f32.argtypes = [vp, vp, i64, ci, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_float, vp, vp, vp, vp]
        L.orc_threaded_sph_rows_f32.restype = ci
        L.orc_threaded_sph_rows_f32.argtypes = [vp, vp, i64, vp, i64, ci, ci, dbl, vp, vp, vp, vp, vp, vp]
        L.orc_nbody.restype = None
        L.orc_nbody.argtypes = [ci, dbl, dbl, vp, i64, vp, vp, vp, vp, vp]
        L.orc_wcsph_interact.restype = None
        L.orc_wcsph_interact.argtypes = [ci, ci, dbl, dbl, dbl, dbl, dbl, dbl, dbl, vp, i64, vp, vp, vp, vp, vp,
                                         vp, vp]
        L.orc_sph_density.restype = None
        L.orc_sph_density.argtypes = [ci, ci, dbl, vp, i64, vp, vp, vp, vp, vp]
        L.orc_tlsph_grad.restype = None
        L.orc_tlsph_grad.argtypes = [ci, ci, dbl, vp, vp, i64, vp, vp, vp, vp, vp]
        L.orc_tlsph_force.restype = None
        L.orc_tlsph_force.argtypes = [ci, ci, dbl, vp, vp, i64, vp, vp, vp, vp, vp, vp, dbl, vp]
        L.orc_tlsph_pk1_rho2.restype = None
        L.orc_tlsph_pk1_rho2.argtypes = [ci, vp, vp, vp, i64, dbl, dbl, vp]
        L.orc_sph_pressure_accel.restype = None
        L.orc_sph_pressure_accel.argtypes = [ci, ci, dbl, vp, i64, vp, vp, vp, vp, vp, vp]
        L.orc_pair_energy.restype = None
        L.orc_pair_energy.argtypes = [ci, dbl, dbl, vp, i64, vp, vp, vp]
        L.orc_loss_dual.restype = None
        L.orc_loss_dual.argtypes = [ci, ci, ci, dbl, dbl, dbl, vp, i64, vp, vp, vp, dbl, dbl, dbl, i64, ci,
                                    vp, vp]
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def code(dt) -> int:
    return 1 if np.dtype(dt) == np.float64 else 0


def floor_to_int(x, dtype=np.float64) -> int:
    if np.dtype(dtype) == np.float32:
        return lib().orc_floor_to_int_f32(ctypes.c_float(x))
    return lib().orc_floor_to_int_f64(float(x))


def fullgrid_geometry(user_min, user_max, radius, user_dtype=np.float64, r_dtype=np.float64):
    """Returns (padded_min, padded_max, n_cells or None for a template, corner dtype)."""
    nd = len(user_min)
    umin = np.asarray(user_min, dtype=np.float64)
    umax = np.asarray(user_max, dtype=np.float64)
    omin, omax = np.zeros(nd), np.zeros(nd)
    on = np.zeros(nd, dtype=np.int64)
    tm = lib().orc_fullgrid_geometry(nd, code(user_dtype), code(r_dtype), _ptr(umin), _ptr(umax),
                                     float(radius), _ptr(omin), _ptr(omax), _ptr(on))
    corner_dtype = np.float64 if (code(user_dtype) or code(r_dtype)) else np.float32
    return omin, omax, (None if tm < 0 else on), corner_dtype


def periodic_sizing(box_size, radius, r_dtype=np.float64):
    """(n_cells, cell_size) of a GridNeighborhoodSearch with a PeriodicBox."""
    nd = len(box_size)
    bs = np.asarray(box_size, dtype=np.float64)
    n = np.zeros(nd, dtype=np.int64)
    cs = np.zeros(nd)
    lib().orc_periodic_sizing(nd, code(r_dtype), _ptr(bs), float(radius), _ptr(n), _ptr(cs))
    return n, cs


def trivial_periodic(x_nd, y_nd, radius, box_size, coord_dtype=np.float64, radius_dtype=np.float64,
                     lists=False):
    """TrivialNeighborhoodSearch with a PeriodicBox: counts (and sorted lists)."""
    x = np.ascontiguousarray(x_nd, dtype=coord_dtype)
    y = np.ascontiguousarray(y_nd, dtype=coord_dtype)
    nd = x.shape[1]
    bs = np.asarray(box_size, dtype=np.float64)
    counts = np.zeros(x.shape[0], dtype=np.int32)
    L = lib()
    L.orc_trivial_count_periodic(nd, code(coord_dtype), code(radius_dtype), _ptr(x), _ptr(y), x.shape[0],
                                 y.shape[0], float(radius), _ptr(bs), _ptr(counts), None, None)
    if not lists:
        return counts
    rowptr = np.zeros(x.shape[0] + 1, dtype=np.int64)
    rowptr[1:] = np.cumsum(counts)
    cols = np.zeros(max(int(rowptr[-1]), 1), dtype=np.int32)
    L.orc_trivial_count_periodic(nd, code(coord_dtype), code(radius_dtype), _ptr(x), _ptr(y), x.shape[0],
                                 y.shape[0], float(radius), _ptr(bs), None, _ptr(rowptr), _ptr(cols))
    return counts, [sorted(cols[rowptr[i]:rowptr[i + 1]].tolist()) for i in range(x.shape[0])]


class Grid:
    """Oracle GridNeighborhoodSearch + FullGridCellList (DVoV storage)."""

    def __init__(self, ndims, min_corner, n_cells, radius, coord_dtype=np.float64,
                 radius_dtype=np.float64, cell_dtype=np.float64, max_points_per_cell=100):
        self.nd = ndims
        self.tc = np.dtype(coord_dtype)
        self.tr = np.dtype(radius_dtype)
        self.tcell = np.dtype(cell_dtype)
        mc = np.asarray(min_corner, dtype=np.float64)
        nc = np.asarray(n_cells, dtype=np.int64)
        self.n_cells = tuple(int(v) for v in nc)
        self.ncells = int(np.prod(nc))
        self.cap = max_points_per_cell
        self.h = lib().orc_grid_create(ndims, code(coord_dtype), code(radius_dtype), code(cell_dtype),
                                       _ptr(mc), _ptr(nc), float(radius), int(max_points_per_cell))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_grid_destroy(self.h)
            self.h = None

    def set_periodic(self, cell_size, n_periodic, box_size):
        """Periodic box: per-dimension cell size, periodic cell counts and box size."""
        cs = np.asarray(cell_size, dtype=np.float64)
        n = np.asarray(n_periodic, dtype=np.int64)
        bs = np.asarray(box_size, dtype=np.float64)
        lib().orc_grid_set_periodic(self.h, _ptr(cs), _ptr(n), _ptr(bs))

    def cell_coords(self, x):
        """1-based cell coordinates of one point (periodic wrap included)."""
        xv = np.ascontiguousarray(x, dtype=self.tc)
        out = np.zeros(3, dtype=np.int64)
        lib().orc_grid_cell_coords(self.h, _ptr(xv), _ptr(out))
        return tuple(int(v) for v in out[: self.nd])

    def _pts(self, a):
        a = np.ascontiguousarray(a, dtype=self.tc)
        assert a.ndim == 2 and a.shape[1] == self.nd
        return a

    def build(self, y_nd, active=None):
        """initialize!/update!; y_nd is (N, D).  Returns (status, bad_point, bad_cell)."""
        y = self._pts(y_nd)
        act = np.arange(y.shape[0], dtype=np.int64) if active is None else np.asarray(active, dtype=np.int64)
        bp = ctypes.c_int64(-1)
        bc = np.zeros(3, dtype=np.int64)
        st = lib().orc_grid_build(self.h, _ptr(y), _ptr(act), len(act), ctypes.byref(bp), _ptr(bc))
        self._y = y
        return st, bp.value, tuple(bc[: self.nd])

    def lengths(self):
        out = np.zeros(self.ncells, dtype=np.int32)
        lib().orc_grid_lengths(self.h, _ptr(out))
        return out

    def cell(self, linear0):
        out = np.zeros(self.cap, dtype=np.int32)
        n = lib().orc_grid_cell(self.h, int(linear0), _ptr(out))
        return out[:n].tolist()

    def traverse(self, x_nd, y_nd, points=None, lists=False):
        """Counts, index sums and (optionally) neighbour lists in traversal order."""
        x = self._pts(x_nd)
        y = self._pts(y_nd)
        pts = np.arange(x.shape[0], dtype=np.int64) if points is None else np.asarray(points, dtype=np.int64)
        counts = np.zeros(len(pts), dtype=np.int32)
        sums = np.zeros(len(pts), dtype=np.int64)
        bp = ctypes.c_int64(-1)
        st = lib().orc_traverse(self.h, _ptr(x), _ptr(y), _ptr(pts), len(pts), _ptr(counts), _ptr(sums),
                                None, None, ctypes.byref(bp))
        if st != OK:
            return st, bp.value, None, None, None
        if not lists:
            return st, -1, counts, sums, None
        rowptr = np.zeros(len(pts) + 1, dtype=np.int64)
        rowptr[1:] = np.cumsum(counts)
        cols = np.zeros(max(int(rowptr[-1]), 1), dtype=np.int32)
        st = lib().orc_traverse(self.h, _ptr(x), _ptr(y), _ptr(pts), len(pts), None, None, _ptr(rowptr),
                                _ptr(cols), ctypes.byref(bp))
        return st, -1, counts, sums, (rowptr, cols[: rowptr[-1]])

    def precomputed(self, x_nd, y_nd, max_neighbors):
        x = self._pts(x_nd)
        y = self._pts(y_nd)
        st, bp, counts, _, _ = self.traverse(x, y)
        if st != OK:
            return st, bp, None
        rowptr = np.zeros(x.shape[0] + 1, dtype=np.int64)
        rowptr[1:] = np.cumsum(counts)
        cols = np.zeros(max(int(rowptr[-1]), 1), dtype=np.int32)
        bpc = ctypes.c_int64(-1)
        st = lib().orc_precomputed_lists(self.h, _ptr(x), _ptr(y), x.shape[0], int(max_neighbors),
                                         _ptr(rowptr), _ptr(cols), ctypes.byref(bpc))
        return st, bpc.value, (rowptr, cols[: rowptr[-1]])

    def threaded_wcsph_f32(self, y_nd, nthreads, mass, h, rho0, c, gamma=7.0):
        y = self._pts(y_nd)
        rho = np.zeros(y.shape[0], dtype=np.float32)
        dv = np.zeros_like(y)
        tb, tp = ctypes.c_double(), ctypes.c_double()
        st = lib().orc_threaded_wcsph_f32(self.h, _ptr(y), y.shape[0], int(nthreads), mass, h, rho0, c, gamma,
                                          _ptr(rho), _ptr(dv), ctypes.byref(tb), ctypes.byref(tp))
        return st, rho, dv, tb.value, tp.value

    def threaded_sph_rows_f32(self, y_nd, rows, nthreads, mass, h, kernel="cubic_spline", P=None):
        """Density (and, with P, pressure acceleration) in double over the grid's neighbour
        sets for the query rows `rows` of y (the grid must be built from y).  Returns
        (status, rho, rho_scale, dv, dv_scale), scales = sums of |term|."""
        y = self._pts(y_nd)
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        m = np.ascontiguousarray(np.broadcast_to(mass, (y.shape[0],)), dtype=np.float64)
        Pd = None if P is None else np.ascontiguousarray(P, dtype=np.float64)
        nr = len(rows)
        rho, rs = np.zeros(nr), np.zeros(nr)
        dv, ds = np.zeros((nr, 3)), np.zeros(nr)
        st = lib().orc_threaded_sph_rows_f32(self.h, _ptr(y), y.shape[0], _ptr(rows), nr, int(nthreads),
                                             KERNELS[kernel], float(h), _ptr(m),
                                             _ptr(Pd) if Pd is not None else None, _ptr(rho), _ptr(rs),
                                             _ptr(dv), _ptr(ds))
        return st, rho, rs, (dv if P is not None else None), (ds if P is not None else None)

    def threaded_count_f32(self, y_nd, nthreads):
        y = self._pts(y_nd)
        counts = np.zeros(y.shape[0], dtype=np.int32)
        tb, tt = ctypes.c_double(), ctypes.c_double()
        st = lib().orc_threaded_count_f32(self.h, _ptr(y), y.shape[0], int(nthreads), _ptr(counts),
                                          ctypes.byref(tb), ctypes.byref(tt))
        return st, counts, tb.value, tt.value


def host_threads() -> int:
    """Threads for the oracle's threaded paths: the host cores this process may use."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def trivial(x_nd, y_nd, radius, coord_dtype=np.float64, radius_dtype=np.float64, lists=False):
    """TrivialNeighborhoodSearch brute force: counts and (optionally) lists (ascending)."""
    x = np.ascontiguousarray(x_nd, dtype=coord_dtype)
    y = np.ascontiguousarray(y_nd, dtype=coord_dtype)
    nd = x.shape[1]
    counts = np.zeros(x.shape[0], dtype=np.int32)
    lib().orc_trivial_count(nd, code(coord_dtype), code(radius_dtype), _ptr(x), _ptr(y), x.shape[0],
                            y.shape[0], float(radius), _ptr(counts), None, None)
    if not lists:
        return counts, None
    rowptr = np.zeros(x.shape[0] + 1, dtype=np.int64)
    rowptr[1:] = np.cumsum(counts)
    cols = np.zeros(max(int(rowptr[-1]), 1), dtype=np.int32)
    lib().orc_trivial_count(nd, code(coord_dtype), code(radius_dtype), _ptr(x), _ptr(y), x.shape[0],
                            y.shape[0], float(radius), None, _ptr(rowptr), _ptr(cols))
    return counts, (rowptr, cols[: rowptr[-1]])


def candidates_dict(y_nd, xq, cell_size):
    y = np.ascontiguousarray(y_nd, dtype=np.float64)
    q = np.ascontiguousarray(xq, dtype=np.float64)
    out = np.zeros(y.shape[0], dtype=np.int64)
    m = lib().orc_candidates_dict(y.shape[1], _ptr(y), y.shape[0], _ptr(q), float(cell_size), _ptr(out))
    return out[:m].tolist()


def cell_coords(x, min_corner, cell_size, coord_dtype=np.float64, cell_dtype=np.float64):
    xa = np.ascontiguousarray(x, dtype=coord_dtype)
    mc = np.ascontiguousarray(min_corner, dtype=np.float64)
    out = np.zeros(3, dtype=np.int64)
    lib().orc_cell_coords(len(xa), code(coord_dtype), code(cell_dtype), _ptr(xa), _ptr(mc), float(cell_size),
                          _ptr(out))
    return tuple(int(v) for v in out[: len(xa)])


def sorted_lists(rowptr, cols):
    """Sort each CSR row (for comparing traversal-order lists as sets)."""
    rows = np.repeat(np.arange(len(rowptr) - 1), np.diff(rowptr))
    order = np.lexsort((cols, rows))
    return cols[order]


# ---- physics over fixed neighbour lists (x == y), double precision ---------------------------

KERNELS = {"cubic_spline": 0, "wendland_c2": 1}


def _csr(lists):
    rowptr, cols = lists
    return np.ascontiguousarray(rowptr, dtype=np.int64), np.ascontiguousarray(cols, dtype=np.int32)


def sph_density(x_nd, lists, mass, h, kernel="cubic_spline"):
    """(rho, scale) with scale_i = sum_j |m_j W_ij|."""
    x = np.ascontiguousarray(x_nd, dtype=np.float64)
    rp, cols = _csr(lists)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    rho = np.zeros(x.shape[0])
    scale = np.zeros(x.shape[0])
    lib().orc_sph_density(x.shape[1], KERNELS[kernel], float(h), _ptr(x), x.shape[0], _ptr(rp), _ptr(cols),
                          _ptr(m), _ptr(rho), _ptr(scale))
    return rho, scale


def tlsph_grad(X_nd, x_nd, lists, volume, h, kernel="wendland_c2", L=None):
    """TLSPH deformation gradient (oracle restatement): (N, D, D) with out[a] = G_a L_a."""
    X = np.ascontiguousarray(X_nd, dtype=np.float64)
    x = np.ascontiguousarray(x_nd, dtype=np.float64)
    rp, cols = _csr(lists)
    v = np.ascontiguousarray(volume, dtype=np.float64)
    n, nd = X.shape
    out = np.zeros((n, nd, nd))
    Lp = None if L is None else np.ascontiguousarray(L, dtype=np.float64)
    lib().orc_tlsph_grad(nd, KERNELS[kernel], float(h), _ptr(X), _ptr(x), n, _ptr(rp), _ptr(cols), _ptr(v),
                         None if Lp is None else _ptr(Lp), _ptr(out))
    return out


def tlsph_force(X_nd, x_nd, lists, mass, rho, pk1_rho2, h, kernel="wendland_c2", F=None, young_modulus=0.0,
                penalty_alpha=0.0):
    """TLSPH interaction dv (N, D) (oracle restatement); pk1_rho2 and F are (N, D, D) with
    [a] the matrix of point a; the penalty term needs F and penalty_alpha > 0."""
    X = np.ascontiguousarray(X_nd, dtype=np.float64)
    x = np.ascontiguousarray(x_nd, dtype=np.float64)
    rp, cols = _csr(lists)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    r = np.ascontiguousarray(rho, dtype=np.float64)
    P = np.ascontiguousarray(pk1_rho2, dtype=np.float64)
    Fp = None if (F is None or penalty_alpha == 0.0) else np.ascontiguousarray(F, dtype=np.float64)
    n, nd = X.shape
    dv = np.zeros((n, nd))
    lib().orc_tlsph_force(nd, KERNELS[kernel], float(h), _ptr(X), _ptr(x), n, _ptr(rp), _ptr(cols), _ptr(m), _ptr(r),
                          _ptr(P), _ptr(Fp), 0.5 * penalty_alpha * young_modulus, _ptr(dv))
    return dv


def tlsph_pk1_rho2(F, rho, lame_lambda, lame_mu, L=None):
    """Corrected SVK PK1 / rho^2, (N, D, D) (oracle restatement)."""
    Fa = np.ascontiguousarray(F, dtype=np.float64)
    r = np.ascontiguousarray(rho, dtype=np.float64)
    Lp = None if L is None else np.ascontiguousarray(L, dtype=np.float64)
    out = np.zeros_like(Fa)
    lib().orc_tlsph_pk1_rho2(Fa.shape[1], _ptr(Fa), _ptr(Lp), _ptr(r), Fa.shape[0], float(lame_lambda),
                             float(lame_mu), _ptr(out))
    return out


def sph_pressure_accel(x_nd, lists, mass, P, h, kernel="cubic_spline"):
    """(dv (N, D), scale) with scale_i = sum_j sum_d |term|."""
    x = np.ascontiguousarray(x_nd, dtype=np.float64)
    rp, cols = _csr(lists)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    Pa = np.ascontiguousarray(P, dtype=np.float64)
    dv = np.zeros_like(x)
    scale = np.zeros(x.shape[0])
    lib().orc_sph_pressure_accel(x.shape[1], KERNELS[kernel], float(h), _ptr(x), x.shape[0], _ptr(rp), _ptr(cols),
                                 _ptr(m), _ptr(Pa), _ptr(dv), _ptr(scale))
    return dv, scale


def pair_energy(x_nd, lists, eps, rc):
    x = np.ascontiguousarray(x_nd, dtype=np.float64)
    rp, cols = _csr(lists)
    e = np.zeros(x.shape[0])
    lib().orc_pair_energy(x.shape[1], float(eps), float(rc), _ptr(x), x.shape[0], _ptr(rp), _ptr(cols), _ptr(e))
    return e


LOSS_ENERGY, LOSS_DENSITY, LOSS_FORCE = 0, 1, 2


def loss_dual(which, x_nd, lists, k, dim, *, mass=None, kernel="cubic_spline", h=1.0, eps=1.0, rc=1.0,
              rho0=1.0, c=1.0, gamma=7.0):
    """(loss, d loss / d x[k, dim]) by forward-mode dual numbers."""
    x = np.ascontiguousarray(x_nd, dtype=np.float64)
    rp, cols = _csr(lists)
    m = np.ascontiguousarray(np.ones(x.shape[0]) if mass is None else mass, dtype=np.float64)
    v, d = ctypes.c_double(), ctypes.c_double()
    lib().orc_loss_dual(which, x.shape[1], KERNELS[kernel], float(h), float(eps), float(rc), _ptr(x), x.shape[0],
                        _ptr(rp), _ptr(cols), _ptr(m), float(rho0), float(c), float(gamma), int(k), int(dim),
                        ctypes.byref(v), ctypes.byref(d))
    return v.value, d.value


SQRT_EPS = {np.dtype(np.float32): float(np.sqrt(np.finfo(np.float32).eps)),
            np.dtype(np.float64): float(np.sqrt(np.finfo(np.float64).eps))}


def nbody(x_nd, lists, mass, G, eltype=np.float64):
    """benchmarks/n_body.jl:34-51 over CSR lists in double: (dv (N, D), scale)."""
    x = np.ascontiguousarray(x_nd, dtype=np.float64)
    rp, cols = _csr(lists)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    dv = np.zeros_like(x)
    scale = np.zeros(x.shape[0])
    lib().orc_nbody(x.shape[1], float(G), SQRT_EPS[np.dtype(eltype)], _ptr(x), x.shape[0], _ptr(rp), _ptr(cols),
                    _ptr(m), _ptr(dv), _ptr(scale))
    return dv, scale


def wcsph_interact(x_nd, lists, mass, v, p, h, c, *, kernel="wendland_c2", alpha=0.02, beta=0.0, epsilon=0.01,
                   delta=0.1, eltype=np.float64):
    """TrixiParticles' fluid-fluid interact! (benchmark_wcsph) over CSR lists in double.
    v: (N, D+1) velocity then density.  Returns (dv (N, D+1), scale (N, 2))."""
    x = np.ascontiguousarray(x_nd, dtype=np.float64)
    rp, cols = _csr(lists)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    vv = np.ascontiguousarray(v, dtype=np.float64)
    pp = np.ascontiguousarray(p, dtype=np.float64)
    n, nd = x.shape
    dv = np.zeros((n, nd + 1))
    scale = np.zeros((n, 2))
    lib().orc_wcsph_interact(nd, KERNELS[kernel], float(h), float(c), float(alpha), float(beta), float(epsilon),
                             float(delta), SQRT_EPS[np.dtype(eltype)], _ptr(x), n, _ptr(rp), _ptr(cols), _ptr(m),
                             _ptr(vv), _ptr(pp), _ptr(dv), _ptr(scale))
    return dv, scale


def tait_eos(rho, rho0, c, gamma=7.0):
    B = rho0 * c * c / gamma
    return B * ((np.asarray(rho, dtype=np.float64) / rho0) ** gamma - 1.0)

