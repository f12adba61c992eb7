"""GridNeighborhoodSearch with FullGridCellList (src/nhs_grid.jl) on the MI355X engine.

initialize!/update! bin the points of y into cells with the device counting sort
(`pn_grid_update`); foreach_point_neighbor runs the wave-per-cell traversal kernels.
"""

from __future__ import annotations

import ctypes
import itertools

import numpy as np

from . import _device
from ._lib import (PN_ERR_OUT_OF_BOUNDS, PN_QUERIES_BINNED, PN_QUERIES_GRID, ArgumentError,
                   BoundsError, PointNeighborsError, check, load)
from .cell_lists import FullGridCellList, SpatialHashingCellList
from .util import FLOAT64_EPS, dtype_code, floor_to_int, promote, scalar_dtype, wrap_int64


class _Strategy:
    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))

    def __repr__(self):
        return f"{type(self).__name__}()"


class ParallelUpdate(_Strategy):
    """src/nhs_grid.jl:144.  On the MI355X engine: the counting-sort rebuild."""


class ParallelIncrementalUpdate(_Strategy):
    """src/nhs_grid.jl:156.  Accepted for drop-in; runs the same counting-sort rebuild
    (identical neighbour sets)."""


class SemiParallelUpdate(_Strategy):
    """src/nhs_grid.jl:168.  Accepted for drop-in; runs the counting-sort rebuild."""


class SerialIncrementalUpdate(_Strategy):
    """src/nhs_grid.jl:180.  Accepted for drop-in; runs the counting-sort rebuild."""


class SerialUpdate(_Strategy):
    """src/nhs_grid.jl:192.  Accepted for drop-in; runs the counting-sort rebuild, whose
    order inside a cell (original index ascending) equals the serial insertion order."""


# src/cell_lists/full_grid.jl:39-42: ParallelUpdate first, so it is the default.
SUPPORTED_UPDATE_STRATEGIES = (ParallelUpdate, ParallelIncrementalUpdate, SemiParallelUpdate,
                               SerialIncrementalUpdate, SerialUpdate)


class PeriodicBox:
    """PeriodicBox(; min_corner, max_corner) (src/neighborhood_search.jl:129-141): corners
    and size = max_corner - min_corner in the corners' element type (Float32 or Float64)."""

    def __init__(self, *, min_corner, max_corner):
        dt = promote(np.asarray(min_corner).dtype, np.asarray(max_corner).dtype)
        dt = np.dtype(np.float32) if dt == np.float32 else np.dtype(np.float64)
        self.min_corner = np.asarray(min_corner, dtype=dt)
        self.max_corner = np.asarray(max_corner, dtype=dt)
        self.size = (self.max_corner - self.min_corner).astype(dt)

    def eltype(self):
        return self.size.dtype

    def __repr__(self):
        return f"PeriodicBox(min_corner={self.min_corner.tolist()}, max_corner={self.max_corner.tolist()})"


def periodic_grid_size(periodic_box, search_radius, radius_dtype):
    """n_cells and cell_size of a GridNeighborhoodSearch with a PeriodicBox
    (src/nhs_grid.jl:109-124).  n_cells = floor.(Int, (size .+ 10eps()) / search_radius)
    evaluates in Float64 (10eps() is Float64); cell_size = size ./ n_cells in the radius
    type."""
    if np.dtype(radius_dtype) != periodic_box.eltype():
        raise ArgumentError("the `search_radius` and the `PeriodicBox` must have the same element type")
    size64 = periodic_box.size.astype(np.float64)
    n = tuple(int(np.floor((v + 10.0 * FLOAT64_EPS) / float(search_radius))) for v in size64)
    T = np.dtype(radius_dtype).type
    cell_size = tuple(T(T(periodic_box.size[d]) / T(n[d])) for d in range(len(n)))
    if any(v < 3 for v in n):
        raise ArgumentError("the `GridNeighborhoodSearch` needs at least 3 cells in each dimension when "
                            "used with periodicity. Please use no NHS for very small problems.")
    return n, cell_size


class _GridMeta(type):
    def __getitem__(cls, ndims):
        return lambda **kw: cls(ndims, **kw)


class GridNeighborhoodSearch(metaclass=_GridMeta):
    """GridNeighborhoodSearch{NDIMS}(; search_radius = 0.0, n_points = 0,
    periodic_box = nothing, cell_list, update_strategy = nothing)
    (src/nhs_grid.jl:77-129).  Write `GridNeighborhoodSearch[3](...)` or
    `GridNeighborhoodSearch(3, ...)`.

    The MI355X engine requires a FullGridCellList (the reference's GPU-compatible cell
    list); the DictionaryCellList default of the reference is a CPU hash table.
    """

    def __init__(self, ndims, *, search_radius=0.0, n_points=0, periodic_box=None,
                 cell_list=None, update_strategy=None):
        if cell_list is None:
            raise ArgumentError(
                "the MI355X engine needs `cell_list = FullGridCellList(...)`; "
                "DictionaryCellList is a CPU hash table")
        if not isinstance(cell_list, (FullGridCellList, SpatialHashingCellList)):
            raise ArgumentError(f"cell list {type(cell_list).__name__} is not supported")
        if cell_list.ndims != ndims:
            raise ArgumentError(f"a {ndims}D cell list is required for a GridNeighborhoodSearch{{{ndims}}}")
        if update_strategy is None:
            update_strategy = SUPPORTED_UPDATE_STRATEGIES[0]()
        elif type(update_strategy) not in SUPPORTED_UPDATE_STRATEGIES:
            names = ", ".join(t.__name__ for t in SUPPORTED_UPDATE_STRATEGIES)
            raise ArgumentError(f"{update_strategy} is not a valid update strategy for this cell list. "
                                f"Available options are ({names})")
        r_dt = scalar_dtype(search_radius)
        if r_dt.kind in "iub":
            raise ArgumentError("`search_radius` cannot be an integer type, since computed "
                                "distances will be converted to this type")
        self.ndims = ndims
        self.search_radius = search_radius
        self.radius_dtype = r_dt
        self.n_points = n_points
        self.periodic_box = periodic_box
        self.cell_list = cell_list
        self.update_strategy = update_strategy
        if float(search_radius) < FLOAT64_EPS or periodic_box is None:
            # src/nhs_grid.jl:102-105: no periodicity
            self.cell_size = tuple(search_radius for _ in range(ndims))
            self.n_cells = tuple(-1 for _ in range(ndims))
        else:
            if len(periodic_box.size) != ndims:
                raise ArgumentError(f"a {ndims}D PeriodicBox is required")
            self.n_cells, self.cell_size = periodic_grid_size(periodic_box, search_radius, r_dt)
        self._handle = None
        self._stream_handle = None  # HIP stream last given to the handle
        self._coord_dtype = None
        self._built = False
        self._empty = False     # initialized with search_radius < eps()
        # Coordinate bookkeeping.  A key is (data_ptr, shape, stride, torch version counter).
        self._y_key = None      # key of y as binned by the last update!; None once it was
                                # declared changed (coordinates_changed)
        self._planes_key = None  # key of the y the sorted coordinate planes now hold
        self._y_ref = None      # the binned y: held so its memory cannot be handed to another tensor
        self._stale = False     # coordinates_changed() was called since the planes were filled
        self._y_full = True     # eachindex_y covered all columns
        self._n_y = 0
        self._tuning = {}

    # -- reference interface -------------------------------------------------------
    def requires_update(self):
        """src/nhs_grid.jl:133"""
        return (False, True)

    def eltype(self):
        return self.radius_dtype

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None:
            try:
                load().pn_grid_destroy(h)
            except Exception:
                pass

    # -- device handle -------------------------------------------------------------
    def _ensure_handle(self, coord_dtype):
        if self._handle is not None and self._coord_dtype == coord_dtype:
            return
        lib = load()
        if self._handle is not None:
            lib.pn_grid_destroy(self._handle)
            self._handle = None
        self._stream_handle = None
        cl = self.cell_list
        nd = self.ndims
        if self._hashed():
            h = ctypes.c_void_p()
            check(lib.pn_grid_create_spatial_hash(nd, dtype_code(coord_dtype), dtype_code(self.radius_dtype),
                                                  float(self.search_radius), cl.list_size,
                                                  int(cl.max_points_per_cell), ctypes.byref(h)),
                  "GridNeighborhoodSearch")
            self._handle = h
            self._coord_dtype = coord_dtype
            self._cell_dtype = promote(coord_dtype, self.radius_dtype)
            if self._periodic():
                cs = (ctypes.c_double * nd)(*[float(v) for v in self.cell_size])
                npc = (ctypes.c_int64 * nd)(*self.n_cells)
                bs = (ctypes.c_double * nd)(*[float(v) for v in self.periodic_box.size])
                check(lib.pn_grid_set_periodic(h, cs, npc, bs), "GridNeighborhoodSearch")
            for k, v in self._tuning.items():
                check(lib.pn_set_tuning(h, k.encode(), int(v)), "set_tuning")
            return
        cell_dt = promote(coord_dtype, cl.corner_dtype)
        mins = (ctypes.c_double * nd)(*[float(v) for v in cl.min_corner])
        ncs = (ctypes.c_int64 * nd)(*cl.n_cells)
        h = ctypes.c_void_p()
        check(lib.pn_grid_create(nd, dtype_code(coord_dtype), dtype_code(self.radius_dtype),
                                 dtype_code(cell_dt), mins, ncs, float(self.search_radius),
                                 int(cl.max_points_per_cell), ctypes.byref(h)),
              "GridNeighborhoodSearch")
        self._handle = h
        self._coord_dtype = coord_dtype
        self._cell_dtype = cell_dt
        if self._periodic():
            pb = self.periodic_box
            cs = (ctypes.c_double * nd)(*[float(v) for v in self.cell_size])
            npc = (ctypes.c_int64 * nd)(*self.n_cells)
            bs = (ctypes.c_double * nd)(*[float(v) for v in pb.size])
            check(lib.pn_grid_set_periodic(h, cs, npc, bs), "GridNeighborhoodSearch")
        for k, v in self._tuning.items():
            check(lib.pn_set_tuning(h, k.encode(), int(v)), "set_tuning")

    def _periodic(self):
        return self.periodic_box is not None and self.n_cells[0] > 0

    def _hashed(self):
        return isinstance(self.cell_list, SpatialHashingCellList)

    def set_tuning(self, key, value):
        """Autotuner knobs: "algorithm" (0: candidate-major cells, 1: per-query candidate
        slices), "persist_blocks" (traversal workgroups, multiple of 8), "chunk" (-1 per
        functor, 0 small, 1 large candidate chunk), "canonical_order" (1: in-cell order by
        original index; 0: atomic order, nondeterministic inside a cell)."""
        self._tuning[key] = int(value)
        if self._handle is not None:
            check(load().pn_set_tuning(self._handle, key.encode(), int(value)), "set_tuning")

    def _set_stream(self, t):
        h = _device.current_stream_handle(t.device)
        if h != self._stream_handle:
            load().pn_grid_set_stream(self._handle, ctypes.c_void_p(h))
            self._stream_handle = h

    def host_cell_coords(self, coords):
        """cell_coords (src/nhs_grid.jl:622-628 with src/cell_lists/full_grid.jl:84-94) of
        one point given on the host, evaluated in the same types as the device."""
        cl = self.cell_list
        if self._hashed():
            # nonperiodic_cell_coords for a SpatialHashingCellList, src/nhs_grid.jl:636-638
            ct = self._cell_dtype.type if self._handle is not None else promote(
                np.asarray(coords).dtype, self.radius_dtype).type
            out = []
            for d in range(self.ndims):
                with np.errstate(all="ignore"):
                    c = floor_to_int(float(ct(coords[d]) / ct(self.cell_size[d])))
                if self._periodic():
                    c = (wrap_int64(c - 2) % self.n_cells[d]) + 2
                out.append(c)
            return tuple(out)
        ct = self._cell_dtype.type if self._handle is not None else promote(
            np.asarray(coords).dtype, cl.corner_dtype).type
        out = []
        for d in range(self.ndims):
            cs = ct(self.cell_size[d])
            with np.errstate(all="ignore"):
                q = (ct(coords[d]) - ct(cl.min_corner[d])) / cs
            c = wrap_int64(floor_to_int(float(q)) + 1)
            if self._periodic():
                # periodic_cell_index, src/nhs_grid.jl:609-620
                c = (wrap_int64(c - 2) % self.n_cells[d]) + 2
            out.append(c)
        return tuple(out)

    # -- build -----------------------------------------------------------------------
    def _build(self, y, eachindex_y, check_errors=True):
        nd = self.ndims
        y = _device.as_coordinates(y, nd, "neighbor coordinates")
        self._built = False
        if float(self.search_radius) < FLOAT64_EPS:
            # src/nhs_grid.jl:263-267: cannot initialize with zero search radius
            self._empty = True
            return
        if self.cell_list.is_template:
            raise PointNeighborsError("`search_radius` is not defined for this cell list")
        self._empty = False
        dt = _device.np_dtype(y)
        self._ensure_handle(dt)
        n = y.shape[1]
        first, count, ptr, keep = _device.selection(eachindex_y, n, y.device)
        self._set_stream(y)
        lib = load()
        fn = lib.pn_grid_update if check_errors else lib.pn_grid_update_async
        st = fn(self._handle, ctypes.c_void_p(y.data_ptr()), n, first, count,
                ctypes.c_void_p(ptr) if ptr else None)
        if st in (PN_ERR_OUT_OF_BOUNDS,):
            p = ctypes.c_int64()
            cell = (ctypes.c_int64 * nd)()
            lib.pn_grid_error(self._handle, ctypes.byref(p), cell)
            raise PointNeighborsError(self.cell_list.bounds_error_message(list(cell)))
        check(st, "initialize!/update!")
        del keep
        self._built = True
        self._y_key = (y.data_ptr(), tuple(y.shape), y.stride(), y._version)
        self._planes_key = self._y_key
        self._y_ref = y
        self._stale = False
        self._y_full = _device.is_full(eachindex_y, n)
        self._n_y = n

    def initialize(self, x, y, *, eachindex_y=None):
        """initialize!(nhs, x, y; eachindex_y) (src/nhs_grid.jl:220-225)."""
        self._build(y, eachindex_y)
        return self

    def update(self, x, y, *, points_moving=(True, True), eachindex_y=None, check_errors=True):
        """update!(nhs, x, y; points_moving, eachindex_y) (src/nhs_grid.jl:283-292): only
        y's motion matters; the update is a full re-binning (:470-477).  check_errors=False
        only enqueues the rebuild (no synchronization); check() raises its errors later."""
        if not points_moving[1]:
            return self
        self._build(y, eachindex_y, check_errors)
        return self

    def coordinates_changed(self):
        """Declares that the binned y was written without torch seeing it (through a raw
        pointer, another library or a HIP kernel of the caller's): the next traversal
        re-reads the current coordinates, as the reference always does
        (src/nhs_grid.jl:543-545).  Cell membership stays that of the last update!."""
        self._stale = True
        return self

    def check(self):
        """Raises the errors of the last update(check_errors=False)."""
        check(load().pn_grid_check(self._handle), "update!")

    def copy(self, search_radius, n_points):
        return copy_grid_search(self, search_radius, n_points)

    def initialize_grid(self, y, *, parallelization_backend=None, eachindex_y=None):
        self._build(y, eachindex_y)
        return self

    def update_grid(self, y, *, parallelization_backend=None, eachindex_y=None):
        self._build(y, eachindex_y)
        return self

    # -- traversal support -------------------------------------------------------------
    def _prepare(self, x, y, points):
        """Returns the query mode for a traversal of x against the binned y."""
        nd = self.ndims
        x = _device.as_coordinates(x, nd, "system coordinates")
        y = _device.as_coordinates(y, nd, "neighbor coordinates")
        if self._empty or not self._built:
            if self._empty:
                raise BoundsError("the neighborhood search was initialized with a zero search radius")
            raise PointNeighborsError("the neighborhood search is not initialized")
        if _device.np_dtype(x) != self._coord_dtype or _device.np_dtype(y) != self._coord_dtype:
            raise ArgumentError("system and neighbor coordinates must have the element type "
                                "the neighborhood search was updated with")
        lib = load()
        self._set_stream(x)
        # The sorted coordinate planes hold the y they were last filled from.  They are
        # current while that is this y (the binned y is held in _y_ref, so its address cannot
        # be reused), torch's version counter has not moved (every in-place torch op and every
        # C-ABI writer of this package bumps it) and coordinates_changed() was not called.
        y_key = (y.data_ptr(), tuple(y.shape), y.stride(), y._version)
        if y_key != self._planes_key or self._stale:
            # cell membership from the last update!, coordinates read now
            # (src/nhs_grid.jl:543-545)
            if y.shape[1] != self._n_y:
                raise BoundsError("neighbor coordinates have a different number of points than "
                                  "the last update!")
            check(lib.pn_grid_refresh(self._handle, ctypes.c_void_p(y.data_ptr()), y.shape[1]),
                  "foreach_point_neighbor")
            if self._stale:
                self._y_key = None  # the binned tensor no longer holds what was binned
            self._planes_key = y_key
            self._stale = False
        nx = x.shape[1]
        # the grid's own cells serve as query cells only while y is exactly what was binned;
        # otherwise the queries are binned from the current x (their cells follow x now)
        x_is_y = _device.same_storage(x, y) and y_key == self._y_key
        if x_is_y and self._y_full and _device.is_full(points, nx):
            return PN_QUERIES_GRID, x
        first, count, ptr, keep = _device.selection(points, nx, x.device)
        st = lib.pn_grid_bin_queries(self._handle, ctypes.c_void_p(x.data_ptr()), nx, first, count,
                                     ctypes.c_void_p(ptr) if ptr else None)
        check(st, "foreach_point_neighbor")
        del keep
        return PN_QUERIES_BINNED, x

    def cell_table(self):
        """(cell_start, sorted_index) of the last build as host numpy arrays: the cell
        lists of the reference's DVoV (src/vector_of_vectors.jl:37-44) in CSR form."""
        import torch
        if not self._built:
            raise PointNeighborsError("the neighborhood search is not initialized")
        lib = load()
        nb, ne, mx = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib.pn_grid_stats(self._handle, ctypes.byref(nb), ctypes.byref(ne), ctypes.byref(mx)))
        dev = torch.device("cuda", torch.cuda.current_device())
        C = self.cell_list.list_size if self._hashed() else self.cell_list.n_cells_total
        start = torch.empty(C + 1, dtype=torch.int32, device=dev)
        idx = torch.empty(max(nb.value, 1), dtype=torch.int32, device=dev)
        check(lib.pn_grid_copy_cells(self._handle, ctypes.c_void_p(start.data_ptr()),
                                     ctypes.c_void_p(idx.data_ptr())))
        torch.cuda.synchronize()
        return start.cpu().numpy(), idx[: nb.value].cpu().numpy()

    def points_in_cell(self, cell):
        """Points of one cell (1-based cell tuple), src/nhs_grid.jl:593-597."""
        start, idx = self.cell_table()
        c = self._cell_row(cell)
        return idx[start[c]:start[c + 1]].tolist()

    def _cell_row(self, cell):
        """0-based row of the cell table holding `cell`: its linear index, or its bucket."""
        if self._hashed():
            return self.cell_list.bucket(cell) - 1
        return self.cell_list.cell_index(cell) - 1

    def eachneighbor(self, coords):
        """Candidate points of the 3^D cells around `coords` (src/nhs_grid.jl:585-591)."""
        start, idx = self.cell_table()
        cell = self.host_cell_coords(coords)
        out = []
        rng = [range(c - 1, c + 2) for c in cell]
        for nc in itertools.product(*reversed(rng)):
            nc = tuple(reversed(nc))
            c = self._cell_row(nc)
            out.extend(idx[start[c]:start[c + 1]].tolist())
        return out

    def launch_times(self, max_launches=4096):
        """Durations (ms) of the traversal kernels launched since set_tuning("time_launches", 1),
        in launch order, from HIP events the library records on its stream around each launch;
        waits for the last one and clears the records."""
        buf = (ctypes.c_float * max_launches)()
        n = ctypes.c_int64()
        check(load().pn_grid_launch_times(self._handle, buf, max_launches, ctypes.byref(n)), "launch_times")
        return [buf[i] for i in range(min(n.value, max_launches))]

    def stats(self):
        lib = load()
        nb, ne, mx = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib.pn_grid_stats(self._handle, ctypes.byref(nb), ctypes.byref(ne), ctypes.byref(mx)))
        return nb.value, ne.value, mx.value


def initialize_grid(nhs, y, **kw):
    return nhs.initialize_grid(y, **kw)


def update_grid(nhs, y, **kw):
    return nhs.update_grid(y, **kw)


def copy_grid_search(nhs: GridNeighborhoodSearch, search_radius, n_points, eachpoint=None):
    """copy_neighborhood_search (src/nhs_grid.jl:640-649)."""
    cell_list = nhs.cell_list.copy_cell_list(search_radius, nhs.periodic_box)
    return GridNeighborhoodSearch(nhs.ndims, search_radius=search_radius, n_points=n_points,
                                  periodic_box=nhs.periodic_box, cell_list=cell_list,
                                  update_strategy=nhs.update_strategy)
