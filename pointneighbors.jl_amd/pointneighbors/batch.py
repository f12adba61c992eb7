"""Many independent point clouds searched in one launch.

The reference searches each point set with its own GridNeighborhoodSearch (made with
copy_neighborhood_search, src/neighborhood_search.jl:107-110) and launches per set.  A
BatchedGridNeighborhoodSearch concatenates the clouds' cell tables: one update bins every
cloud and one traversal serves every cloud, with a persistent traversal grid whose shape
does not depend on the data, so an update + query step can be captured in a graph.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _device
from ._lib import PN_QUERIES_GRID, ArgumentError, PointNeighborsError, check, load
from .cell_lists import FullGridCellList
from .util import dtype_code, promote, scalar_dtype


class BatchedGridNeighborhoodSearch:
    """`cell_lists[c]` is cloud c's FullGridCellList (all with `search_radius`);
    `n_points[c]` its number of points.  Points are passed concatenated, cloud by
    cloud, as one D x N matrix."""

    def __init__(self, ndims, *, search_radius, cell_lists, n_points, max_points_per_cell=100):
        if len(cell_lists) != len(n_points) or not cell_lists:
            raise ArgumentError("one cell list and one point count per cloud are required")
        for cl in cell_lists:
            if not isinstance(cl, FullGridCellList) or cl.ndims != ndims or cl.is_template:
                raise ArgumentError(f"every cloud needs a {ndims}D FullGridCellList with a search radius")
        r_dt = scalar_dtype(search_radius)
        if r_dt.kind in "iub":
            raise ArgumentError("`search_radius` cannot be an integer type, since computed "
                                "distances will be converted to this type")
        self.ndims = ndims
        self.search_radius = search_radius
        self.radius_dtype = r_dt
        self.cell_lists = list(cell_lists)
        self.n_points = np.asarray(n_points, dtype=np.int64)
        self.point_offsets = np.concatenate([[0], np.cumsum(self.n_points)]).astype(np.int64)
        self.max_points_per_cell = max_points_per_cell
        self._handle = None
        self._coord_dtype = None

    @property
    def n_clouds(self):
        return len(self.cell_lists)

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None:
            try:
                load().pn_grid_destroy(h)
            except Exception:
                pass

    def _ensure_handle(self, coord_dtype):
        if self._handle is not None and self._coord_dtype == coord_dtype:
            return
        lib = load()
        if self._handle is not None:
            lib.pn_grid_destroy(self._handle)
            self._handle = None
        corner_dt = promote(*[cl.corner_dtype for cl in self.cell_lists])
        cell_dt = promote(coord_dtype, corner_dt)
        mins = np.concatenate([cl.min_corner.astype(np.float64) for cl in self.cell_lists])
        ncs = np.concatenate([np.asarray(cl.n_cells, dtype=np.int64) for cl in self.cell_lists])
        offs = np.ascontiguousarray(self.point_offsets)
        h = ctypes.c_void_p()
        check(lib.pn_batch_create(self.ndims, dtype_code(coord_dtype), dtype_code(self.radius_dtype),
                                  dtype_code(cell_dt), self.n_clouds,
                                  mins.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  ncs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), float(self.search_radius),
                                  int(self.max_points_per_cell),
                                  offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(h)),
              "BatchedGridNeighborhoodSearch")
        self._handle = h
        self._coord_dtype = coord_dtype

    def _stream(self, t):
        load().pn_grid_set_stream(self._handle, ctypes.c_void_p(_device.current_stream_handle(t.device)))

    def update(self, y, *, check_errors=True):
        """Bins every cloud.  check_errors=False only enqueues (for graph capture); call
        check() later to raise the reference's errors."""
        y = _device.as_coordinates(y, self.ndims, "neighbor coordinates")
        if y.shape[1] != int(self.point_offsets[-1]):
            raise ArgumentError("the batch is updated with the concatenated points of all clouds")
        self._ensure_handle(_device.np_dtype(y))
        self._stream(y)
        lib = load()
        fn = lib.pn_grid_update if check_errors else lib.pn_grid_update_async
        check(fn(self._handle, ctypes.c_void_p(y.data_ptr()), y.shape[1], 0, y.shape[1], None),
              "BatchedGridNeighborhoodSearch.update")
        return self

    initialize = update

    def check(self):
        check(load().pn_grid_check(self._handle), "BatchedGridNeighborhoodSearch")

    def count_neighbors(self, out, accumulate=False):
        """Neighbour counts of every point of every cloud (concatenated order)."""
        if self._handle is None:
            raise PointNeighborsError("the neighborhood search is not initialized")
        self._stream(out)
        check(load().pn_count_neighbors(self._handle, PN_QUERIES_GRID, ctypes.c_void_p(out.data_ptr()),
                                        1 if accumulate else 0), "count_neighbors")

    def set_tuning(self, key, value):
        """Same knobs as GridNeighborhoodSearch.set_tuning."""
        if self._handle is None:
            raise PointNeighborsError("update the batch once before tuning it")
        check(load().pn_set_tuning(self._handle, key.encode(), int(value)), "set_tuning")


def displace_uniform(x, amplitude, seed, step):
    """x += U[-amplitude, amplitude] per coordinate (seeded by seed and the int64 device
    counter `step`, which is incremented)."""
    import torch
    code = 0 if x.dtype == torch.float32 else 1
    check(load().pn_displace_uniform(ctypes.c_void_p(x.data_ptr()), x.numel(), code, float(amplitude),
                                     int(seed) & 0xFFFFFFFFFFFFFFFF, ctypes.c_void_p(step.data_ptr()),
                                     ctypes.c_void_p(_device.current_stream_handle(x.device))), "displace_uniform")
    # an in-place write torch did not see: bump the version counter so a search that
    # binned x re-reads its coordinates
    torch.autograd.graph.increment_version(x)
