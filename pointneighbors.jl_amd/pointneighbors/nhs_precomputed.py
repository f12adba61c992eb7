"""PrecomputedNeighborhoodSearch (src/nhs_precomputed.jl) on the MI355X engine.

The neighbour lists are CSR (row_ptr, col_idx) built by count -> scan -> fill ->
sort-rows kernels, instead of the reference's max_neighbors x N column matrix filled by
pushes and sorted with a full vendor column sort (src/nhs_precomputed.jl:187-206,
src/vector_of_vectors.jl:187-212).  Rows are exact-length and ascending.
"""

from __future__ import annotations

import ctypes

from . import _device
from ._lib import ArgumentError, BoundsError, PointNeighborsError, check, load
from .cell_lists import FullGridCellList
from .nhs_grid import GridNeighborhoodSearch, _GridMeta
from .util import FLOAT64_EPS, scalar_dtype

# max_points_per_cell of an inner grid the search makes for itself: the grid's largest
# accepted value (a 3^D stencil's candidate count must fit in int32), i.e. no capacity
AUTO_GRID_CAPACITY = 2**31 // 27


def max_neighbors_default(ndims):
    """src/nhs_precomputed.jl:114-124 (the code's values; the docstring says 324)."""
    if ndims == 1:
        return 32
    if ndims == 2:
        return 64
    if ndims == 3:
        return 320
    raise ArgumentError("`NDIMS` must be 1, 2, or 3")


class NeighborLists:
    """CSR neighbour lists: row i holds the sorted 0-based neighbours of point i."""

    def __init__(self, row_ptr=None, col_idx=None, n=0, max_neighbors=0):
        self.row_ptr = row_ptr
        self.col_idx = col_idx
        self.n = n
        self.max_neighbors = max_neighbors

    def __len__(self):
        return self.n

    def __getitem__(self, i):
        if not (0 <= i < self.n):
            raise BoundsError(f"attempt to access {self.n}-element neighbor lists at index [{i}]")
        rp = self.row_ptr[i:i + 2].cpu().numpy()
        return self.col_idx[int(rp[0]):int(rp[1])].cpu().numpy().tolist()

    def lengths(self):
        return (self.row_ptr[1:] - self.row_ptr[:-1])

    def to_host(self):
        return self.row_ptr.cpu().numpy(), self.col_idx.cpu().numpy()

    def to_dvov(self, max_neighbors=None, transpose_backend=False, max_outer_length=None):
        """The reference's DynamicVectorOfVectors{Int32} layout (src/vector_of_vectors.jl:3-31) of
        these lists, on the device: (backend, lengths).  backend is an (N, max_neighbors) int32
        tensor whose row i holds point i's neighbours (Julia's Array{Int32,2}(max_neighbors, N));
        with transpose_backend it is a (max_neighbors, max_outer_length) tensor whose row k holds
        the k-th neighbours of all points (the row-major array behind the reference's
        PermutedDimsArray).  Slots past a list's length are zero here."""
        import torch
        max_neighbors = int(self.max_neighbors if max_neighbors is None else max_neighbors)
        n = self.n
        if max_neighbors <= 0:
            raise ArgumentError("max_neighbors must be positive")
        longest = int(self.lengths().max().item()) if n else 0
        if longest > max_neighbors:
            raise PointNeighborsError(f"a neighbor list has {longest} entries, more than max_neighbors = "
                                      f"{max_neighbors}")
        dev = self.row_ptr.device
        ld = n if max_outer_length is None else int(max_outer_length)
        if ld < n:
            raise ArgumentError("max_outer_length must be at least the number of points")
        shape = (max_neighbors, ld) if transpose_backend else (n, max_neighbors)
        backend = torch.zeros(shape, dtype=torch.int32, device=dev)
        lengths = torch.zeros(n, dtype=torch.int32, device=dev)
        check(load().pn_csr_to_dvov(ctypes.c_void_p(self.row_ptr.data_ptr()), ctypes.c_void_p(self.col_idx.data_ptr()),
                                    n, max_neighbors, 1 if transpose_backend else 0, ld,
                                    ctypes.c_void_p(backend.data_ptr()), ctypes.c_void_p(lengths.data_ptr()),
                                    ctypes.c_void_p(_device.current_stream_handle(dev))), "to_dvov")
        return backend, lengths


class PrecomputedNeighborhoodSearch(metaclass=_GridMeta):
    """PrecomputedNeighborhoodSearch{NDIMS}(; search_radius = 0.0, n_points = 0,
    periodic_box = nothing, update_strategy = nothing, update_neighborhood_search,
    backend, transpose_backend = false, max_neighbors, sort_neighbor_lists = true)
    (src/nhs_precomputed.jl:89-111).

    `update_neighborhood_search` must be a GridNeighborhoodSearch with a FullGridCellList;
    when omitted, one over the bounding box of the points (padded by the search radius)
    is created at initialize!.  Lists are always sorted (CSR fill + row sort);
    `transpose_backend` only selects the reference's memory layout and has no effect on
    the CSR storage.
    """

    def __init__(self, ndims, *, search_radius=0.0, n_points=0, periodic_box=None,
                 update_strategy=None, update_neighborhood_search=None, backend=None,
                 transpose_backend=False, max_neighbors=None, sort_neighbor_lists=True,
                 _lists=None):
        r_dt = scalar_dtype(search_radius)
        if r_dt.kind in "iub":
            raise ArgumentError("`search_radius` cannot be an integer type, since computed "
                                "distances will be converted to this type")
        if update_neighborhood_search is not None and not isinstance(
                update_neighborhood_search, GridNeighborhoodSearch):
            raise ArgumentError("update_neighborhood_search must be a GridNeighborhoodSearch "
                                "with a FullGridCellList")
        self.ndims = ndims
        self.search_radius = search_radius
        self.radius_dtype = r_dt
        self.n_points = n_points
        self.periodic_box = periodic_box
        self.update_strategy = update_strategy
        self.neighborhood_search = update_neighborhood_search
        self.transpose_backend = transpose_backend
        self.max_neighbors = max_neighbors_default(ndims) if max_neighbors is None else int(max_neighbors)
        self.sort_neighbor_lists = sort_neighbor_lists
        self.neighbor_lists = _lists if _lists is not None else NeighborLists(max_neighbors=self.max_neighbors)
        self._frozen = False
        self._auto = False  # the inner grid was made by this search (no user grid given)

    def requires_update(self):
        """src/nhs_precomputed.jl:128"""
        return (True, True)

    def eltype(self):
        return self.radius_dtype

    def _auto_grid(self, x, y):
        import torch
        r = self.search_radius
        if self.periodic_box is not None:
            # the grid spans the periodic box; cells wrap, coordinates need not be inside
            pb = self.periodic_box
            cl = FullGridCellList(min_corner=pb.min_corner, max_corner=pb.max_corner, search_radius=r)
            return GridNeighborhoodSearch(self.ndims, search_radius=r, n_points=self.n_points,
                                          periodic_box=pb, cell_list=cl, update_strategy=self.update_strategy)
        sets = [t for t in ((x,) if _device.same_storage(x, y) else (x, y)) if t.shape[1]]
        if sets:
            # row-major D x N copies reduce along contiguous rows (a reduction over the
            # strided point dimension runs on a few workgroups only)
            ext = [torch.aminmax(t.contiguous(), dim=1) for t in sets]
            lo = torch.stack([e.min for e in ext]).amin(dim=0).cpu().numpy()
            hi = torch.stack([e.max for e in ext]).amax(dim=0).cpu().numpy()
        else:
            import numpy as np
            lo = hi = np.zeros(self.ndims, dtype=_device.np_dtype(x))
        # no per-cell capacity (the reference's default inner search is an unbounded
        # DictionaryCellList); the largest limit the grid accepts
        cl = FullGridCellList(min_corner=lo - r, max_corner=hi + r, search_radius=r,
                              max_points_per_cell=AUTO_GRID_CAPACITY)
        return GridNeighborhoodSearch(self.ndims, search_radius=r, n_points=self.n_points,
                                      cell_list=cl, update_strategy=self.update_strategy)

    def _update_inner(self, x, y, points_moving, rebuild_lists):
        """Updates the inner search and the lists.  A grid this search made for itself
        covers the bounding box it was made for; when points have left it (the grid's
        bounds error), it is remade over the current points and the update repeated, so
        points may move anywhere, as with the reference's unbounded default."""
        if self.neighborhood_search is None:
            self.neighborhood_search = self._auto_grid(x, y)
            self._auto = True
        try:
            self.neighborhood_search.update(x, y, points_moving=points_moving)
            if rebuild_lists:
                self._build_lists(x, y)
        except (PointNeighborsError, BoundsError):
            if not self._auto:
                raise
            self.neighborhood_search = self._auto_grid(x, y)
            self.neighborhood_search.update(x, y, points_moving=points_moving)
            if rebuild_lists:
                self._build_lists(x, y)

    def initialize(self, x, y, *, eachindex_y=None):
        """src/nhs_precomputed.jl:130-147"""
        if not _device.is_full(eachindex_y, y.shape[1]):
            raise PointNeighborsError("this neighborhood search does not support inactive points")
        if self._frozen:
            raise PointNeighborsError("a frozen neighborhood search cannot be initialized")
        self._update_inner(x, y, (True, True), True)
        return self

    def update(self, x, y, *, points_moving=(True, True), eachindex_y=None):
        """src/nhs_precomputed.jl:149-169"""
        if not _device.is_full(eachindex_y, y.shape[1]):
            raise PointNeighborsError("this neighborhood search does not support inactive points")
        if self._frozen:
            raise PointNeighborsError("a frozen neighborhood search cannot be updated")
        self._update_inner(x, y, points_moving, any(points_moving))
        return self

    def _build_lists(self, x, y):
        """initialize_neighbor_lists! (src/nhs_precomputed.jl:187-206) as count -> scan ->
        fill -> sort rows."""
        import torch
        nhs = self.neighborhood_search
        if float(self.search_radius) < FLOAT64_EPS:
            n = x.shape[1]
            self.neighbor_lists = NeighborLists(torch.zeros(n + 1, dtype=torch.int32, device=x.device),
                                                torch.zeros(0, dtype=torch.int32, device=x.device), n,
                                                self.max_neighbors)
            return
        mode, xx = nhs._prepare(x, y, None)
        n = xx.shape[1]
        lib = load()
        row_ptr = torch.empty(n + 1, dtype=torch.int32, device=x.device)
        nnz, mx = ctypes.c_int64(), ctypes.c_int64()
        check(lib.pn_csr_count(nhs._handle, mode, n, ctypes.c_void_p(row_ptr.data_ptr()),
                               int(self.max_neighbors), ctypes.byref(nnz), ctypes.byref(mx)),
              "PrecomputedNeighborhoodSearch")
        col = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=x.device)
        check(lib.pn_csr_fill(nhs._handle, mode, n, ctypes.c_void_p(row_ptr.data_ptr()),
                              ctypes.c_void_p(col.data_ptr()), 0), "PrecomputedNeighborhoodSearch")
        self.neighbor_lists = NeighborLists(row_ptr, col[: nnz.value], n, self.max_neighbors)

    def foreach_point_neighbor(self, f, x, y, points=None):
        """mapreduce_neighbor_inner(::PrecomputedNeighborhoodSearch)
        (src/nhs_precomputed.jl:210-247): stored neighbours, no radius test."""
        import torch
        from .neighborhood_search import CountNeighbors, SumNeighborIndices, _check_out, _menu_error
        from . import sph
        nl = self.neighbor_lists
        if nl.row_ptr is None:
            raise PointNeighborsError("the neighborhood search is not initialized")
        n = nl.n
        if x.shape[1] > n:
            raise BoundsError("system coordinates have more points than the neighbor lists")
        if isinstance(f, sph.KernelFunctor):
            return f.run_lists(self, x, y, points)
        first, count, ptr, keep = _device.selection(points, x.shape[1], x.device)
        stream = ctypes.c_void_p(_device.current_stream_handle(x.device))
        lib = load()
        sel = ctypes.c_void_p(ptr) if ptr else None
        if isinstance(f, CountNeighbors):
            _check_out(f.out, x.shape[1], torch.int32, "CountNeighbors.out")
            check(lib.pn_list_count_neighbors(ctypes.c_void_p(nl.row_ptr.data_ptr()), n, first, count, sel,
                                              ctypes.c_void_p(f.out.data_ptr()), 1 if f.accumulate else 0,
                                              stream))
        elif isinstance(f, SumNeighborIndices):
            _check_out(f.out, x.shape[1], torch.int64, "SumNeighborIndices.out")
            check(lib.pn_list_sum_neighbor_indices(ctypes.c_void_p(nl.row_ptr.data_ptr()),
                                                   ctypes.c_void_p(nl.col_idx.data_ptr()), n, 0, first,
                                                   count, sel, ctypes.c_void_p(f.out.data_ptr()),
                                                   int(f.index_base), 1 if f.accumulate else 0, stream))
        else:
            _menu_error(f)
        del keep
        return None

    def copy(self, search_radius, n_points):
        """copy_neighborhood_search (src/nhs_precomputed.jl:249-266)."""
        inner = None
        if self.neighborhood_search is not None:
            from .nhs_grid import copy_grid_search
            inner = copy_grid_search(self.neighborhood_search, search_radius, n_points)
        return PrecomputedNeighborhoodSearch(self.ndims, search_radius=search_radius, n_points=n_points,
                                             periodic_box=self.periodic_box,
                                             update_neighborhood_search=inner,
                                             transpose_backend=self.transpose_backend,
                                             max_neighbors=self.max_neighbors,
                                             sort_neighbor_lists=self.sort_neighbor_lists)

    def freeze(self):
        """freeze_neighborhood_search (src/nhs_precomputed.jl:268-277): drops the inner
        grid; the lists stay."""
        p = PrecomputedNeighborhoodSearch(self.ndims, search_radius=self.search_radius,
                                          n_points=self.n_points, max_neighbors=self.max_neighbors,
                                          sort_neighbor_lists=self.sort_neighbor_lists,
                                          transpose_backend=self.transpose_backend,
                                          _lists=self.neighbor_lists)
        p._frozen = True
        return p
