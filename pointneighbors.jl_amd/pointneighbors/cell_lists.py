"""FullGridCellList (src/cell_lists/full_grid.jl) on the MI355X engine.

The cell storage is not the reference's fixed-capacity column matrix but the result of
the device counting sort: `cell_start` offsets into a cell-sorted permutation.  The
reference's capacity, `max_points_per_cell`, is kept as a limit so the "cell list is
full" error fires exactly where the reference raises it.
"""

from __future__ import annotations

import math

import numpy as np

from ._lib import ArgumentError, PointNeighborsError
from .util import FLOAT64_EPS, array_dtype, promote, scalar_dtype, to_float_array


class DynamicVectorOfVectors:
    """Marker for the reference's default cell storage backend
    (`DynamicVectorOfVectors{Int32}`, src/vector_of_vectors.jl:3-31).  On the MI355X
    engine it selects the counting-sort storage."""

    def __class_getitem__(cls, eltype):
        return cls


def _check_backend(backend):
    if backend is None or backend is DynamicVectorOfVectors or isinstance(backend, DynamicVectorOfVectors):
        return
    raise ArgumentError(
        f"backend {backend!r} is not supported by the MI355X engine; "
        "use DynamicVectorOfVectors (the default)")


class FullGridCellList:
    """FullGridCellList(; min_corner, max_corner, search_radius = 0, backend,
    max_points_per_cell = 100) — src/cell_lists/full_grid.jl:48-82.

    Pads both corners by `1001 // 1000 * search_radius` in the radius type
    (`Rational{Int} * Float` converts 1001//1000 as Float(1001)/Float(1000)), stores the
    padded corners in promote(corner type, radius type) and sets
    `n_cells = ceil.(Int, (max_corner .- min_corner) ./ search_radius)`.
    With `search_radius < eps()` it is an empty template for copy_neighborhood_search.
    """

    def __init__(self, *, min_corner, max_corner, search_radius=0.0,
                 backend=DynamicVectorOfVectors, max_points_per_cell=100):
        if len(min_corner) != len(max_corner):
            raise ArgumentError("min_corner and max_corner must have the same length")
        if len(min_corner) > 100:
            raise ArgumentError("FullGridCellList only supports up to 100 dimensions, "
                                "check your `min_corner` and `max_corner`")
        _check_backend(backend)
        user_dt = promote(array_dtype(min_corner), array_dtype(max_corner))
        r_dt = scalar_dtype(search_radius)
        if r_dt.kind != "f":
            r_dt = np.dtype(np.float64) if r_dt.kind in "iu" else r_dt
        corner_dt = promote(user_dt, r_dt)
        self.user_min_corner = to_float_array(min_corner, user_dt)
        self.user_max_corner = to_float_array(max_corner, user_dt)
        self.search_radius = search_radius
        self.radius_dtype = r_dt
        self.corner_dtype = corner_dt
        self.max_points_per_cell = int(max_points_per_cell)
        rr = r_dt.type(search_radius)
        pad = (r_dt.type(1001) / r_dt.type(1000)) * rr
        self.min_corner = (self.user_min_corner.astype(corner_dt) - corner_dt.type(pad)).astype(corner_dt)
        self.max_corner = (self.user_max_corner.astype(corner_dt) + corner_dt.type(pad)).astype(corner_dt)
        ndims = len(self.min_corner)
        if float(search_radius) < FLOAT64_EPS:
            self.n_cells = tuple(0 for _ in range(ndims))
            self.is_template = True
        else:
            q = (self.max_corner - self.min_corner) / corner_dt.type(rr)
            self.n_cells = tuple(int(math.ceil(float(v))) for v in q.astype(corner_dt))
            self.is_template = False

    @property
    def ndims(self) -> int:
        return len(self.min_corner)

    @property
    def linear_indices_size(self) -> tuple:
        return self.n_cells

    @property
    def n_cells_total(self) -> int:
        return int(np.prod(self.n_cells)) if self.n_cells else 0

    def cell_index(self, cell) -> int:
        """LinearIndices(n_cells)[cell...] (src/cell_lists/full_grid.jl:157-161),
        1-based cell tuple -> 1-based linear index."""
        c, stride = 0, 1
        for d, cd in enumerate(cell):
            if not (1 <= cd <= self.n_cells[d]):
                from ._lib import BoundsError
                raise BoundsError(f"cell {tuple(cell)} out of bounds of LinearIndices{self.n_cells}")
            c += (cd - 1) * stride
            stride *= self.n_cells[d]
        return c + 1

    def copy_cell_list(self, search_radius, periodic_box=None):
        """src/cell_lists/full_grid.jl:179-185 (re-pads the stored corners, as the
        reference does)."""
        return FullGridCellList(min_corner=self.min_corner, max_corner=self.max_corner,
                                search_radius=search_radius,
                                max_points_per_cell=self.max_points_per_cell)

    def bounds_error_message(self, cell) -> str:
        """The CPU-variant message of check_cell_bounds (src/cell_lists/full_grid.jl:194-198)."""
        size_ = ", ".join(f"2:{n - 1}" for n in self.n_cells)
        return ("particle coordinates are NaN or outside the domain bounds of the cell list\n"
                f"cell {tuple(int(c) for c in cell)} is out of bounds for cell grid of size [{size_}]")

    def __repr__(self):
        return (f"FullGridCellList(min_corner={self.min_corner.tolist()}, "
                f"max_corner={self.max_corner.tolist()}, n_cells={self.n_cells})")


def spatial_hash(cell, list_size) -> int:
    """spatial_hash (src/cell_lists/spatial_hashing.jl:150-164, Ihmsen et al. 2011), 1-based:
    mod(xor(i * 73856093, j * 19349663, k * 83492791), list_size) + 1 in wrapping Int64."""
    m64 = (1 << 64) - 1
    h = 0
    for c, k in zip(cell, (73856093, 19349663, 83492791)):
        h ^= (int(c) * k) & m64
    if h >= 1 << 63:
        h -= 1 << 64
    return h % int(list_size) + 1


class _HashMeta(type):
    def __getitem__(cls, ndims):
        return lambda **kw: cls(ndims, **kw)


class SpatialHashingCellList(metaclass=_HashMeta):
    """SpatialHashingCellList{NDIMS}(; list_size, backend, max_points_per_cell = 100) —
    src/cell_lists/spatial_hashing.jl:1-183.  Cells of size search_radius cover all of R^D;
    a cell's points live in bucket spatial_hash(cell) of `list_size` buckets (about twice the
    number of points is the reference's recommendation), and `max_points_per_cell` bounds a
    bucket.  Use as `SpatialHashingCellList[D](list_size=2 * n)`."""

    def __init__(self, ndims, *, list_size, backend=DynamicVectorOfVectors, max_points_per_cell=100):
        _check_backend(backend)
        if int(ndims) not in (1, 2, 3):
            raise ArgumentError("SpatialHashingCellList supports 1, 2 or 3 dimensions")
        if int(list_size) < 1:
            raise ArgumentError("list_size must be positive")
        self.ndims = int(ndims)
        self.list_size = int(list_size)
        self.max_points_per_cell = int(max_points_per_cell)
        self.is_template = False

    def copy_cell_list(self, search_radius, periodic_box=None):
        """src/cell_lists/spatial_hashing.jl:125-134: same list size and capacity."""
        return SpatialHashingCellList(self.ndims, list_size=self.list_size,
                                      max_points_per_cell=self.max_points_per_cell)

    def bucket(self, cell) -> int:
        """1-based bucket of a cell (the reference's `cell_list[cell]` reads this one)."""
        return spatial_hash(cell, self.list_size)

    def bounds_error_message(self, cell) -> str:
        return ("InexactError: a cell coordinate does not fit in Int32 (coordinates_flattened, "
                "src/cell_lists/spatial_hashing.jl:170-183); NaN or very large particle coordinates")

    def __repr__(self):
        return f"SpatialHashingCellList{{{self.ndims}}}(list_size={self.list_size})"


def empty_template_error():
    raise PointNeighborsError("`search_radius` is not defined for this cell list")
