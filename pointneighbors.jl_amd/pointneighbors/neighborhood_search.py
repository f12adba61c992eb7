"""The traversal API of src/neighborhood_search.jl on the MI355X engine.

`foreach_point_neighbor(f, x, y, nhs)` (src/neighborhood_search.jl:183-201) takes a
Julia closure in the reference.  A closure cannot be compiled into a HIP kernel across a
C-ABI, so `f` is one of a fixed menu of functors, each a hand-written kernel:

  CountNeighbors(out)          out[i] += 1 per neighbour   (benchmarks/count_neighbors.jl:24-27)
  SumNeighborIndices(out)      out[i] += j per neighbour   (test/neighborhood_search.jl:389-399)

plus the SPH and energy functors of the `sph` module.  Anything else raises
ArgumentError naming the menu; there is no silent fallback.
"""

from __future__ import annotations

import ctypes
import operator

from ._lib import ArgumentError, BoundsError, check, load


class CountNeighbors:
    """Closure `(i, j, pos_diff, d) -> out[i] += 1`; `out` is an int32 device vector
    indexed by the original point index.  accumulate=False overwrites instead."""

    def __init__(self, out, accumulate=True):
        self.out = out
        self.accumulate = accumulate


class SumNeighborIndices:
    """Closure `(i, j, pos_diff, d) -> out[i] += j + index_base` (int64 device vector)."""

    def __init__(self, out, index_base=0, accumulate=True):
        self.out = out
        self.index_base = index_base
        self.accumulate = accumulate


# Per-pair value functions for mapreduce_neighbor (the `f` of
# mapreduce_neighbor(f, op, ...), src/neighborhood_search.jl:340-348).
class NeighborIndex:
    """`(i, j, pos_diff, d) -> j` (0-based j)."""


class One:
    """`(i, j, pos_diff, d) -> 1`."""


class Distance:
    """`(i, j, pos_diff, d) -> d` (the traversal's distance, sqrt of d^2 in the radius type)."""


class DistanceSquared:
    """`(i, j, pos_diff, d) -> d^2` (dot(pos_diff, pos_diff) as the pair predicate computes it)."""


class PosDiff:
    """`(i, j, pos_diff, d) -> pos_diff[k]` (0-based component k)."""

    def __init__(self, k):
        self.k = int(k)


MENU = ("CountNeighbors", "SumNeighborIndices", "SPHDensity", "SPHPressureAcceleration",
        "WendlandPairEnergy", "NBodyAcceleration", "WCSPHInteraction")
# value functions and operators of mapreduce_neighbor (pn_mapreduce_neighbors codes)
_MR_FN = {One: 0, NeighborIndex: 1, Distance: 2, DistanceSquared: 3}
_MR_OP = {operator.add: 0, min: 1, max: 2}


def _menu_error(f):
    raise ArgumentError(
        f"{f!r} cannot run on the MI355X engine: foreach_point_neighbor takes one of the "
        f"kernel functors {', '.join(MENU)}; arbitrary closures cannot cross the C-ABI")


def _check_out(t, n, dtype, name):
    import torch
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise ArgumentError(f"{name} must be a GPU tensor")
    if t.dtype != dtype:
        raise ArgumentError(f"{name} must have dtype {dtype}")
    if t.dim() != 1 or t.shape[0] < n or not t.is_contiguous():
        raise BoundsError(f"{name} must be a contiguous vector of at least {n} entries")


def foreach_point_neighbor(f, system_coords, neighbor_coords, neighborhood_search, *,
                           parallelization_backend=None, points=None):
    """src/neighborhood_search.jl:183-201.  `points` is a 0-based range or index list
    (default: every column of system_coords).  Returns None; outputs land in the
    functor's arrays.  Work is enqueued on torch's current stream."""
    from .nhs_grid import GridNeighborhoodSearch
    from .nhs_precomputed import PrecomputedNeighborhoodSearch
    from . import sph
    nhs = neighborhood_search
    if isinstance(nhs, PrecomputedNeighborhoodSearch):
        return nhs.foreach_point_neighbor(f, system_coords, neighbor_coords, points=points)
    if not isinstance(nhs, GridNeighborhoodSearch):
        raise ArgumentError(f"unsupported neighborhood search {type(nhs).__name__}")
    import torch
    nd = nhs.ndims
    if not isinstance(system_coords, torch.Tensor) or system_coords.dim() != 2 \
            or system_coords.shape[0] != nd:
        raise BoundsError(f"system coordinates must be a {nd} x N matrix")
    nx = system_coords.shape[1]
    if isinstance(f, sph.KernelFunctor):
        return f.run_grid(nhs, system_coords, neighbor_coords, points)
    if not isinstance(f, (CountNeighbors, SumNeighborIndices)):
        _menu_error(f)
    mode, x = nhs._prepare(system_coords, neighbor_coords, points)
    lib = load()
    if isinstance(f, CountNeighbors):
        _check_out(f.out, nx, torch.int32, "CountNeighbors.out")
        check(lib.pn_count_neighbors(nhs._handle, mode, ctypes.c_void_p(f.out.data_ptr()),
                                     1 if f.accumulate else 0), "foreach_point_neighbor")
    else:
        _check_out(f.out, nx, torch.int64, "SumNeighborIndices.out")
        check(lib.pn_sum_neighbor_indices(nhs._handle, mode, ctypes.c_void_p(f.out.data_ptr()),
                                          int(f.index_base), 1 if f.accumulate else 0),
              "foreach_point_neighbor")
    return None


def foreach_point_neighbor_unsafe(f, system_coords, neighbor_coords, neighborhood_search, *,
                                  parallelization_backend=None, points=None):
    """src/neighborhood_search.jl:221-234.  The engine's kernels never read outside the
    grid (queries are bounds-checked while binned), so this is the safe variant."""
    return foreach_point_neighbor(f, system_coords, neighbor_coords, neighborhood_search,
                                  parallelization_backend=parallelization_backend, points=points)


def _mr_codes(f, op, ndims):
    if isinstance(f, PosDiff):
        if not 0 <= f.k < ndims:
            raise ArgumentError(f"PosDiff({f.k}) is out of range for {ndims} dimensions")
        fn = 4 + f.k
    else:
        fn = _MR_FN.get(f if isinstance(f, type) else type(f))
    if fn is None:
        raise ArgumentError(
            f"{f!r} cannot run on the MI355X engine: mapreduce_neighbor takes one of the value "
            "functions One, NeighborIndex, Distance, DistanceSquared, PosDiff(k); arbitrary "
            "closures cannot cross the C-ABI")
    if op not in _MR_OP:
        raise ArgumentError("mapreduce_neighbor on the MI355X engine supports op = +, min, max")
    return fn, _MR_OP[op]


def mapreduce_point_neighbors(f, op, system_coords, neighbor_coords, neighborhood_search, *,
                              init, points=None):
    """mapreduce_neighbor (src/neighborhood_search.jl:340-348) for every selected point in one
    launch: returns (values, counts), float64 and int32 device vectors over the columns of
    system_coords.  values[i] = op(init, reduction over the neighbours of i), or init where i
    has no neighbour (or is not selected); the reduction runs in double."""
    import torch
    from .nhs_grid import GridNeighborhoodSearch
    nhs = neighborhood_search
    if not isinstance(nhs, GridNeighborhoodSearch):
        raise ArgumentError("mapreduce_point_neighbors runs on a GridNeighborhoodSearch")
    fn, opc = _mr_codes(f, op, nhs.ndims)
    mode, x = nhs._prepare(system_coords, neighbor_coords, points)
    n = x.shape[1]
    values = torch.full((n,), float("nan"), dtype=torch.float64, device=x.device)
    counts = torch.zeros(n, dtype=torch.int32, device=x.device)
    check(load().pn_mapreduce_neighbors(nhs._handle, mode, fn, opc, 0, ctypes.c_void_p(values.data_ptr()),
                                        ctypes.c_void_p(counts.data_ptr())), "mapreduce_neighbor")
    init_t = torch.full_like(values, float(init))
    if op is operator.add:
        red = init_t + values
    elif op is min:
        red = torch.minimum(init_t, values)
    else:
        red = torch.maximum(init_t, values)
    return torch.where(counts > 0, red, init_t), counts


def mapreduce_neighbor(f, op, system_coords, neighbor_coords, neighborhood_search, point, *,
                       init):
    """src/neighborhood_search.jl:340-348 for one point: op(init, f(neighbour 1), ...).  f is
    one of One, NeighborIndex (0-based j), Distance, DistanceSquared, PosDiff(k); op is +, min
    or max.  One launch, one host synchronization.  Returns `init` for an empty neighbourhood
    without evaluating f or op (test/neighborhood_search.jl:445-466).  On a precomputed search
    the stored lists give One and NeighborIndex with +."""
    import torch
    from .nhs_precomputed import PrecomputedNeighborhoodSearch
    nhs = neighborhood_search
    n = system_coords.shape[1]
    if not 0 <= point < n:
        raise BoundsError(f"attempt to access {n}-element columns at index {point}")
    integral = (f in (One, NeighborIndex) or isinstance(f, (One, NeighborIndex))) and op is operator.add
    if isinstance(nhs, PrecomputedNeighborhoodSearch):
        _mr_codes(f, op, nhs.ndims)
        if not integral:
            raise ArgumentError("mapreduce_neighbor over precomputed lists supports One and "
                                "NeighborIndex with op = +")
        dev = system_coords.device
        cnt = torch.zeros(n, dtype=torch.int32, device=dev)
        foreach_point_neighbor(CountNeighbors(cnt), system_coords, neighbor_coords, nhs, points=[point])
        c = int(cnt[point])
        if c == 0:
            return init
        if f is One or isinstance(f, One):
            return init + c
        out = torch.zeros(n, dtype=torch.int64, device=dev)
        foreach_point_neighbor(SumNeighborIndices(out), system_coords, neighbor_coords, nhs, points=[point])
        return init + int(out[point])
    values, counts = mapreduce_point_neighbors(f, op, system_coords, neighbor_coords, nhs, init=init,
                                               points=[point])
    v, c = values[point].item(), int(counts[point].item())
    if c == 0:
        return init
    return int(round(v)) if integral and isinstance(init, int) else v


mapreduce_neighbor_unsafe = mapreduce_neighbor


def foreach_neighbor(f, system_coords, neighbor_coords, neighborhood_search, point):
    """src/neighborhood_search.jl:250-263 for one point."""
    return foreach_point_neighbor(f, system_coords, neighbor_coords, neighborhood_search,
                                  points=[point])


foreach_neighbor_unsafe = foreach_neighbor


def initialize(search, x, y, *, parallelization_backend=None, eachindex_y=None):
    """initialize!(search, x, y; eachindex_y) (src/nhs_grid.jl:220-225,
    src/nhs_precomputed.jl:130-147).  `eachindex_y` is a 0-based range or index list."""
    return search.initialize(x, y, eachindex_y=eachindex_y)


def update(search, x, y, *, points_moving=(True, True), parallelization_backend=None,
           eachindex_y=None, **kw):
    """update!(search, x, y; points_moving, eachindex_y) (src/nhs_grid.jl:283-292,
    src/nhs_precomputed.jl:149-169)."""
    return search.update(x, y, points_moving=points_moving, eachindex_y=eachindex_y, **kw)


def requires_update(search):
    return search.requires_update()


def copy_neighborhood_search(search, search_radius, n_points, *, eachpoint=None):
    """src/neighborhood_search.jl:107-110 with the grid and precomputed methods."""
    return search.copy(search_radius, n_points)


def freeze_neighborhood_search(search):
    """src/neighborhood_search.jl:112-118."""
    return search.freeze() if hasattr(search, "freeze") else search


def search_radius(search):
    return search.search_radius
