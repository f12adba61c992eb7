"""PointNeighbors.jl's GPU hot path, MI355X-native.

Host mirror of the reference's API (GridNeighborhoodSearch + FullGridCellList +
PrecomputedNeighborhoodSearch, initialize!/update!/foreach_point_neighbor) over the
C-ABI of libpointneighbors_hip.so.  Julia's `initialize!`/`update!` are `initialize` /
`update` here; indices are 0-based; coordinate matrices are D x N torch tensors on the
GPU (strides (1, D), the memory of Julia's column-major matrix).
"""

from ._lib import ArgumentError, BoundsError, PointNeighborsError, load
from .cell_lists import DynamicVectorOfVectors, FullGridCellList, SpatialHashingCellList
from .neighborhood_search import (CountNeighbors, Distance, DistanceSquared, NeighborIndex, One, PosDiff,
                                  SumNeighborIndices, mapreduce_point_neighbors,
                                  copy_neighborhood_search, foreach_neighbor,
                                  foreach_neighbor_unsafe, foreach_point_neighbor,
                                  foreach_point_neighbor_unsafe, freeze_neighborhood_search,
                                  initialize, mapreduce_neighbor, mapreduce_neighbor_unsafe,
                                  requires_update, search_radius, update)
from .nhs_grid import (GridNeighborhoodSearch, ParallelIncrementalUpdate, ParallelUpdate,
                       PeriodicBox, SemiParallelUpdate, SerialIncrementalUpdate, SerialUpdate,
                       initialize_grid, update_grid)
from .nhs_precomputed import NeighborLists, PrecomputedNeighborhoodSearch
from .interactions import NBodyAcceleration, WCSPHInteraction, state_equation_cole
from .util import floor_to_int

__all__ = [
    "ArgumentError", "BoundsError", "PointNeighborsError", "load",
    "DynamicVectorOfVectors", "FullGridCellList", "SpatialHashingCellList",
    "CountNeighbors", "Distance", "DistanceSquared", "NeighborIndex", "One", "PosDiff",
    "SumNeighborIndices", "mapreduce_point_neighbors", "NBodyAcceleration", "WCSPHInteraction",
    "state_equation_cole",
    "copy_neighborhood_search", "foreach_neighbor", "foreach_neighbor_unsafe",
    "foreach_point_neighbor", "foreach_point_neighbor_unsafe", "freeze_neighborhood_search",
    "initialize", "mapreduce_neighbor", "mapreduce_neighbor_unsafe", "requires_update",
    "search_radius", "update",
    "GridNeighborhoodSearch", "ParallelIncrementalUpdate", "ParallelUpdate", "PeriodicBox",
    "SemiParallelUpdate", "SerialIncrementalUpdate", "SerialUpdate", "initialize_grid",
    "update_grid", "NeighborLists", "PrecomputedNeighborhoodSearch", "floor_to_int",
]
