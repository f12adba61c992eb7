"""SpatialHashingCellList host logic (src/cell_lists/spatial_hashing.jl): the bucket function
pinned by closed forms (single-term cells reduce to c * prime mod L; two and three terms are
checked against numpy's wrapping int64 arithmetic, a separate implementation), the reference's
collision fixtures (test/cell_lists/spatial_hashing.jl:2-70: cells (-1, 0) and (-2, -1) share a
bucket for list_size 1 and 2), cell coordinates without corner or +1, and the API.  CPU only."""

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors.cell_lists import spatial_hash

P = (73856093, 19349663, 83492791)


@pytest.mark.parametrize("L", [1, 2, 7, 1000003, 2 * 1030301])
def test_single_term_closed_forms(L):
    for c in (0, 1, -1, 5, -12345, 2**31 - 1, -(2**31)):
        assert spatial_hash((c,), L) == (c * P[0]) % L + 1
        assert spatial_hash((0, c), L) == (c * P[1]) % L + 1
        assert spatial_hash((0, 0, c), L) == (c * P[2]) % L + 1
    assert spatial_hash((-1,), 10) == 10 - (P[0] % 10) + 1  # Julia mod: non-negative


@pytest.mark.parametrize("L", [3, 1000, 2**31 - 2])
def test_xor_terms_match_numpy_int64(L):
    rng = np.random.default_rng(4)
    cells = rng.integers(-(2**31), 2**31, size=(200, 3))
    cells[:5] = [[2**31 - 1, -(2**31), 5], [-1, -1, -1], [1, 1, 1], [0, 0, 0], [-(2**31), -(2**31), -(2**31)]]
    with np.errstate(over="ignore"):
        h = (cells[:, 0].astype(np.int64) * np.int64(P[0])) ^ (cells[:, 1].astype(np.int64) * np.int64(P[1])) \
            ^ (cells[:, 2].astype(np.int64) * np.int64(P[2]))
    want = np.mod(h, np.int64(L)) + 1
    got = [spatial_hash(tuple(int(v) for v in c), L) for c in cells]
    assert got == want.tolist()


def test_reference_collision_cells():
    for L in (1, 2):
        assert spatial_hash((-1, 0), L) == spatial_hash((-2, -1), L)


def test_cell_coordinates_and_api():
    r = 0.1 + 10 * np.finfo(np.float64).eps
    cl = pn.SpatialHashingCellList[2](list_size=1)
    nhs = pn.GridNeighborhoodSearch(2, search_radius=r, n_points=1, cell_list=cl)
    # "The point is in cell (-1, 0)" (test/cell_lists/spatial_hashing.jl:3)
    assert nhs.host_cell_coords(np.array([-0.05, 0.05])) == (-1, 0)
    assert nhs.host_cell_coords(np.array([-0.15, -0.05])) == (-2, -1)
    c2 = cl.copy_cell_list(0.5)
    assert (c2.list_size, c2.max_points_per_cell, c2.ndims) == (1, 100, 2)
    per = pn.GridNeighborhoodSearch(2, search_radius=1.0, cell_list=pn.SpatialHashingCellList[2](list_size=4),
                                    periodic_box=pn.PeriodicBox(min_corner=[0.0, 0.0], max_corner=[4.0, 4.0]))
    # periodic_cell_index of floor_to_int.(x ./ cell_size): 4 cells of size 1, wrapped to 2:5
    assert per.host_cell_coords(np.array([0.5, 3.5])) == (4, 3)
    assert per.host_cell_coords(np.array([-0.5, 4.5])) == (3, 4)
    with pytest.raises(pn.ArgumentError):
        pn.SpatialHashingCellList[2](list_size=0)
    cp = pn.copy_neighborhood_search(nhs, 0.3, 5)
    assert isinstance(cp.cell_list, pn.SpatialHashingCellList) and cp.cell_list.list_size == 1
