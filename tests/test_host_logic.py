"""Host-side logic of the Python mirror (constructors, geometry, errors, inputs) and the
C-ABI library's exports.  CPU only: no GPU compute is called."""

import ctypes
import math
import os
import re

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import synthetic
from pointneighbors._lib import SIGNATURES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---- FullGridCellList geometry against the oracle and the fixture ---------------------------

@pytest.mark.parametrize("n", [1, 2, 3])
def test_fullgrid_padding_matches_fixture(golden, orc, n):
    pc = golden["padded_corners"]
    cl = pn.FullGridCellList(min_corner=[pc["min_corner"]] * n, max_corner=[pc["max_corner"]] * n,
                             search_radius=pc["search_radius"])
    assert cl.min_corner.tolist() == [pc["padded_min"]] * n
    assert cl.max_corner.tolist() == [10.0 + 1.001] * n
    omin, omax, on, _ = orc.fullgrid_geometry([0.0] * n, [10.0] * n, 1.0)
    assert cl.n_cells == tuple(int(v) for v in on)


@pytest.mark.parametrize("user_dt,r_dt", [(np.float32, np.float32), (np.float32, np.float64),
                                          (np.float64, np.float32), (np.float64, np.float64)])
def test_fullgrid_geometry_types(orc, user_dt, r_dt):
    lo = np.array([0.013, -2.7, 5.1], dtype=user_dt)
    hi = np.array([1.37, 3.3, 9.9], dtype=user_dt)
    r = r_dt(0.024)
    cl = pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=r)
    omin, omax, on, cdt = orc.fullgrid_geometry(lo.astype(np.float64), hi.astype(np.float64), float(r),
                                                user_dtype=user_dt, r_dtype=r_dt)
    assert cl.corner_dtype == np.dtype(cdt)
    assert np.array_equal(cl.min_corner.astype(np.float64), omin)
    assert np.array_equal(cl.max_corner.astype(np.float64), omax)
    assert cl.n_cells == tuple(int(v) for v in on)


def test_fullgrid_template():
    cl = pn.FullGridCellList(min_corner=(0.0, 0.0), max_corner=(1.0, 1.0))
    assert cl.is_template
    c2 = cl.copy_cell_list(0.5)
    assert not c2.is_template and c2.ndims == 2


def test_fullgrid_constructor_errors(golden):
    ce = golden["constructor_errors"]
    with pytest.raises(pn.ArgumentError, match=re.escape(ce["too_many_dims"])):
        pn.FullGridCellList(min_corner=[0.0] * 101, max_corner=[1.0] * 101)
    with pytest.raises(pn.ArgumentError, match=re.escape(ce["corner_length"])):
        pn.FullGridCellList(min_corner=[0.0] * 3, max_corner=[1.0] * 2)


def test_grid_constructor_errors(golden):
    ce = golden["constructor_errors"]
    cl3 = pn.FullGridCellList(min_corner=[0.0] * 3, max_corner=[1.0] * 3, search_radius=0.1)
    with pytest.raises(pn.ArgumentError, match=re.escape(ce["dims_mismatch"])):
        pn.GridNeighborhoodSearch[2](cell_list=cl3)
    cl2 = pn.FullGridCellList(min_corner=[0.0] * 2, max_corner=[1.0] * 2, search_radius=0.1)
    with pytest.raises(pn.ArgumentError, match=ce["invalid_strategy"]):
        pn.GridNeighborhoodSearch[2](cell_list=cl2, update_strategy="test")
    with pytest.raises(pn.ArgumentError, match=re.escape(ce["integer_radius"])):
        pn.GridNeighborhoodSearch[2](search_radius=1, cell_list=cl2)
    with pytest.raises(pn.ArgumentError, match="FullGridCellList"):
        pn.GridNeighborhoodSearch[2](search_radius=0.1)


def test_default_strategy_and_copy(golden):
    cl = pn.FullGridCellList(min_corner=(0.0, 0.0), max_corner=(1.0, 1.0))
    nhs = pn.GridNeighborhoodSearch[2](cell_list=cl)
    assert nhs.update_strategy == pn.ParallelUpdate()
    cp = pn.copy_neighborhood_search(nhs, 1.0, 10)
    assert cp.update_strategy == pn.ParallelUpdate()
    assert cp.search_radius == 1.0 and cp.ndims == 2
    cl101 = pn.FullGridCellList(min_corner=(0.0, 0.0), max_corner=(1.0, 1.0), max_points_per_cell=101)
    nhs = pn.GridNeighborhoodSearch[2](cell_list=cl101, update_strategy=pn.SerialUpdate())
    cp = pn.copy_neighborhood_search(nhs, 1.0, 10)
    assert cp.update_strategy == pn.SerialUpdate()
    assert cp.cell_list.max_points_per_cell == 101
    assert nhs.requires_update() == (False, True)


def test_precomputed_defaults(golden):
    md = golden["precomputed_defaults"]["max_neighbors"]
    for d in (1, 2, 3):
        assert pn.PrecomputedNeighborhoodSearch[d]().max_neighbors == md[str(d)]
    p = pn.PrecomputedNeighborhoodSearch[3]()
    assert p.requires_update() == (True, True)
    with pytest.raises(pn.ArgumentError):
        pn.PrecomputedNeighborhoodSearch[2](search_radius=2)


def test_host_cell_coords_saturation(golden):
    fg = golden["floor_to_int"]["full_grid_cells"]
    cl = pn.FullGridCellList(min_corner=fg["min_corner"], max_corner=fg["max_corner"],
                             search_radius=fg["search_radius"])
    nhs = pn.GridNeighborhoodSearch[2](search_radius=1.0, cell_list=cl)
    tmax, tmin = 2**63 - 1, -(2**63)
    assert nhs.host_cell_coords([math.inf, -math.inf]) == (tmin, tmin + 1)
    assert nhs.host_cell_coords([math.nan, 0.0]) == (tmin, 2)
    assert nhs.host_cell_coords([float(tmax) + 1.0, -float(tmax) - 1.0]) == (tmin, tmin + 1)


def test_functor_menu_rejects_closures():
    cl = pn.FullGridCellList(min_corner=(0.0, 0.0), max_corner=(1.0, 1.0), search_radius=0.1)
    nhs = pn.GridNeighborhoodSearch[2](search_radius=0.1, cell_list=cl)
    import torch
    x = torch.zeros(2, 3)
    with pytest.raises(pn.ArgumentError, match="closures"):
        pn.foreach_point_neighbor(lambda i, j, p, d: None, x, x, nhs)


# ---- synthetic inputs ----------------------------------------------------------------------------

def test_splitmix_reference_values():
    # SplitMix64 with seed 0: first outputs of Vigna's reference implementation
    z = synthetic.splitmix64(0, np.arange(3, dtype=np.uint64))
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_jittered_lattice_layout():
    pts = synthetic.jittered_lattice((3, 2), 1.0, 0.0, dtype=np.float64)
    assert pts.tolist() == [[1, 1], [2, 1], [3, 1], [1, 2], [2, 2], [3, 2]]
    a = synthetic.jittered_lattice((20, 20, 20), 0.01, 0.001, dtype=np.float32, seed=4)
    b = synthetic.jittered_lattice((20, 20, 20), 0.01, 0.001, dtype=np.float32, seed=4)
    assert np.array_equal(a, b)
    base = synthetic.jittered_lattice((20, 20, 20), 0.01, 0.0, dtype=np.float64)
    assert np.abs(a.astype(np.float64) - base).max() <= 0.001 + 1e-7


# ---- the C-ABI library: loads and exports every declared symbol --------------------------------

def header_symbols():
    with open(os.path.join(ROOT, "include", "pointneighbors_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pn_\w+)\s*\(", text, re.M)))


def test_library_exports_header_symbols():
    lib = pn.load()
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
        assert s in SIGNATURES, f"{s} has no ctypes signature"
    assert lib.pn_abi_version() == 1


def test_library_argument_errors_without_gpu():
    lib = pn.load()
    h = ctypes.c_void_p()
    mins = (ctypes.c_double * 2)(0.0, 0.0)
    ncs = (ctypes.c_int64 * 2)(2, 10)
    st = lib.pn_grid_create(2, 0, 0, 0, mins, ncs, 0.1, 100, ctypes.byref(h))
    assert st == 1 and "at least 3 cells" in lib.pn_last_error().decode()
    ncs = (ctypes.c_int64 * 2)(10, 10)
    st = lib.pn_grid_create(2, 0, 1, 0, mins, ncs, 0.1, 100, ctypes.byref(h))
    assert st == 1 and "promotion" in lib.pn_last_error().decode()
    st = lib.pn_grid_create(2, 0, 0, 0, mins, ncs, 0.1, 100, ctypes.byref(h))
    assert st == 0 and h.value
    assert lib.pn_count_neighbors(h, 0, None, 0) == 1
    assert lib.pn_count_neighbors(h, 0, ctypes.c_void_p(8), 0) == 5  # not initialized
    assert lib.pn_grid_destroy(h) == 0
    # the stencil's candidate count must fit in int32
    st = lib.pn_grid_create(2, 0, 0, 0, mins, ncs, 0.1, 2**31 // 27 + 1, ctypes.byref(h))
    assert st == 1 and "max_points_per_cell" in lib.pn_last_error().decode()
    # SpatialHashingCellList handles: argument checks before any device work
    st = lib.pn_grid_create_spatial_hash(3, 0, 0, 0.1, 0, 100, ctypes.byref(h))
    assert st == 1 and "list_size" in lib.pn_last_error().decode()
    st = lib.pn_grid_create_spatial_hash(3, 0, 0, -1.0, 10, 100, ctypes.byref(h))
    assert st == 1 and "search_radius" in lib.pn_last_error().decode()
    st = lib.pn_grid_create_spatial_hash(3, 0, 0, 0.1, 10, 2**31 // 27 + 1, ctypes.byref(h))
    assert st == 1 and "max_points_per_cell" in lib.pn_last_error().decode()


def test_traversal_lds_padding_fits():
    """k_traverse holds CAP = CH + 64 staged candidates; a chunk takes at most
    cmax = CAP - 2S + 1 of them and is padded to S * L, L = ceil(mc / S) | 1.  Check every
    (queries per cell, chunk size) fits."""
    for ch in (512, 256, 128):
        cap = ch + 64
        for nq in range(1, 65):
            s = 64 // nq
            cmax = cap - 2 * s + 1
            assert cmax >= 1
            for mc in range(1, cmax + 1):
                L = (mc + s - 1) // s | 1
                assert s * L <= cap


# ---- periodic boxes (src/nhs_grid.jl:102-125, src/neighborhood_search.jl:129-141) ----------

def _periodic_nhs(lo, hi, r, dtype=np.float64):
    lo = np.asarray(lo, dtype=dtype)
    hi = np.asarray(hi, dtype=dtype)
    pb = pn.PeriodicBox(min_corner=lo, max_corner=hi)
    cl = pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=dtype(r))
    return pn.GridNeighborhoodSearch(len(lo), search_radius=dtype(r), periodic_box=pb, cell_list=cl)


def test_periodic_sizing_matches_oracle(orc, golden):
    r = golden["periodic"]["search_radius"]
    for case in golden["periodic"]["cases"]:
        nhs = _periodic_nhs(case["min_corner"], case["max_corner"], r)
        pb = nhs.periodic_box
        n, cs = orc.periodic_sizing(pb.size, r)
        assert list(nhs.n_cells) == n.tolist()
        assert [float(v) for v in nhs.cell_size] == cs.tolist()
    nhs = _periodic_nhs([0, 0, 0], [1.0, 1.3, 0.9], 0.29, dtype=np.float32)
    n, cs = orc.periodic_sizing(nhs.periodic_box.size.astype(np.float64), float(np.float32(0.29)),
                                r_dtype=np.float32)
    assert list(nhs.n_cells) == n.tolist()
    assert [float(v) for v in nhs.cell_size] == cs.tolist()
    assert all(np.asarray(nhs.cell_size).dtype == np.float32 for _ in [0])


def test_periodic_constructor_errors(golden):
    E = golden["periodic"]["errors"]
    pb = pn.PeriodicBox(min_corner=[0.0, 0.0], max_corner=[1.0, 1.0])
    cl = pn.FullGridCellList(min_corner=[0.0, 0.0], max_corner=[1.0, 1.0], search_radius=np.float32(0.1))
    with pytest.raises(pn.ArgumentError, match="must have the same element type"):
        pn.GridNeighborhoodSearch(2, search_radius=np.float32(0.1), periodic_box=pb, cell_list=cl)
    assert E["type_mismatch"].startswith("the `search_radius` and the `PeriodicBox`")
    with pytest.raises(pn.ArgumentError, match="needs at least 3 cells in each dimension"):
        _periodic_nhs([0.0, 0.0], [0.25, 1.0], 0.1)
    # search_radius < eps(): no periodicity (src/nhs_grid.jl:102-105)
    nhs = pn.GridNeighborhoodSearch(2, search_radius=0.0, periodic_box=pb,
                                    cell_list=pn.FullGridCellList(min_corner=[0.0, 0.0], max_corner=[1.0, 1.0]))
    assert nhs.n_cells == (-1, -1)


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_periodic_host_cell_coords_match_oracle(orc, golden, T):
    R = golden["periodic"]["rounding"]
    nhs = _periodic_nhs(R["min_corner"], R["max_corner"], R["search_radius"], dtype=T)
    cl = nhs.cell_list
    g = orc.Grid(2, cl.min_corner.astype(np.float64), cl.n_cells, float(T(R["search_radius"])),
                 coord_dtype=T, radius_dtype=T, cell_dtype=cl.corner_dtype)
    g.set_periodic(np.asarray(nhs.cell_size, dtype=np.float64), nhs.n_cells,
                   nhs.periodic_box.size.astype(np.float64))
    vals = [np.nextafter(T(0), T(-1)), T(0), np.nextafter(T(0), T(1)), np.nextafter(T(1), T(0)), T(1),
            np.nextafter(T(1), T(2)), T(-3.7), T(5.2), T(np.nan)]
    for a in vals:
        x = np.asarray([a, T(0.5)], dtype=T)
        assert nhs.host_cell_coords(x) == g.cell_coords(x), a
