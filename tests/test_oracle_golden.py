"""Pins the oracle (oracle/pn_oracle.cpp) against the reference's known answers
(tests/golden/reference_known_answers.json, transcribed from the reference's test files),
closed forms, and independent brute force.  CPU only."""

import itertools
import math

import numpy as np
import pytest

from pointneighbors import synthetic

TYPEMAX = 2**63 - 1
TYPEMIN = -(2**63)


def wrap(v):
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >= (1 << 63) else v


def parse(v):
    return {"Inf": math.inf, "-Inf": -math.inf, "NaN": math.nan,
            "typemax+1": float(TYPEMAX) + 1.0, "-typemax-1": -float(TYPEMAX) - 1.0}.get(v, v)


def parse_cell(v):
    return {"typemax": TYPEMAX, "typemin": TYPEMIN}.get(v, v)


# ---- floor_to_int and cell coordinates (test/nhs_grid.jl:62-86) --------------------------

def test_floor_to_int_dictionary_cells(orc, golden):
    from pointneighbors.util import floor_to_int
    for case in golden["floor_to_int"]["dictionary_cells"]:
        coords = [parse(c) for c in case["coords"]]
        want = tuple(parse_cell(c) for c in case["cell"])
        got_orc = tuple(orc.floor_to_int(c / 1.0) for c in coords)
        got_host = tuple(floor_to_int(c / 1.0) for c in coords)
        assert got_orc == want
        assert got_host == want


def test_floor_to_int_float32(orc):
    assert orc.floor_to_int(np.float32(2.0**63), np.float32) == TYPEMAX
    assert orc.floor_to_int(np.float32(-(2.0**63)), np.float32) == TYPEMIN
    assert orc.floor_to_int(np.float32(-1.5), np.float32) == -2
    assert orc.floor_to_int(np.float32(np.nan), np.float32) == TYPEMAX


def test_full_grid_cell_coords_wrap(orc, golden):
    fg = golden["floor_to_int"]["full_grid_cells"]
    pmin, pmax, ncells, _ = orc.fullgrid_geometry(fg["min_corner"], fg["max_corner"], fg["search_radius"])
    for case in golden["floor_to_int"]["dictionary_cells"]:
        coords = [parse(c) for c in case["coords"]]
        want = tuple(wrap(parse_cell(c) + 1) for c in case["cell"])
        if case["coords"][0] == "NaN":
            # test/nhs_grid.jl:84: (typemax(Int), 1) .+ 1 -- the 0.0 coordinate lies in
            # padded cell 1 because the padded min corner is -1.001
            want = (wrap(TYPEMAX + 1), 2)
        assert orc.cell_coords(coords, pmin, 1.0) == want


# ---- FullGridCellList geometry (test/cell_lists/full_grid.jl:20-29) ------------------------

@pytest.mark.parametrize("n", [1, 2, 3])
def test_padded_corners(orc, golden, n):
    pc = golden["padded_corners"]
    pmin, pmax, ncells, _ = orc.fullgrid_geometry([pc["min_corner"]] * n, [pc["max_corner"]] * n,
                                                  pc["search_radius"])
    assert list(pmin) == [pc["padded_min"]] * n
    assert list(pmax) == [10.0 + 1.001] * n
    assert list(ncells) == [math.ceil((10.0 + 1.001 - -1.001) / 1.0)] * n


def test_template_geometry(orc):
    _, _, ncells, _ = orc.fullgrid_geometry([0.0], [1.0], 0.0)
    assert ncells is None


# ---- bounds and overflow errors ---------------------------------------------------------------

def _grid_for_box(orc, n, lo, hi, r, cap=100, dtype=np.float64):
    pmin, pmax, ncells, cdt = orc.fullgrid_geometry([lo] * n, [hi] * n, r)
    return orc.Grid(n, pmin, ncells, r, coord_dtype=dtype, radius_dtype=np.float64, cell_dtype=np.float64,
                    max_points_per_cell=cap)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_bounds_errors(orc, golden, n):
    be = golden["bounds_errors"]
    g = _grid_for_box(orc, n, be["min_corner"], be["max_corner"], be["search_radius"])
    y = synthetic.uniform(11 + n, n * be["n_points"]).reshape(be["n_points"], n)
    k = be["index"] - 1
    values = {"NaN": math.nan, "min-0.01": be["min_corner"] - 0.01, "max+1.01": be["max_corner"] + 1.01}
    for key in be["initialize_errors"] + be["update_errors"]:
        y[k, 0] = values[key]
        st, bp, _ = g.build(y)
        assert st == orc.ERR_OUT_OF_BOUNDS and bp == k
    for v in be["initialize_ok"] + be["update_ok"]:
        y[k, 0] = v
        st, _, _ = g.build(y)
        assert st == orc.OK


def test_overflow(orc):
    g = _grid_for_box(orc, 2, 0.0, 10.0, 1.0, cap=100)
    y = np.full((100, 2), 5.5)
    assert g.build(y)[0] == orc.OK
    y = np.full((101, 2), 5.5)
    assert g.build(y)[0] == orc.ERR_CELL_FULL


# ---- candidate lists (test/nhs_grid.jl:88-226) ---------------------------------------------

def _rect(ndims, golden_key, golden):
    rng = golden[golden_key]["range"]
    vals = np.array([-0.25, -0.15, -0.05, 0.05, 0.15, 0.25, 0.35])
    assert vals[0] == rng[0] and vals[-1] == rng[2]
    # Iterators.product: first coordinate fastest
    pts = np.array([p[::-1] for p in itertools.product(*([vals] * ndims))])
    return pts


def test_rectangular_2d_dictionary(orc, golden):
    gk = golden["rectangular_2d"]
    c1 = _rect(2, "rectangular_2d", golden)
    p1 = np.array(gk["point_position1"])
    r = gk["search_radius"]
    one = lambda l: [v + 1 for v in l]
    assert one(orc.candidates_dict(c1, p1, r)) == gk["neighbors1"]
    c2 = c1 + np.array(gk["shift1"])
    assert one(orc.candidates_dict(c2, p1, r)) == gk["neighbors2"]
    assert one(orc.candidates_dict(c2, p1 + np.array(gk["shift1"]), r)) == gk["neighbors3"]
    assert one(orc.candidates_dict(c1, p1, 2 * r)) == gk["neighbors4"]
    c3 = c1 + np.array(gk["shift2"])
    assert one(orc.candidates_dict(c3, p1, 2 * r)) == gk["neighbors5"]


def test_rectangular_3d_dictionary(orc, golden):
    gk = golden["rectangular_3d"]
    c1 = _rect(3, "rectangular_3d", golden)
    p1 = np.array(gk["point_position1"])
    r = gk["search_radius"]
    one = lambda l: [v + 1 for v in l]
    assert one(orc.candidates_dict(c1, p1, r)) == gk["neighbors1"]
    c2 = c1 + np.array(gk["shift1"])
    assert one(orc.candidates_dict(c2, p1, r)) == gk["neighbors2"]
    assert one(orc.candidates_dict(c2, p1 + np.array(gk["shift1"]), r)) == gk["neighbors3"]


def oracle_eachneighbor(orc, g, pmin, r, q):
    cell = orc.cell_coords(q, pmin, r)
    out = []
    for off in itertools.product(*([(-1, 0, 1)] * len(cell))):
        c = [cell[d] + off[::-1][d] for d in range(len(cell))]
        lin, stride = 0, 1
        for d in range(len(cell)):
            lin += (c[d] - 1) * stride
            stride *= g.n_cells[d]
        out.extend(g.cell(lin))
    return sorted(out)


def test_full_grid_eachindex_y(orc, golden):
    gk = golden["rectangular_3d"]
    fk = gk["full_grid_eachindex_y"]
    c1 = _rect(3, "rectangular_3d", golden)
    c2 = c1 + np.array(gk["shift1"])
    r = gk["search_radius"]
    lo = np.minimum(c1.min(axis=0), c2.min(axis=0))
    hi = np.maximum(c1.max(axis=0), c2.max(axis=0))
    pmin, pmax, ncells, _ = orc.fullgrid_geometry(lo, hi, r)
    g = orc.Grid(3, pmin, ncells, r)
    assert g.build(c1)[0] == orc.OK
    first, last = fk["eachindex_y"]
    assert g.build(c2, active=np.arange(first - 1, last))[0] == orc.OK
    p1 = np.array(gk["point_position1"])
    p2 = p1 + np.array(gk["shift1"])
    assert [v + 1 for v in oracle_eachneighbor(orc, g, pmin, r, p1)] == fk["neighbors2"]
    assert [v + 1 for v in oracle_eachneighbor(orc, g, pmin, r, p2)] == fk["neighbors3"]


# ---- known answers of the GPU tutorial and closed-form lattice counts -----------------------

def lattice_int(shape, dtype):
    return synthetic.jittered_lattice(shape, 1.0, 0.0, dtype=dtype)


def grid_for_points(orc, pts, r, coord_dtype, radius_dtype=None, corner_from_points=True, cap=100):
    radius_dtype = radius_dtype or coord_dtype
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    pmin, pmax, ncells, cdt = orc.fullgrid_geometry(lo, hi, r, user_dtype=pts.dtype, r_dtype=radius_dtype)
    cell_dt = np.float64 if (np.dtype(coord_dtype) == np.float64 or cdt == np.float64) else np.float32
    return orc.Grid(pts.shape[1], pmin, ncells, r, coord_dtype=coord_dtype, radius_dtype=radius_dtype,
                    cell_dtype=cell_dt, max_points_per_cell=cap), pmin


def test_gpu_tutorial_extrema(orc, golden):
    t = golden["gpu_tutorial"]
    pts = lattice_int(t["n_points_per_dimension"], np.float32)
    g, _ = grid_for_points(orc, pts, t["search_radius"], np.float32)
    assert g.build(pts)[0] == orc.OK
    st, _, counts, _, _ = g.traverse(pts, pts)
    assert st == orc.OK
    assert (int(counts.min()), int(counts.max())) == tuple(t["extrema"])


def count_ball(ndims, r, octant):
    """Integer points with |p|^2 <= r^2 (all or one closed octant): exact enumeration."""
    R = int(r)
    rng = range(0, R + 1) if octant else range(-R, R + 1)
    return sum(1 for p in itertools.product(rng, repeat=ndims) if sum(v * v for v in p) <= r * r)


@pytest.mark.parametrize("ndims,interior,corner", [(2, 29, 11), (3, 123, 29)])
def test_closed_form_lattice_counts(orc, ndims, interior, corner):
    assert count_ball(ndims, 3, False) == interior
    assert count_ball(ndims, 3, True) == corner
    shape = (12,) * ndims
    pts = lattice_int(shape, np.float64)
    g, _ = grid_for_points(orc, pts, 3.0, np.float64)
    assert g.build(pts)[0] == orc.OK
    st, _, counts, _, _ = g.traverse(pts, pts)
    assert (int(counts.min()), int(counts.max())) == (corner, interior)


# ---- grid traversal against brute force (test/neighborhood_search.jl:186-343) ---------------

def numpy_bruteforce(x, y, r):
    """Independent brute force: per-row float64 step-rounded d^2 (separate ufuncs, no FMA)."""
    rows = []
    r2 = r * r
    for i in range(0, x.shape[0], 512):
        xi = x[i:i + 512]
        pd = xi[:, None, :] - y[None, :, :]
        d2 = pd[..., 0] * pd[..., 0]
        for d in range(1, x.shape[1]):
            d2 = d2 + pd[..., d] * pd[..., d]
        for row in d2 <= r2:
            rows.append(np.nonzero(row)[0])
    return rows


def cloud(shape, seed):
    return synthetic.jittered_lattice(shape, 1.0, 0.25, dtype=np.float64, seed=seed)


@pytest.mark.parametrize("shape", [(10, 11), (100, 90), (9, 10, 7), (39, 40, 41)])
@pytest.mark.parametrize("seed", [1, 2])
def test_grid_vs_bruteforce(orc, golden, shape, seed):
    r = golden["trivial_comparison"]["search_radius"]
    coords = cloud(shape, seed)
    init = cloud(shape, 1)
    lo, hi = coords.min(axis=0) - r, coords.max(axis=0) + r
    pmin, _, ncells, _ = orc.fullgrid_geometry(lo, hi, r)
    g = orc.Grid(len(shape), pmin, ncells, r)
    st, _, _ = g.build(init)
    if seed != 1:
        # initialize! on the seed-1 cloud, then update! with this cloud
        st, _, _ = g.build(coords)
    assert st == orc.OK
    n = coords.shape[0]
    sample = np.arange(n) if n <= 10000 else np.sort(
        (synthetic.uniform(seed, 2000) * n).astype(np.int64))
    st, _, counts, sums, (rowptr, cols) = g.traverse(coords, coords, points=sample, lists=True)
    assert st == orc.OK
    srt = orc.sorted_lists(rowptr, cols)
    expected = numpy_bruteforce(coords[sample], coords, r)
    for a, e in enumerate(expected):
        assert np.array_equal(srt[rowptr[a]:rowptr[a + 1]], e)
        assert sums[a] == int(e.sum())
    tcounts, (trp, tcols) = orc.trivial(coords[sample], coords, r, lists=True)
    assert np.array_equal(tcounts, counts)
    assert np.array_equal(tcols, srt)


# ---- d^2 evaluation order: unfused, left to right ----------------------------------------------

def adversarial_pairs_f32(n_try=400000, seed=5):
    """Pairs whose acceptance differs between unfused ((a*a + b*b) + c*c) and an FMA form."""
    u = synthetic.uniform(seed, 3 * n_try).reshape(n_try, 3)
    r = np.float32(1.2) * np.float32(0.01) * np.float32(2.0)
    r2 = np.float32(r * r)
    direction = u - 0.5
    direction /= np.linalg.norm(direction, axis=1, keepdims=True)
    xi = np.float32(0.5)
    xj = (np.float64(xi) + direction * float(r) * (1.0 + 1e-6 * (u[:, :1] - 0.5))).astype(np.float32)
    pd = (np.float32(xi) - xj).astype(np.float32)
    p = pd * pd
    unf = (p[:, 0] + p[:, 1]) + p[:, 2]
    fused = (pd[:, 2].astype(np.float64) * pd[:, 2] + (p[:, 0] + p[:, 1]).astype(np.float64)).astype(np.float32)
    sel = (unf <= r2) != (fused <= r2)
    return xj[sel], r, unf[sel] <= r2


def test_d2_unfused_left_fold(orc):
    xj, r, want = adversarial_pairs_f32()
    assert len(xj) >= 5, "the search must find FMA-sensitive pairs"
    xi = np.full((1, 3), 0.5, dtype=np.float32)
    for k in range(min(len(xj), 50)):
        pts = np.vstack([xi, xj[k:k + 1]])
        lo = pts.min(axis=0) - r
        hi = pts.max(axis=0) + r
        pmin, _, ncells, _ = orc.fullgrid_geometry(lo, hi, r, user_dtype=np.float32, r_dtype=np.float32)
        g = orc.Grid(3, pmin, ncells, float(r), coord_dtype=np.float32, radius_dtype=np.float32,
                     cell_dtype=np.float32)
        assert g.build(pts)[0] == orc.OK
        st, _, counts, _, _ = g.traverse(pts[:1], pts)
        assert st == orc.OK
        assert counts[0] == 1 + int(want[k])


# ---- precomputed lists -----------------------------------------------------------------------

def test_precomputed_sorted_lists(orc):
    coords = cloud((30, 31), 3)
    r = 2.5
    lo, hi = coords.min(axis=0) - r, coords.max(axis=0) + r
    pmin, _, ncells, _ = orc.fullgrid_geometry(lo, hi, r)
    g = orc.Grid(2, pmin, ncells, r)
    assert g.build(coords)[0] == orc.OK
    st, _, (rowptr, cols) = g.precomputed(coords, coords, 64)
    assert st == orc.OK
    expected = numpy_bruteforce(coords, coords, r)
    for i, e in enumerate(expected):
        assert np.array_equal(cols[rowptr[i]:rowptr[i + 1]], e)
    st, _, _ = g.precomputed(coords, coords, 10)
    assert st == orc.ERR_CELL_FULL


def test_query_bounds(orc):
    coords = cloud((10, 10), 1)
    r = 2.5
    pmin, _, ncells, _ = orc.fullgrid_geometry(coords.min(axis=0), coords.max(axis=0), r)
    g = orc.Grid(2, pmin, ncells, r)
    assert g.build(coords)[0] == orc.OK
    far = coords[:1] + 100.0
    st, bp, _, _, _ = g.traverse(far, coords)
    assert st == orc.ERR_QUERY_BOUNDS and bp == 0
