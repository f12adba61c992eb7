"""Pins the oracle's periodic-box path (GridNeighborhoodSearch with a PeriodicBox) against
the reference's known answers (tests/golden/reference_known_answers.json, "periodic"),
closed-form grid sizing, and an image-enumeration brute force.  CPU only."""

import itertools

import numpy as np
import pytest


def periodic_grid(orc, lo, hi, r, dtype=np.float64, cap=100):
    """FullGridCellList(min_corner = box min, max_corner = box max, search_radius = r) with
    the box as PeriodicBox, as in test/neighborhood_search.jl:110-118."""
    lo = np.asarray(lo, dtype=dtype)
    hi = np.asarray(hi, dtype=dtype)
    size = (hi - lo).astype(dtype)
    n, cs = orc.periodic_sizing(size.astype(np.float64), float(dtype(r)), r_dtype=dtype)
    pmin, _, ncells, cdt = orc.fullgrid_geometry(lo.astype(np.float64), hi.astype(np.float64), float(dtype(r)),
                                                 user_dtype=dtype, r_dtype=dtype)
    g = orc.Grid(len(lo), pmin, ncells, float(dtype(r)), coord_dtype=dtype, radius_dtype=dtype, cell_dtype=cdt,
                 max_points_per_cell=cap)
    g.set_periodic(cs, n, size.astype(np.float64))
    return g, n, size


def candidates(g, n, pts, i):
    """Points in the 3^D wrapped neighbouring cells of point i (eachneighbor)."""
    cell = g.cell_coords(pts[i])
    nd = len(cell)
    out = set()
    for off in itertools.product((-1, 0, 1), repeat=nd):
        nc = [((cell[d] + off[d] - 2) % n[d]) + 2 for d in range(nd)]
        lin = 0
        stride = 1
        for d in range(nd):
            lin += (nc[d] - 1) * stride
            stride *= g.n_cells[d]
        out.update(g.cell(lin))
    return sorted(out)


def test_periodic_sizing_closed_form(orc):
    # 0.3 / 0.1 == 2.9999999999999996 in Float64: the 10eps() tolerance gives 3 cells
    n, cs = orc.periodic_sizing([0.3, 0.6], 0.1)
    assert n.tolist() == [3, 6]
    assert cs.tolist() == [0.3 / 3, 0.6 / 6]
    n, cs = orc.periodic_sizing([0.305, 0.63], 0.1)
    assert n.tolist() == [3, 6]
    assert cs.tolist() == [0.305 / 3, 0.63 / 6]
    n32, cs32 = orc.periodic_sizing([np.float32(0.3)], np.float32(0.1), r_dtype=np.float32)
    assert cs32[0] == float(np.float32(np.float32(0.3) / np.float32(n32[0])))


def test_periodic_neighbors_reference_cases(orc, golden):
    P = golden["periodic"]
    r = P["search_radius"]
    want = [[v - 1 for v in l] for l in P["neighbors"]]
    for case in P["cases"]:
        pts = np.asarray(case["coords"])
        g, n, size = periodic_grid(orc, case["min_corner"], case["max_corner"], r)
        assert g.build(pts)[0] == orc.OK
        st, _, counts, _, (rp, cols) = g.traverse(pts, pts, lists=True)
        assert st == orc.OK
        got = [sorted(cols[rp[i]:rp[i + 1]].tolist()) for i in range(len(pts))]
        assert got == want, case["name"]
        _, triv = orc.trivial_periodic(pts, pts, r, size, lists=True)
        assert triv == want, case["name"]


def test_periodic_candidates_reference_cases(orc, golden):
    P = golden["periodic"]
    r = P["search_radius"]
    want = [[v - 1 for v in l] for l in P["candidates"]]
    for case in P["cases"]:
        off = np.asarray(case.get("candidates_offset", [0.0] * len(case["min_corner"])))
        pts = np.asarray(case["coords"]) - off
        lo = np.asarray(case["min_corner"]) - off
        hi = np.asarray(case["max_corner"]) - off
        g, n, _ = periodic_grid(orc, lo, hi, r)
        assert g.build(pts)[0] == orc.OK
        got = [candidates(g, n, pts, i) for i in range(len(pts))]
        assert got == want, case["name"]


def test_periodic_split_cells(orc, golden):
    S = golden["periodic"]["split_cells"]
    pts = np.asarray(S["coords"])
    g, n, _ = periodic_grid(orc, S["min_corner"], S["max_corner"], S["search_radius"])
    assert n.tolist() == [4, 3]
    assert g.build(pts)[0] == orc.OK
    want = [[v - 1 for v in l] for l in S["candidates"]]
    assert [candidates(g, n, pts, i) for i in range(len(pts))] == want


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_periodic_rounding_at_box_faces(orc, golden, T):
    """test/neighborhood_search.jl:4-58: cells of points at the box faces equal the cells of
    their periodic_coords, and lie in 2:n_cells+1."""
    R = golden["periodic"]["rounding"]
    lo = np.asarray(R["min_corner"], dtype=T)
    hi = np.asarray(R["max_corner"], dtype=T)
    size = hi - lo
    g, n, _ = periodic_grid(orc, lo, hi, R["search_radius"], dtype=T)
    zero, one, half = T(0), T(1), T(0.5)
    pts = []
    for v in (zero, one):
        for w in (np.nextafter(v, T(-1)), v, np.nextafter(v, T(2))):
            pts.append((w, half))
            pts.append((half, w))
    for x in pts:
        x = np.asarray(x, dtype=T)
        # periodic_coords (src/neighborhood_search.jl:439-449), in T
        off = np.floor((x - lo) / size)
        xp = (x - off * size).astype(T)
        xp = np.minimum(np.maximum(xp, lo), hi)
        c = g.cell_coords(x)
        assert c == g.cell_coords(xp)
        assert all(2 <= c[d] <= n[d] + 1 for d in range(2))


def image_brute_force(x, y, size):
    """Minimum-image distances by enumerating periodic images |x - y + k L|, k in
    {-1, 0, 1}^D, in float64 (an independent formulation of the minimum image)."""
    nd = x.shape[1]
    d2 = np.full((len(x), len(y)), np.inf)
    for k in itertools.product((-1, 0, 1), repeat=nd):
        diff = x[:, None, :] - y[None, :, :] + np.asarray(k) * size
        d2 = np.minimum(d2, (diff ** 2).sum(-1))
    return np.sqrt(d2)


@pytest.mark.parametrize("nd,T,box", [
    (2, np.float64, ([0.0, 0.0], [3.0, 2.05])),
    (3, np.float64, ([-1.0, 0.0, 0.5], [1.5, 1.2, 3.0])),
    (3, np.float32, ([0.0, 0.0, 0.0], [1.0, 1.3, 0.9])),
    (2, np.float32, ([-0.5, -0.5], [0.4, 0.7])),   # 3 cells in x: the smallest periodic grid
])
def test_periodic_grid_equals_image_brute_force(orc, nd, T, box):
    rng = np.random.default_rng(11 + nd)
    lo = np.asarray(box[0], dtype=T)
    hi = np.asarray(box[1], dtype=T)
    size = (hi - lo).astype(T)
    r = T(0.29)
    n_pts = 400
    # points in the box and a few just outside it (cells wrap, coordinates do not)
    u = rng.uniform(-0.02, 1.02, size=(n_pts, nd))
    pts = (lo + u * size).astype(T)
    g, n, _ = periodic_grid(orc, lo, hi, r, dtype=T)
    assert min(n) >= 3
    assert g.build(pts)[0] == orc.OK
    st, _, counts, _, (rp, cols) = g.traverse(pts, pts, lists=True)
    assert st == orc.OK
    got = [set(cols[rp[i]:rp[i + 1]].tolist()) for i in range(n_pts)]
    _, triv = orc.trivial_periodic(pts, pts, r, size.astype(np.float64), coord_dtype=T, radius_dtype=T, lists=True)
    assert [sorted(s) for s in got] == triv
    d = image_brute_force(pts.astype(np.float64), pts.astype(np.float64), size.astype(np.float64))
    # pairs within rounding of the boundary may go either way in T
    tol = 1e-9 if T == np.float64 else 1e-5
    for i in range(n_pts):
        want = set(np.nonzero(d[i] <= float(r))[0].tolist())
        amb = set(np.nonzero(np.abs(d[i] - float(r)) < tol * float(r))[0].tolist())
        assert got[i] - amb == want - amb, i
