"""Parity of the HIP path (through the C-ABI, via the host mirror) with the oracle on the
same seeded inputs: neighbour counts, index sums and sorted neighbour lists bit-exact."""

import math
import re

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import synthetic

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(pts):
    return synthetic.to_device_coords(pts)


def make_nhs(pts, r, lo=None, hi=None, cap=100, strategy=None, corner_dtype=None):
    lo = pts.min(axis=0) if lo is None else lo
    hi = pts.max(axis=0) if hi is None else hi
    if corner_dtype is not None:
        lo, hi = np.asarray(lo, dtype=corner_dtype), np.asarray(hi, dtype=corner_dtype)
    cl = pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=r, max_points_per_cell=cap)
    return pn.GridNeighborhoodSearch(pts.shape[1], search_radius=r, n_points=pts.shape[0], cell_list=cl,
                                     update_strategy=strategy)


def oracle_like(orc, nhs, coord_dtype):
    cl = nhs.cell_list
    cell_dt = np.float64 if (np.dtype(coord_dtype) == np.float64 or cl.corner_dtype == np.float64) else np.float32
    return orc.Grid(nhs.ndims, cl.min_corner.astype(np.float64), cl.n_cells, float(nhs.search_radius),
                    coord_dtype=coord_dtype, radius_dtype=nhs.radius_dtype, cell_dtype=cell_dt,
                    max_points_per_cell=cl.max_points_per_cell)


def gpu_counts(nhs, x, y, points=None):
    n = x.shape[1]
    out = torch.zeros(n, dtype=torch.int32, device=x.device)
    pn.foreach_point_neighbor(pn.CountNeighbors(out), x, y, nhs, points=points)
    return out.cpu().numpy()


def gpu_sums(nhs, x, y, points=None):
    out = torch.zeros(x.shape[1], dtype=torch.int64, device=x.device)
    pn.foreach_point_neighbor(pn.SumNeighborIndices(out), x, y, nhs, points=points)
    return out.cpu().numpy()


def cloud(shape, seed, jitter=0.25, dtype=np.float64):
    return synthetic.jittered_lattice(shape, 1.0, jitter, dtype=dtype, seed=seed)


# ---- known answers --------------------------------------------------------------------------

def test_gpu_tutorial_extrema(orc, golden):
    t = golden["gpu_tutorial"]
    pts = synthetic.jittered_lattice(t["n_points_per_dimension"], 1.0, 0.0, dtype=np.float32)
    r = np.float32(t["search_radius"])
    nhs = make_nhs(pts, r)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    c = gpu_counts(nhs, x, x)
    assert (int(c.min()), int(c.max())) == tuple(t["extrema"])
    og = oracle_like(orc, nhs, np.float32)
    assert og.build(pts)[0] == orc.OK
    assert np.array_equal(c, og.traverse(pts, pts)[2])


def test_full_grid_eachindex_y_fixture(golden):
    gk = golden["rectangular_3d"]
    fk = gk["full_grid_eachindex_y"]
    vals = np.array([-0.25, -0.15, -0.05, 0.05, 0.15, 0.25, 0.35])
    import itertools
    c1 = np.array([p[::-1] for p in itertools.product(vals, vals, vals)])
    c2 = c1 + np.array(gk["shift1"])
    r = gk["search_radius"]
    lo = np.minimum(c1.min(axis=0), c2.min(axis=0))
    hi = np.maximum(c1.max(axis=0), c2.max(axis=0))
    for strategy in (pn.SerialUpdate(), pn.ParallelUpdate()):
        nhs = make_nhs(c1, r, lo, hi, strategy=strategy)
        d1, d2 = dev(c1), dev(c2)
        pn.initialize(nhs, d1, d1)
        first, last = fk["eachindex_y"]
        pn.update(nhs, d2, d2, eachindex_y=range(first - 1, last))
        p1 = np.array(gk["point_position1"])
        p2 = p1 + np.array(gk["shift1"])
        assert sorted(v + 1 for v in nhs.eachneighbor(p1)) == fk["neighbors2"]
        assert sorted(v + 1 for v in nhs.eachneighbor(p2)) == fk["neighbors3"]


# ---- jittered clouds against the oracle (test/neighborhood_search.jl:186-343) ----------------

def check_lists(orc, og, nhs, pts, x):
    p = pn.PrecomputedNeighborhoodSearch(pts.shape[1], search_radius=nhs.search_radius,
                                         n_points=pts.shape[0], update_neighborhood_search=nhs,
                                         max_neighbors=100000)
    pn.initialize(p, x, x)
    rp, cols = p.neighbor_lists.to_host()
    st, _, (orp, ocols) = og.precomputed(pts, pts, 100000)
    assert st == orc.OK
    assert np.array_equal(rp.astype(np.int64), orp)
    assert np.array_equal(cols, ocols)


@pytest.mark.parametrize("shape", [(10, 11), (100, 90), (9, 10, 7), (39, 40, 41)])
@pytest.mark.parametrize("seed", [1, 2])
def test_clouds_vs_oracle(orc, shape, seed):
    r = 2.5
    coords = cloud(shape, seed)
    init = cloud(shape, 1)
    lo, hi = coords.min(axis=0) - r, coords.max(axis=0) + r
    nhs = make_nhs(coords, r, lo, hi)
    xi, x = dev(init), dev(coords)
    pn.initialize(nhs, xi, xi)
    if seed != 1:
        pn.update(nhs, x, x)
    og = oracle_like(orc, nhs, np.float64)
    assert og.build(coords)[0] == orc.OK
    st, _, ocnt, osum, _ = og.traverse(coords, coords)
    assert np.array_equal(gpu_counts(nhs, x, x), ocnt)
    assert np.array_equal(gpu_sums(nhs, x, x), osum)
    check_lists(orc, og, nhs, coords, x)
    # mapreduce_neighbor with + over neighbour indices (test/neighborhood_search.jl:389-399)
    for point in (0, coords.shape[0] // 2):
        got = pn.mapreduce_neighbor(pn.NeighborIndex, __import__("operator").add, x, x, nhs, point, init=0)
        assert got == osum[point]


def test_build_deterministic():
    pts = cloud((40, 40, 40), 7)
    nhs = make_nhs(pts, 2.5, pts.min(axis=0) - 2.5, pts.max(axis=0) + 2.5)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    s1, i1 = nhs.cell_table()
    pn.update(nhs, x, x)
    s2, i2 = nhs.cell_table()
    assert np.array_equal(s1, s2) and np.array_equal(i1, i2)
    # cells ascending, original index ascending inside every cell
    for c in np.nonzero(np.diff(s1))[0][:2000]:
        seg = i1[s1[c]:s1[c + 1]]
        assert np.all(np.diff(seg) > 0)


# ---- errors -------------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3])
def test_bounds_errors(golden, n):
    be = golden["bounds_errors"]
    msg = re.escape(be["message"])
    cl = pn.FullGridCellList(min_corner=[0.0] * n, max_corner=[10.0] * n, search_radius=1.0)
    nhs = pn.GridNeighborhoodSearch(n, cell_list=cl, search_radius=1.0)
    y = synthetic.uniform(11 + n, n * 10).reshape(10, n)
    k = be["index"] - 1
    bad = {"NaN": math.nan, "min-0.01": -0.01, "max+1.01": 11.01}
    for key in be["initialize_errors"]:
        y[k, 0] = bad[key]
        d = dev(y)
        with pytest.raises(pn.PointNeighborsError, match=msg):
            pn.initialize(nhs, d, d)
    y[k, 0] = 0.0
    d = dev(y)
    pn.initialize(nhs, d, d)
    pn.update(nhs, d, d)
    y[k, 0] = 10.0
    d = dev(y)
    pn.update(nhs, d, d)
    for key in be["update_errors"]:
        y[k, 0] = bad[key]
        d = dev(y)
        with pytest.raises(pn.PointNeighborsError, match=msg):
            pn.update(nhs, d, d)


def test_overflow_error(golden):
    msg = re.escape(golden["overflow"]["message"])
    cl = pn.FullGridCellList(min_corner=[0.0] * 2, max_corner=[10.0] * 2, search_radius=1.0,
                             max_points_per_cell=100)
    nhs = pn.GridNeighborhoodSearch(2, cell_list=cl, search_radius=1.0)
    d = dev(np.full((100, 2), 5.5))
    pn.initialize(nhs, d, d)
    d = dev(np.full((101, 2), 5.5))
    with pytest.raises(pn.PointNeighborsError, match=msg):
        pn.initialize(nhs, d, d)


def test_status_of_each_update(orc, golden):
    """update! raises the errors of its own binning and no other: valid, NaN, overfull,
    deferred and valid again in a row on one search (each binning publishes its status
    under a new sequence number), and the last valid one counts like the oracle."""
    msg = re.escape(golden["overflow"]["message"])
    msg_bounds = re.escape(golden["bounds_errors"]["message"])
    pts = cloud((12, 11), 3)
    r = 1.6
    nhs = make_nhs(pts, r, pts.min(axis=0) - r, pts.max(axis=0) + r, cap=20)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    assert nhs.stats()[0] == len(pts)
    bad = pts.copy()
    bad[7, 1] = np.nan
    crowd = pts.copy()
    crowd[:40] = pts[60]
    for _ in range(3):
        with pytest.raises(pn.PointNeighborsError, match=msg_bounds):
            pn.update(nhs, dev(bad), dev(bad))
        with pytest.raises(pn.PointNeighborsError, match=msg):
            pn.update(nhs, dev(crowd), dev(crowd))
        nhs.update(dev(crowd), dev(crowd), check_errors=False)
        with pytest.raises(pn.PointNeighborsError, match=msg):
            nhs.check()
        pn.update(nhs, x, x)
        assert nhs.stats()[0] == len(pts)
    og = oracle_like(orc, nhs, np.float64)
    assert og.build(pts)[0] == orc.OK
    assert np.array_equal(gpu_counts(nhs, x, x), og.traverse(pts, pts)[2])


def test_query_bounds_error():
    pts = cloud((10, 10), 1)
    nhs = make_nhs(pts, 2.5)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    far = dev(pts + 100.0)
    with pytest.raises(pn.BoundsError):
        gpu_counts(nhs, far, x)


# ---- x != y, subsets, empty neighbourhoods ---------------------------------------------------------

def test_different_point_sets_and_subsets(orc):
    r = 2.5
    y = cloud((30, 32), 4)
    xs = cloud((30, 32), 5) + 0.5
    lo = np.minimum(y.min(axis=0), xs.min(axis=0)) - r
    hi = np.maximum(y.max(axis=0), xs.max(axis=0)) + r
    nhs = make_nhs(y, r, lo, hi)
    dy, dx = dev(y), dev(xs)
    pn.initialize(nhs, dx, dy)
    og = oracle_like(orc, nhs, np.float64)
    assert og.build(y)[0] == orc.OK
    _, _, ocnt, osum, _ = og.traverse(xs, y)
    assert np.array_equal(gpu_counts(nhs, dx, dy), ocnt)
    assert np.array_equal(gpu_sums(nhs, dx, dy), osum)
    sel = range(5, 700)
    c = gpu_counts(nhs, dx, dy, points=sel)
    assert np.array_equal(c[5:700], ocnt[5:700]) and not c[:5].any() and not c[700:].any()
    idx = [3, 17, 959, 400, 1]
    c = gpu_counts(nhs, dx, dy, points=idx)
    mask = np.zeros(len(c), dtype=bool)
    mask[idx] = True
    assert np.array_equal(c[mask], ocnt[mask]) and not c[~mask].any()
    # eachindex_y subset: only these neighbours are binned
    pn.update(nhs, dy, dy, eachindex_y=range(100, 600))
    og.build(y, active=np.arange(100, 600))
    _, _, ocnt2, _, _ = og.traverse(y, y)
    assert np.array_equal(gpu_counts(nhs, dy, dy), ocnt2)


def test_repeated_selections(orc):
    """A point selected twice is visited twice, as the reference's loops over `points` and
    `eachindex_y` do: a repeated y is binned twice (each copy its own slot in the cell, so
    counts and index sums see it twice); a repeated query point gets its result twice."""
    r = 2.5
    y = cloud((30, 32), 4)
    nhs = make_nhs(y, r, y.min(axis=0) - r, y.max(axis=0) + r)
    dy = dev(y)
    n = len(y)
    active = np.concatenate([np.arange(n), [5, 5, 700, 3]]).astype(np.int64)
    pn.initialize(nhs, dy, dy, eachindex_y=torch.as_tensor(active, device="cuda"))
    og = oracle_like(orc, nhs, np.float64)
    assert og.build(y, active=active)[0] == orc.OK
    _, _, ocnt, osum, _ = og.traverse(y, y)
    assert np.array_equal(gpu_counts(nhs, dy, dy), ocnt)
    assert np.array_equal(gpu_sums(nhs, dy, dy), osum)
    start, idx = nhs.cell_table()
    assert np.array_equal(np.sort(idx), np.sort(active))
    for c in range(len(start) - 1):  # canonical order: ascending index inside every cell
        assert (np.diff(idx[start[c]:start[c + 1]]) >= 0).all()
    pn.initialize(nhs, dy, dy)
    assert og.build(y)[0] == orc.OK
    _, _, ocnt, _, _ = og.traverse(y, y)
    c = gpu_counts(nhs, dy, dy, points=[3, 3, 17, 959, 17])
    mask = np.zeros(n, dtype=bool)
    mask[[3, 17, 959]] = True
    assert np.array_equal(c[mask], ocnt[mask]) and not c[~mask].any()


def test_coordinates_written_through_the_c_abi_are_reread(orc):
    """y moved in place by a C-ABI kernel (the replayable displacement) after update!: the next
    traversal reads the moved coordinates with the cells of the last update!, as the reference
    reads current coordinates (src/nhs_grid.jl:543-545)."""
    from pointneighbors import batch
    r = 2.5
    a = cloud((25, 25), 1)
    nhs = make_nhs(a, r, a.min(axis=0) - 2 * r, a.max(axis=0) + 2 * r)
    x = dev(a)
    pn.initialize(nhs, x, x)
    step = torch.zeros(1, dtype=torch.int64, device="cuda")
    batch.displace_uniform(x, 0.3, 5, step)
    b = x.t().cpu().numpy().copy()
    assert not np.array_equal(a, b)
    og = oracle_like(orc, nhs, np.float64)
    assert og.build(a)[0] == orc.OK
    _, _, ocnt, osum, _ = og.traverse(b, b)
    assert np.array_equal(gpu_counts(nhs, x, x), ocnt)
    assert np.array_equal(gpu_sums(nhs, x, x), osum)
    # a write torch cannot see, declared with coordinates_changed()
    raw = x.clone()
    pn.update(nhs, x, x)
    torch.cuda.synchronize()
    pn.foreach_point_neighbor(pn.CountNeighbors(torch.zeros(len(a), dtype=torch.int32, device="cuda")), x, x, nhs)
    x.data.copy_(dev(a))  # through .data: the version counter does not move
    nhs.coordinates_changed()
    og2 = oracle_like(orc, nhs, np.float64)
    assert og2.build(raw.t().cpu().numpy())[0] == orc.OK
    _, _, ocnt2, _, _ = og2.traverse(a, a)
    assert np.array_equal(gpu_counts(nhs, x, x), ocnt2)


def test_empty_neighborhood_returns_init():
    r = 2.5
    coords = cloud((10, 11), 1)
    nhs = make_nhs(coords, r, coords.min(axis=0) - r, coords.max(axis=0) + r)
    x = dev(coords)
    empty = torch.empty((0, 2), dtype=torch.float64, device="cuda").t()
    pn.initialize(nhs, x, empty)
    assert pn.mapreduce_neighbor(pn.NeighborIndex, __import__("operator").add, x, empty, nhs, 0, init=123) == 123


def test_points_moving_false_keeps_cells(orc):
    r = 2.5
    a = cloud((20, 20), 1)
    b = cloud((20, 20), 2)
    lo = np.minimum(a.min(axis=0), b.min(axis=0)) - r
    hi = np.maximum(a.max(axis=0), b.max(axis=0)) + r
    nhs = make_nhs(a, r, lo, hi)
    da, db = dev(a), dev(b)
    pn.initialize(nhs, da, da)
    s1, i1 = nhs.cell_table()
    pn.update(nhs, db, db, points_moving=(True, False))
    s2, i2 = nhs.cell_table()
    assert np.array_equal(s1, s2) and np.array_equal(i1, i2)


def test_stale_cells_read_current_coordinates(orc):
    """Moving y in place without update!: cells from the last update!, coordinates read
    at traversal time (src/nhs_grid.jl:543-545), query cells from current x."""
    r = 2.5
    a = cloud((25, 25), 1)
    b = a + (synthetic.uniform(9, a.size).reshape(a.shape) - 0.5) * 0.6
    lo = np.minimum(a.min(axis=0), b.min(axis=0)) - r
    hi = np.maximum(a.max(axis=0), b.max(axis=0)) + r
    nhs = make_nhs(a, r, lo, hi)
    x = dev(a)
    pn.initialize(nhs, x, x)
    x.copy_(dev(b))
    og = oracle_like(orc, nhs, np.float64)
    assert og.build(a)[0] == orc.OK
    _, _, ocnt, osum, _ = og.traverse(b, b)
    assert np.array_equal(gpu_counts(nhs, x, x), ocnt)
    assert np.array_equal(gpu_sums(nhs, x, x), osum)


# ---- types ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("coord_dt,radius_dt,corner_dt", [
    (np.float32, np.float32, np.float32), (np.float32, np.float32, np.float64),
    (np.float64, np.float32, np.float64), (np.float32, np.float64, np.float64),
    (np.float64, np.float64, np.float64)])
def test_type_combinations(orc, coord_dt, radius_dt, corner_dt):
    pts = synthetic.jittered_lattice((24, 25, 26), 0.01, 0.001, dtype=coord_dt, seed=3)
    r = radius_dt(2.0) * (radius_dt(1.2) * radius_dt(0.01))
    lo = pts.min(axis=0).astype(corner_dt) - corner_dt(3 * float(r))
    hi = pts.max(axis=0).astype(corner_dt) + corner_dt(3 * float(r))
    nhs = make_nhs(pts, r, lo, hi, corner_dtype=corner_dt)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    og = oracle_like(orc, nhs, coord_dt)
    assert og.build(pts)[0] == orc.OK
    _, _, ocnt, osum, _ = og.traverse(pts, pts)
    assert np.array_equal(gpu_counts(nhs, x, x), ocnt)
    assert np.array_equal(gpu_sums(nhs, x, x), osum)


@pytest.mark.parametrize("shape", [(200,), (5000,)])
def test_one_dimension(orc, shape):
    pts = cloud(shape, 2)
    r = 2.5
    nhs = make_nhs(pts, r, pts.min(axis=0) - r, pts.max(axis=0) + r)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    og = oracle_like(orc, nhs, np.float64)
    assert og.build(pts)[0] == orc.OK
    _, _, ocnt, osum, _ = og.traverse(pts, pts)
    assert np.array_equal(gpu_counts(nhs, x, x), ocnt)
    assert np.array_equal(gpu_sums(nhs, x, x), osum)


def test_fma_sensitive_pairs(orc):
    from test_oracle_golden import adversarial_pairs_f32
    xj, r, want = adversarial_pairs_f32()
    xi = np.full((1, 3), 0.5, dtype=np.float32)
    pts = np.vstack([xi, xj[:200]]).astype(np.float32)
    lo = pts.min(axis=0) - r
    hi = pts.max(axis=0) + r
    nhs = make_nhs(pts, r, lo, hi, cap=1000)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    c = gpu_counts(nhs, dev(pts[:1]), x)
    assert c[0] == 1 + int(want[:200].sum())


# ---- benchmark configurations -----------------------------------------------------------------------

def test_config_small_2d(orc):
    """62,500 2D Float64 points, jitter 0.25, r = 3.0, grid over [min-3, max+3]."""
    pts = synthetic.jittered_lattice((250, 250), 1.0, 0.25, dtype=np.float64, seed=1)
    r = 3.0
    nhs = make_nhs(pts, r, pts.min(axis=0) - 3.0, pts.max(axis=0) + 3.0)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    og = oracle_like(orc, nhs, np.float64)
    assert og.build(pts)[0] == orc.OK
    _, _, ocnt, osum, _ = og.traverse(pts, pts)
    assert np.array_equal(gpu_counts(nhs, x, x), ocnt)
    assert np.array_equal(gpu_sums(nhs, x, x), osum)


def test_config_wcsph_1m_neighbor_sets(orc):
    """1,000,000 3D Float32 points (spacing 0.01, jitter 0.001), r = 2 * (1.2 * 0.01):
    neighbour sets bit-exact after an update! with displaced points."""
    sp = np.float32(0.01)
    pts = synthetic.jittered_lattice((100, 100, 100), 0.01, 0.001, dtype=np.float32, seed=1)
    r = np.float32(2.0) * (np.float32(1.2) * sp)
    lo = pts.min(axis=0) - np.float32(3) * r
    hi = pts.max(axis=0) + np.float32(3) * r
    nhs = make_nhs(pts, r, lo, hi)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    moved = pts + synthetic.displacement(pts.shape[0], 3, 0.0005, seed=1, step=1)
    x.copy_(dev(moved))
    pn.update(nhs, x, x)
    og = oracle_like(orc, nhs, np.float32)
    assert og.build(moved)[0] == orc.OK
    st, _, ocnt, osum, _ = og.traverse(moved, moved)
    c = gpu_counts(nhs, x, x)
    assert np.array_equal(c, ocnt)
    assert 50 < c.mean() < 70
    check_lists(orc, og, nhs, moved, x)


# ---- precomputed ------------------------------------------------------------------------------------

def test_precomputed_own_grid_follows_moving_points(orc):
    """Without a user grid the search makes one over the points' bounding box; update! after
    the points moved far outside it, or packed more than 100 into a cell, still succeeds
    (the reference's default inner search is unbounded) and the lists equal the oracle's."""
    r = 2.5
    a = cloud((20, 21), 3)
    # lists of up to 1000 neighbours (the 150 packed points exceed the 2D default of 64, which
    # the reference's lists would reject too)
    p = pn.PrecomputedNeighborhoodSearch(2, search_radius=r, n_points=len(a), max_neighbors=1000)
    x = dev(a)
    pn.initialize(p, x, x)
    moved = a * np.array([3.0, 0.5]) + np.array([500.0, -80.0])
    moved[:150] = moved[0] + (synthetic.uniform(3, 300).reshape(150, 2) - 0.5) * 0.4  # 150 in one cell
    xm = dev(moved)
    pn.update(p, xm, xm)
    nl = p.neighbor_lists
    rp = nl.row_ptr.cpu().numpy().astype(np.int64)
    cols = nl.col_idx.cpu().numpy()
    orp, ocols = orc.trivial(moved, moved, r, lists=True)[1]  # ascending lists
    assert np.array_equal(rp, orp)
    assert np.array_equal(cols, ocols)


def test_precomputed_lists_and_freeze(orc, golden):
    r = 2.5
    pts = cloud((40, 41), 6)
    lo, hi = pts.min(axis=0) - r, pts.max(axis=0) + r
    inner = make_nhs(pts, r, lo, hi)
    p = pn.PrecomputedNeighborhoodSearch(2, search_radius=r, n_points=pts.shape[0],
                                         update_neighborhood_search=inner)
    x = dev(pts)
    pn.initialize(p, x, x)
    og = oracle_like(orc, inner, np.float64)
    assert og.build(pts)[0] == orc.OK
    st, _, (orp, ocols) = og.precomputed(pts, pts, 64)
    assert st == orc.OK
    rp, cols = p.neighbor_lists.to_host()
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(cols, ocols)
    assert p.neighbor_lists[5] == ocols[orp[5]:orp[6]].tolist()
    frozen = pn.freeze_neighborhood_search(p)
    c = gpu_counts(frozen, x, x)
    s = gpu_sums(frozen, x, x)
    _, _, ocnt, osum, _ = og.traverse(pts, pts)
    assert np.array_equal(c, ocnt) and np.array_equal(s, osum)
    # default inner grid over the bounding box
    p2 = pn.PrecomputedNeighborhoodSearch(2, search_radius=r, n_points=pts.shape[0])
    pn.initialize(p2, x, x)
    assert np.array_equal(p2.neighbor_lists.to_host()[1], ocols)
    # max_neighbors overflow keeps the reference's message
    p3 = pn.PrecomputedNeighborhoodSearch(2, search_radius=r, n_points=pts.shape[0],
                                          update_neighborhood_search=make_nhs(pts, r, lo, hi),
                                          max_neighbors=10)
    with pytest.raises(pn.PointNeighborsError, match=re.escape(golden["overflow"]["message"])):
        pn.initialize(p3, x, x)
    with pytest.raises(pn.PointNeighborsError, match="inactive points"):
        pn.initialize(p, x, x, eachindex_y=range(3, 10))


@pytest.mark.parametrize("n_pts,extent,longest", [(900, 6.0, 256), (2600, 9.0, 512)])
def test_precomputed_long_rows(orc, n_pts, extent, longest):
    """Rows from a few to ~1,500 neighbours: every row-sort path (1, 2, 4, 8 keys per lane in
    registers, LDS beyond 512) must give the oracle's sorted lists."""
    rng = np.random.default_rng(n_pts)
    # a dense core and a sparse halo: row lengths spread over the whole range
    core = rng.normal(0.0, 0.12 * extent, size=(n_pts // 2, 3))
    halo = rng.uniform(-extent, extent, size=(n_pts - n_pts // 2, 3))
    pts = np.concatenate([core, halo]).astype(np.float64)
    r = 0.3 * extent
    lo, hi = pts.min(axis=0) - r, pts.max(axis=0) + r
    inner = make_nhs(pts, r, lo, hi, cap=4096)
    p = pn.PrecomputedNeighborhoodSearch(3, search_radius=r, n_points=len(pts),
                                         update_neighborhood_search=inner, max_neighbors=4096)
    x = dev(pts)
    pn.initialize(p, x, x)
    og = oracle_like(orc, inner, np.float64)
    assert og.build(pts)[0] == orc.OK
    st, _, (orp, ocols) = og.precomputed(pts, pts, 4096)
    assert st == orc.OK
    lens = np.diff(orp)
    assert lens.max() > longest and (lens <= 64).any() and ((lens > 128) & (lens <= 256)).any()
    rp, cols = p.neighbor_lists.to_host()
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(cols, ocols)


@pytest.mark.parametrize("transpose", [False, True])
def test_dvov_layouts(orc, transpose):
    """CSR lists exported to the reference's DynamicVectorOfVectors backend (plain and
    transpose_backend): every list and length equals the CSR row, in both layouts."""
    coords = cloud((23, 21, 19), 5, dtype=np.float32)
    r = np.float32(2.2)
    nhs = make_nhs(coords, r, coords.min(axis=0) - r, coords.max(axis=0) + r)
    x = dev(coords)
    p = pn.PrecomputedNeighborhoodSearch(3, search_radius=r, n_points=len(coords), update_neighborhood_search=nhs,
                                         max_neighbors=128)
    pn.initialize(p, x, x)
    rp, cols = p.neighbor_lists.to_host()
    longest = int(np.diff(rp).max())
    backend, lengths = p.neighbor_lists.to_dvov(max_neighbors=longest + 3, transpose_backend=transpose,
                                                max_outer_length=len(coords) + 7 if transpose else None)
    b = backend.cpu().numpy()
    if transpose:
        assert b.shape == (longest + 3, len(coords) + 7)
        b = b.T
    assert np.array_equal(lengths.cpu().numpy(), np.diff(rp))
    for i in range(len(coords)):
        assert np.array_equal(b[i, :rp[i + 1] - rp[i]], cols[rp[i]:rp[i + 1]])
    with pytest.raises(pn.PointNeighborsError):
        p.neighbor_lists.to_dvov(max_neighbors=longest - 1)
