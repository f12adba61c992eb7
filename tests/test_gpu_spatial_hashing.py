"""GridNeighborhoodSearch with a SpatialHashingCellList on the GPU: the reference's collision
fixtures (test/cell_lists/spatial_hashing.jl:2-70), neighbour counts, index sums and sorted
lists bit-exact against the oracle's brute force (the method of
test/neighborhood_search.jl:186-469, which runs SpatialHashingCellList against
TrivialNeighborhoodSearch), with collisions forced by tiny list sizes, points far apart (an
unbounded domain), x != y, the bucket overflow and Int32 errors, and SPH density equal to the
FullGridCellList path within the summation-order tolerance."""

import re

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import sph, synthetic
from pointneighbors.cell_lists import spatial_hash
from test_gpu_parity import cloud, dev, gpu_counts, gpu_sums, make_nhs

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def hash_nhs(nd, r, n, list_size, cap=100):
    cl = pn.SpatialHashingCellList[nd](list_size=list_size, max_points_per_cell=cap)
    return pn.GridNeighborhoodSearch(nd, search_radius=r, n_points=n, cell_list=cl)


def lists_of(nhs, x, n):
    p = pn.PrecomputedNeighborhoodSearch(x.shape[0], search_radius=nhs.search_radius, n_points=n,
                                         update_neighborhood_search=nhs, max_neighbors=100000)
    pn.initialize(p, x, x)
    rp, cols = p.neighbor_lists.to_host()
    return [cols[rp[i]:rp[i + 1]].tolist() for i in range(n)]


@pytest.mark.parametrize("case", ["empty", "nonempty"])
def test_reference_collision_fixtures(case):
    r = 0.1 + 10 * np.finfo(np.float64).eps
    if case == "empty":  # the point is in cell (-1, 0), which collides with the empty cell (-2, -1)
        pts = np.array([[-0.05, 0.05]])
        want_bucket, want_lists = [0], [[0]]
    else:  # cell (-1, 0) with point 1 collides with cell (-2, -1) with point 2
        pts = np.array([[-0.05, 0.05], [-0.15, -0.05]])
        want_bucket, want_lists = [0, 1], [[0], [1]]
    n = len(pts)
    nhs = hash_nhs(2, r, n, list_size=n)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    assert spatial_hash((-1, 0), n) == spatial_hash((-2, -1), n)
    assert sorted(nhs.points_in_cell((-1, 0))) == sorted(nhs.points_in_cell((-2, -1))) == want_bucket
    assert lists_of(nhs, x, n) == want_lists
    assert gpu_counts(nhs, x, x).tolist() == [1] * n


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("list_size", [1, 7, None])
@pytest.mark.parametrize("shape", [(23, 21), (9, 10, 8), (40,)])
def test_vs_brute_force(orc, shape, list_size, dt):
    r = dt(2.5)
    pts = (cloud(shape, 3) - np.array(shape) / 2.0).astype(dt)  # negative coordinates too
    n, nd = pts.shape
    nhs = hash_nhs(nd, r, n, list_size or 2 * n, cap=100000)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    cnt, (orp, ocols) = orc.trivial(pts, pts, float(r), coord_dtype=dt, radius_dtype=dt, lists=True)
    assert np.array_equal(gpu_counts(nhs, x, x), cnt)
    want_sums = np.array([ocols[orp[i]:orp[i + 1]].astype(np.int64).sum() for i in range(n)])
    assert np.array_equal(gpu_sums(nhs, x, x), want_sums)
    got = lists_of(nhs, x, n)
    assert all(got[i] == ocols[orp[i]:orp[i + 1]].tolist() for i in range(n))
    # every point sits in the bucket of its own cell
    start, idx = nhs.cell_table()
    L = nhs.cell_list.list_size
    for b in range(L):
        for p in idx[start[b]:start[b + 1]]:
            assert spatial_hash(nhs.host_cell_coords(pts[p]), L) == b + 1


def test_unbounded_domain_and_queries(orc):
    """Two clusters 1e6 apart (a FullGridCellList would need ~1e12 cells) and x != y."""
    a = cloud((12, 11, 10), 5)
    b = cloud((9, 9, 9), 6) + 1.0e6
    pts = np.concatenate([a, b - 2.0e6 * np.array([0.0, 1.0, 0.0])])
    n = len(pts)
    r = 2.2
    nhs = hash_nhs(3, r, n, 2 * n)
    y = dev(pts)
    pn.initialize(nhs, y, y)
    cnt, _ = orc.trivial(pts, pts, r)
    assert np.array_equal(gpu_counts(nhs, y, y), cnt)
    q = pts[::7] + 0.37
    xq = dev(q)
    qcnt, _ = orc.trivial(q, pts, r)
    assert np.array_equal(gpu_counts(nhs, xq, y), qcnt)


def test_bucket_overflow_and_nan(golden):
    msg = re.escape(golden["overflow"]["message"])
    nhs = hash_nhs(2, 1.0, 0, list_size=3, cap=5)
    d = dev(np.array([[0.5 + i, 0.5] for i in range(40)]))  # 40 points in 3 buckets
    with pytest.raises(pn.PointNeighborsError, match=msg):
        pn.initialize(nhs, d, d)
    ok = hash_nhs(2, 1.0, 0, list_size=64, cap=100)
    bad = np.array([[0.5, 0.5], [np.nan, 1.0]])
    with pytest.raises(pn.PointNeighborsError, match="InexactError"):
        pn.initialize(ok, dev(bad), dev(bad))
    big = np.array([[0.5, 0.5], [3.0e9, 1.0]])  # cell 3e9 does not fit in Int32
    with pytest.raises(pn.PointNeighborsError, match="InexactError"):
        pn.initialize(ok, dev(big), dev(big))


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_sph_density_matches_full_grid(dt, tol):
    h = dt(1.2)
    r = dt(2 * h)
    pts = cloud((20, 19, 18), 7, dtype=dt)
    n = len(pts)
    x = dev(pts)
    mass = torch.ones(n, dtype=x.dtype, device=x.device)
    out = []
    for nhs in (hash_nhs(3, r, n, 2 * n), make_nhs(pts, r, pts.min(axis=0) - r, pts.max(axis=0) + r)):
        pn.initialize(nhs, x, x)
        rho = torch.zeros(n, dtype=x.dtype, device=x.device)
        pn.foreach_point_neighbor(sph.SPHDensity(rho, mass, float(h), "cubic_spline"), x, x, nhs)
        out.append(rho)
    assert torch.allclose(out[0], out[1], rtol=tol, atol=0)


def test_stale_coordinates_follow_current_cells():
    """Points moved in place without update!: buckets come from the last update!, and a
    candidate counts only if its current cell is the scanned cell (check_collision,
    src/nhs_grid.jl:479-513), applied to every bucket.  The reference's parallel build never
    records a bucket's cell (push_cell_atomic! compares a UInt128 with an NTuple,
    src/cell_lists/spatial_hashing.jl:103-116), so it flags every bucket that holds a point of
    a cell other than (0, ..., 0) and checks those; a bucket holding only points of cell
    (0, ..., 0) stays unflagged and its stale points are accepted unchecked.  This build checks
    that bucket too (a recorded deviation).  Expected counts from the checked loop restated in
    numpy (src/nhs_grid.jl:519-575)."""
    r = 2.5
    a = cloud((25, 25), 1)
    b = a + (synthetic.uniform(9, a.size).reshape(a.shape) - 0.5) * 0.6
    n = len(a)
    L = 2 * n
    nhs = hash_nhs(2, r, n, L)
    x = dev(a)
    pn.initialize(nhs, x, x)
    x.copy_(dev(b))
    cell = lambda p: tuple(int(v) for v in np.floor(p / r))
    old_bucket = np.array([spatial_hash(cell(p), L) for p in a])
    cur_cell = [cell(p) for p in b]
    want = np.zeros(n, dtype=np.int64)
    for q in range(n):
        cq = cell(b[q])
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                nc = (cq[0] + dx, cq[1] + dy)
                for p in np.nonzero(old_bucket == spatial_hash(nc, L))[0]:
                    if cur_cell[p] == nc and np.sum((b[q] - b[p]) ** 2) <= r * r:
                        want[q] += 1
    assert np.array_equal(gpu_counts(nhs, x, x), want)
