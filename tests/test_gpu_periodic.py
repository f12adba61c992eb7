"""Periodic boxes on the GPU (GridNeighborhoodSearch / PrecomputedNeighborhoodSearch with a
PeriodicBox): neighbour sets equal the reference's hand-made cases and the oracle's,
counts and index sums bit-exact; a periodic lattice has one density everywhere."""

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import sph, synthetic
from test_gpu_parity import dev, gpu_counts, gpu_sums

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def periodic_nhs(lo, hi, r, T=np.float64, hashed=False):
    pb = pn.PeriodicBox(min_corner=np.asarray(lo, dtype=T), max_corner=np.asarray(hi, dtype=T))
    if hashed:  # a small table: wrapped cells collide in buckets
        cl = pn.SpatialHashingCellList[len(lo)](list_size=97)
        return pn.GridNeighborhoodSearch(len(lo), search_radius=T(r), periodic_box=pb, cell_list=cl)
    cl = pn.FullGridCellList(min_corner=pb.min_corner, max_corner=pb.max_corner, search_radius=T(r))
    return pn.GridNeighborhoodSearch(len(lo), search_radius=T(r), periodic_box=pb, cell_list=cl)


def oracle_for(orc, nhs, T):
    cl = nhs.cell_list
    g = orc.Grid(nhs.ndims, cl.min_corner.astype(np.float64), cl.n_cells, float(nhs.search_radius),
                 coord_dtype=T, radius_dtype=nhs.radius_dtype, cell_dtype=cl.corner_dtype,
                 max_points_per_cell=cl.max_points_per_cell)
    g.set_periodic(np.asarray(nhs.cell_size, dtype=np.float64), nhs.n_cells,
                   nhs.periodic_box.size.astype(np.float64))
    return g


@pytest.mark.parametrize("hashed", [False, True])
def test_reference_cases(orc, golden, hashed):
    P = golden["periodic"]
    r = P["search_radius"]
    want = [[v - 1 for v in l] for l in P["neighbors"]]
    for case in P["cases"]:
        pts = np.asarray(case["coords"])
        x = dev(pts)
        nhs = periodic_nhs(case["min_corner"], case["max_corner"], r, hashed=hashed)
        pn.initialize(nhs, x, x)
        assert gpu_counts(nhs, x, x).tolist() == [len(w) for w in want], case["name"]
        assert gpu_sums(nhs, x, x).tolist() == [sum(w) for w in want], case["name"]
        p = pn.PrecomputedNeighborhoodSearch(pts.shape[1], search_radius=r, n_points=len(pts),
                                             periodic_box=nhs.periodic_box,
                                             update_neighborhood_search=nhs if hashed else None)
        pn.initialize(p, x, x)
        assert [p.neighbor_lists[i] for i in range(len(pts))] == want, case["name"]


@pytest.mark.parametrize("nd,T,box,r", [
    (3, np.float32, ([0.0, 0.0, 0.0], [10.0, 12.0, 9.0]), 2.4),
    (3, np.float64, ([-3.0, 0.5, 1.0], [4.5, 7.0, 5.0]), 1.3),
    (2, np.float64, ([0.0, 0.0], [30.0, 20.5]), 2.9),
    (2, np.float32, ([-1.0, -1.0], [8.0, 16.0]), 2.9),   # 3 periodic cells in x
    (3, np.float32, ([0.0, 0.0, 0.0], [7.5, 7.5, 7.5]), 2.4),  # 3 cells per dimension
])
@pytest.mark.parametrize("hashed", [False, True])
def test_clouds_vs_oracle(orc, nd, T, box, r, hashed):
    rng = np.random.default_rng(101 + nd)
    lo = np.asarray(box[0], dtype=T)
    hi = np.asarray(box[1], dtype=T)
    size = hi - lo
    n = int(np.prod(size) * 0.9)
    # mostly inside the box, some a little outside (cells wrap, coordinates do not)
    u = rng.uniform(-0.03, 1.03, size=(n, nd))
    pts = (lo + u * size).astype(T)
    nhs = periodic_nhs(lo, hi, r, T, hashed=hashed)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    og = oracle_for(orc, periodic_nhs(lo, hi, r, T), T)
    assert og.build(pts)[0] == orc.OK
    st, _, ocnt, osum, (orp, ocols) = og.traverse(pts, pts, lists=True)
    assert st == orc.OK
    for algorithm in (0, 1):  # a periodic grid always runs the cell kernel
        nhs.set_tuning("algorithm", algorithm)
        assert np.array_equal(gpu_counts(nhs, x, x), ocnt)
        assert np.array_equal(gpu_sums(nhs, x, x), osum)
    p = pn.PrecomputedNeighborhoodSearch(nd, search_radius=T(r), n_points=n, periodic_box=nhs.periodic_box,
                                         update_neighborhood_search=nhs, max_neighbors=400)
    pn.initialize(p, x, x)
    rp, cols = p.neighbor_lists.to_host()
    want = np.concatenate([np.sort(ocols[orp[i]:orp[i + 1]]) for i in range(n)])
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(cols, want)


def test_queries_outside_grid_set(orc):
    """x != y: query points binned into the periodic grid of y."""
    T = np.float64
    rng = np.random.default_rng(7)
    lo, hi = np.zeros(3), np.array([9.0, 8.0, 7.5])
    y = (lo + rng.uniform(0, 1, size=(600, 3)) * (hi - lo)).astype(T)
    xq = (lo + rng.uniform(-0.1, 1.1, size=(300, 3)) * (hi - lo)).astype(T)
    nhs = periodic_nhs(lo, hi, 2.2, T)
    pn.initialize(nhs, dev(xq), dev(y))
    og = oracle_for(orc, nhs, T)
    assert og.build(y)[0] == orc.OK
    _, _, ocnt, osum, _ = og.traverse(xq, y)
    assert np.array_equal(gpu_counts(nhs, dev(xq), dev(y)), ocnt)
    assert np.array_equal(gpu_sums(nhs, dev(xq), dev(y)), osum)


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_periodic_lattice_uniform_density(T):
    """On a lattice that tiles the periodic box every point sees the same neighbourhood, so
    the density is the same everywhere (up to summation order)."""
    m = 12
    pts = synthetic.jittered_lattice((m, m, m), 1.0, 0.0, dtype=T)
    lo = np.full(3, 0.5, dtype=T)   # lattice points at 1..m; the box [0.5, m + 0.5) tiles it
    hi = np.full(3, m + 0.5, dtype=T)
    h = T(1.2)
    nhs = periodic_nhs(lo, hi, T(2) * h, T)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    rho = torch.zeros(len(pts), dtype=x.dtype, device=x.device)
    mass = torch.ones(len(pts), dtype=x.dtype, device=x.device)
    pn.foreach_point_neighbor(sph.SPHDensity(rho, mass, float(h)), x, x, nhs)
    rho = rho.cpu().numpy()
    tol = 1e-5 if T == np.float32 else 1e-12
    assert np.allclose(rho, rho[0], rtol=tol, atol=0)
    counts = gpu_counts(nhs, x, x)
    assert (counts == counts[0]).all()
    # the interior value of the same lattice without periodicity
    cl = pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=T(2) * h)
    plain = pn.GridNeighborhoodSearch(3, search_radius=T(2) * h, cell_list=cl)
    pn.initialize(plain, x, x)
    rho2 = torch.zeros(len(pts), dtype=x.dtype, device=x.device)
    pn.foreach_point_neighbor(sph.SPHDensity(rho2, mass, float(h)), x, x, plain)
    centre = (m // 2 - 1) + m * (m // 2 - 1) + m * m * (m // 2 - 1)
    assert np.isclose(rho2.cpu().numpy()[centre], rho[0], rtol=tol, atol=0)


def test_periodic_list_density_matches_grid():
    """PrecomputedNeighborhoodSearch with a PeriodicBox: the list density (minimum image of
    the stored pairs) equals the grid traversal's density."""
    T = np.float64
    rng = np.random.default_rng(5)
    lo, hi = np.zeros(3), np.array([8.0, 9.0, 7.5])
    pts = (lo + rng.uniform(-0.02, 1.02, size=(900, 3)) * (hi - lo)).astype(T)
    h = 1.1
    nhs = periodic_nhs(lo, hi, 2 * h, T)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    mass = torch.rand(len(pts), dtype=x.dtype, device=x.device) + 0.5
    rho_g = torch.zeros(len(pts), dtype=x.dtype, device=x.device)
    pn.foreach_point_neighbor(sph.SPHDensity(rho_g, mass, h), x, x, nhs)
    p = pn.PrecomputedNeighborhoodSearch(3, search_radius=T(2 * h), n_points=len(pts),
                                         periodic_box=nhs.periodic_box)
    pn.initialize(p, x, x)
    rho_l = torch.zeros_like(rho_g)
    pn.foreach_point_neighbor(sph.SPHDensity(rho_l, mass, h), x, x, p)
    assert torch.allclose(rho_l, rho_g, rtol=1e-12, atol=0)
