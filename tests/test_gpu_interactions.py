"""The reference's benchmark closures and per-point reductions on the GPU against the oracle:
n-body acceleration (benchmarks/n_body.jl:34-51) over grids and precomputed lists, the WCSPH
fluid interaction timed by benchmark_wcsph (benchmarks/smoothed_particle_hydrodynamics.jl:44-102)
and mapreduce_neighbor (src/neighborhood_search.jl:340-383) with several value functions and
operators.  Sums: |gpu - oracle| <= tol * sum|term|, tol = 1e-5 (Float32) and 1e-12 (Float64)."""

import operator

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import synthetic
from test_gpu_parity import dev, make_nhs, oracle_like

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = {np.float32: 1e-5, np.float64: 1e-12}


def grid_and_lists(orc, pts, r):
    lo = pts.min(axis=0) - np.asarray(r, dtype=pts.dtype) * 3
    hi = pts.max(axis=0) + np.asarray(r, dtype=pts.dtype) * 3
    nhs = make_nhs(pts, r, lo, hi)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    og = oracle_like(orc, nhs, pts.dtype)
    assert og.build(pts)[0] == orc.OK
    st, _, lists = og.precomputed(pts, pts, 100000)
    assert st == orc.OK
    return nhs, x, lists


def tdev(a, dt):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=dt), device="cuda")


def coords_like(a, dt):
    """(N, K) host array -> K x N device matrix with strides (1, K)."""
    return tdev(a, dt).t()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("nd", [2, 3])
def test_nbody_grid_and_lists(orc, dt, nd):
    shape = (24, 25, 23) if nd == 3 else (90, 85)
    pts = synthetic.jittered_lattice(shape, 0.1, 0.03, dtype=dt, seed=11)
    r = dt(0.25)
    nhs, x, lists = grid_and_lists(orc, pts, r)
    n = len(pts)
    m = (1e10 * (synthetic.uniform(12, n) + 1)).astype(dt)
    G = 6.6743e-11
    want, scale = orc.nbody(pts, lists, m, G, eltype=dt)
    dv = torch.zeros((n, nd), dtype=x.dtype, device="cuda").t()
    pn.foreach_point_neighbor(pn.NBodyAcceleration(dv, tdev(m, dt), G), x, x, nhs)
    err = np.abs(dv.t().cpu().numpy().astype(np.float64) - want).max(axis=1)
    assert (err <= TOL[dt] * scale).all(), (err / scale).max()
    # the same closure over the precomputed lists of a frozen search
    pre = pn.PrecomputedNeighborhoodSearch(nd, search_radius=r, n_points=n, update_neighborhood_search=nhs)
    pn.initialize(pre, x, x)
    dvl = torch.zeros_like(dv)
    pn.foreach_point_neighbor(pn.NBodyAcceleration(dvl, tdev(m, dt), G), x, x,
                              pn.freeze_neighborhood_search(pre))
    err = np.abs(dvl.t().cpu().numpy().astype(np.float64) - want).max(axis=1)
    assert (err <= TOL[dt] * scale).all(), (err / scale).max()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("nd", [2, 3])
def test_wcsph_interaction(orc, dt, nd):
    """The benchmark's models (WendlandC2, h = r/2, alpha 0.02, Molteni-Colagrossi delta 0.1) with
    random velocities (viscosity on for approaching pairs), densities, pressures and masses;
    accumulate adds to dv."""
    sp = 0.01
    shape = (22, 23, 21) if nd == 3 else (80, 75)
    pts = synthetic.jittered_lattice(shape, sp, 0.07 * sp, dtype=dt, seed=13)
    r = dt(3) * dt(sp)
    h = float(r) / 2
    nhs, x, lists = grid_and_lists(orc, pts, r)
    n = len(pts)
    m = (0.1 * sp * (1 + 0.1 * synthetic.uniform(14, n))).astype(dt)
    rho = (1000 + synthetic.uniform(15, n)).astype(dt)
    vel = ((synthetic.uniform(16, nd * n).reshape(n, nd) - 0.5) * 0.2).astype(dt)
    v = np.concatenate([vel, rho[:, None]], axis=1)
    c = 10.0
    p = pn.state_equation_cole(tdev(rho, dt), 1000.0, c, 1.0)
    pn_host = p.cpu().numpy().astype(np.float64)
    # p = rho0 c^2 (rho / rho0 - 1): Float32 rounds rho / rho0 to 2^-24, i.e. p to about 6e-3
    np.testing.assert_allclose(pn_host, c**2 * (rho.astype(np.float64) - 1000.0), rtol=0,
                               atol=2e-2 if dt == np.float32 else 1e-9)
    want, scale = orc.wcsph_interact(pts, lists, m, v, pn_host, h, c, eltype=dt)
    vd = coords_like(v, dt)
    dv = torch.full((n, nd + 1), 7.0, dtype=x.dtype, device="cuda").t()
    pn.foreach_point_neighbor(pn.WCSPHInteraction(dv, vd, p, tdev(m, dt), h, c, accumulate=False), x, x, nhs)
    got = dv.t().cpu().numpy().astype(np.float64)
    em = np.abs(got[:, :nd] - want[:, :nd]).max(axis=1)
    ed = np.abs(got[:, nd] - want[:, nd])
    tol = TOL[dt]
    assert (em <= tol * scale[:, 0]).all(), (em / scale[:, 0]).max()
    assert (ed <= tol * scale[:, 1]).all(), (ed / scale[:, 1]).max()
    # momentum conservation of the GPU result itself
    mom = (m.astype(np.float64)[:, None] * got[:, :nd]).sum(axis=0)
    assert np.abs(mom).max() <= (1e-4 if dt == np.float32 else 1e-11) * (m * scale[:, 0]).sum()
    # interact! accumulates into dv: a second call doubles it (the sums are deterministic)
    pn.foreach_point_neighbor(pn.WCSPHInteraction(dv, vd, p, tdev(m, dt), h, c), x, x, nhs)
    assert np.array_equal(dv.t().cpu().numpy(), (2 * got).astype(dt))


def test_wcsph_interaction_rejects_other_queries(orc):
    pts = synthetic.jittered_lattice((10, 10, 10), 1.0, 0.1, dtype=np.float32, seed=1)
    nhs, x, _ = grid_and_lists(orc, pts, np.float32(2.5))
    n = len(pts)
    v = torch.zeros((n, 4), dtype=torch.float32, device="cuda").t()
    f = pn.WCSPHInteraction(torch.zeros_like(v), v, torch.zeros(n, device="cuda"),
                            torch.ones(n, device="cuda"), 1.25, 10.0)
    with pytest.raises(pn.ArgumentError):
        pn.foreach_point_neighbor(f, x, x, nhs, points=range(3, 9))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_mapreduce_posdiff_components_3d(orc, dt):
    """Each pos_diff component in 3D (the component is selected per call): min and max are
    exact, sums within the summation-order tolerance."""
    pts = synthetic.jittered_lattice((11, 12, 13), 1.0, 0.3, dtype=dt, seed=23)
    r = dt(2.2)
    nhs, x, lists = grid_and_lists(orc, pts, r)
    rp, cols = lists
    n = len(pts)
    rows = np.repeat(np.arange(n), np.diff(rp))
    pd = pts[rows].astype(dt) - pts[cols].astype(dt)
    for c in range(3):
        vals = pd[:, c].astype(np.float64)
        lo = np.full(n, 1e9)
        np.minimum.at(lo, rows, vals)
        hi = np.full(n, -1e9)
        np.maximum.at(hi, rows, vals)
        sm = np.zeros(n)
        np.add.at(sm, rows, vals)
        ab = np.zeros(n)
        np.add.at(ab, rows, np.abs(vals))
        got_lo, _ = pn.mapreduce_point_neighbors(pn.PosDiff(c), min, x, x, nhs, init=1e9)
        got_hi, _ = pn.mapreduce_point_neighbors(pn.PosDiff(c), max, x, x, nhs, init=-1e9)
        got_sm, _ = pn.mapreduce_point_neighbors(pn.PosDiff(c), operator.add, x, x, nhs, init=0.0)
        assert np.array_equal(got_lo.cpu().numpy(), lo), c
        assert np.array_equal(got_hi.cpu().numpy(), hi), c
        tol = (1e-5 if dt == np.float32 else 1e-12) * ab
        assert (np.abs(got_sm.cpu().numpy() - sm) <= tol).all(), c


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_mapreduce_menu(orc, dt):
    """Per-point reductions against the oracle's neighbour lists: counts, index sums, distances
    (sum, min, max), squared distances and pos_diff components; init for an empty
    neighbourhood; the single-point form."""
    pts = synthetic.jittered_lattice((30, 31), 1.0, 0.3, dtype=dt, seed=17)
    r = dt(2.5)
    nhs, x, lists = grid_and_lists(orc, pts, r)
    rp, cols = lists
    n = len(pts)
    rows = np.repeat(np.arange(n), np.diff(rp))
    pd = pts[rows].astype(dt) - pts[cols].astype(dt)
    d2 = (pd[:, 0] * pd[:, 0] + pd[:, 1] * pd[:, 1]).astype(dt)
    dist = np.sqrt(d2).astype(dt)

    def reduce(vals, op, init):
        out = np.full(n, float(init))
        if op is operator.add:
            np.add.at(out, rows, vals.astype(np.float64))
        elif op is min:
            np.minimum.at(out, rows, vals.astype(np.float64))
        else:
            np.maximum.at(out, rows, vals.astype(np.float64))
        return out

    cases = [(pn.One, operator.add, 0, np.ones(len(cols))), (pn.NeighborIndex, operator.add, 5, cols),
             (pn.Distance, operator.add, 0.0, dist), (pn.Distance, min, 1e9, dist), (pn.Distance, max, -1.0, dist),
             (pn.DistanceSquared, max, 0.0, d2), (pn.PosDiff(1), operator.add, 0.0, pd[:, 1])]
    for f, op, init, vals in cases:
        got, cnt = pn.mapreduce_point_neighbors(f, op, x, x, nhs, init=init)
        assert np.array_equal(cnt.cpu().numpy(), np.diff(rp))
        want = reduce(np.asarray(vals), op, init)
        if op is operator.add:
            tol = (1e-5 if dt == np.float32 else 1e-12) * reduce(np.abs(np.asarray(vals, dtype=np.float64)),
                                                                  operator.add, abs(init))
            assert (np.abs(got.cpu().numpy() - want) <= tol).all(), f
        else:
            assert np.array_equal(got.cpu().numpy(), want), f
    k = 7 * 31 + 3
    assert pn.mapreduce_neighbor(pn.NeighborIndex, operator.add, x, x, nhs, k, init=0) == int(cols[rp[k]:rp[k + 1]].sum())
    assert pn.mapreduce_neighbor(pn.Distance, min, x, x, nhs, k, init=1e9) == 0.0  # itself
    # an empty neighbourhood returns init, f and op unevaluated
    far = dev(np.full((1, 2), 1e6, dtype=dt))
    lonely = make_nhs(np.array([[1e6, 1e6]], dtype=dt), r, np.full(2, 1e6 - 10, dtype=dt),
                      np.full(2, 1e6 + 10, dtype=dt))
    pn.initialize(lonely, far, far[:, :0])
    assert pn.mapreduce_neighbor(pn.Distance, max, far, far[:, :0], lonely, 0, init=-7.5) == -7.5
