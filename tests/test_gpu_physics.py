"""GPU interaction functors and adjoints against the oracle.

Sums: |gpu - oracle| <= tol * sum_j |term_j| (the summation-error scale; for the density
all terms are positive, so this is the relative error), tol = 1e-5 for Float32 and 1e-12
for Float64 (summation order differs).  Gradients: relative error
|g - g_ref| / max(|g_ref|, 1e-3 max|g_ref|) against forward-mode duals through the oracle
and central finite differences (step 1e-6) of the GPU forward pass."""

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import sph, synthetic
from test_gpu_parity import dev, make_nhs, oracle_like

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def setup(orc, pts, r, coord_dt, radius_dt):
    lo = pts.min(axis=0) - np.asarray(r, dtype=pts.dtype) * 3
    hi = pts.max(axis=0) + np.asarray(r, dtype=pts.dtype) * 3
    nhs = make_nhs(pts, radius_dt(r), lo, hi)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    og = oracle_like(orc, nhs, coord_dt)
    assert og.build(pts)[0] == orc.OK
    st, _, lists = og.precomputed(pts, pts, 100000)
    assert st == orc.OK
    return nhs, x, lists


def tdev(a, dt):
    return torch.as_tensor(np.asarray(a, dtype=dt), device="cuda")


@pytest.mark.parametrize("dt,tol", [(np.float32, 1e-5), (np.float64, 1e-12)])
@pytest.mark.parametrize("kernel", ["cubic_spline", "wendland_c2"])
def test_density_and_pressure(orc, dt, tol, kernel):
    sp = 0.01
    pts = synthetic.jittered_lattice((30, 31, 32), sp, 0.1 * sp, dtype=dt, seed=2)
    h = dt(1.2) * dt(sp)
    r = dt(2) * h
    nhs, x, lists = setup(orc, pts, r, dt, dt)
    n = len(pts)
    rho0 = 1000.0
    m = np.full(n, rho0 * sp**3, dtype=dt)
    rho = torch.zeros(n, dtype=x.dtype, device="cuda")
    pn.foreach_point_neighbor(sph.SPHDensity(rho, tdev(m, dt), float(h), kernel), x, x, nhs)
    orho, scale = orc.sph_density(pts, lists, m, float(h), kernel)
    err = np.abs(rho.cpu().numpy().astype(np.float64) - orho)
    assert (err <= tol * scale).all(), err.max()
    # pressure with P from the oracle's density (same P for both sides)
    p = orc.tait_eos(orho, rho0, 10.0)
    P = (p / orho**2).astype(dt)
    dv = sph._vec_like(x).zero_()
    pn.foreach_point_neighbor(sph.SPHPressureAcceleration(dv, tdev(m, dt), tdev(P, dt), float(h), kernel),
                              x, x, nhs)
    odv, scale = orc.sph_pressure_accel(pts, lists, m, P, float(h), kernel)
    err = np.abs(dv.t().cpu().numpy().astype(np.float64) - odv).max(axis=1)
    assert (err <= tol * np.maximum(scale, 1e-300)).all(), (err / scale).max()


@pytest.mark.parametrize("dt,tol", [(np.float32, 1e-5), (np.float64, 1e-12)])
@pytest.mark.parametrize("kernel", ["cubic_spline", "wendland_c2"])
def test_density_and_pressure_2d(orc, dt, tol, kernel):
    """2D: candidate records without a padding slot, 2D kernel normalisations."""
    sp = 0.02
    pts = synthetic.jittered_lattice((70, 61), sp, 0.15 * sp, dtype=dt, seed=4)
    h = dt(1.3) * dt(sp)
    r = dt(2) * h
    nhs, x, lists = setup(orc, pts, r, dt, dt)
    n = len(pts)
    m = np.full(n, 1000.0 * sp**2, dtype=dt)
    rho = torch.zeros(n, dtype=x.dtype, device="cuda")
    pn.foreach_point_neighbor(sph.SPHDensity(rho, tdev(m, dt), float(h), kernel), x, x, nhs)
    orho, scale = orc.sph_density(pts, lists, m, float(h), kernel)
    err = np.abs(rho.cpu().numpy().astype(np.float64) - orho)
    assert (err <= tol * scale).all(), err.max()
    P = (orc.tait_eos(orho, 1000.0, 10.0) / orho**2).astype(dt)
    dv = sph._vec_like(x).zero_()
    pn.foreach_point_neighbor(sph.SPHPressureAcceleration(dv, tdev(m, dt), tdev(P, dt), float(h), kernel),
                              x, x, nhs)
    odv, scale = orc.sph_pressure_accel(pts, lists, m, P, float(h), kernel)
    err = np.abs(dv.t().cpu().numpy().astype(np.float64) - odv).max(axis=1)
    assert (err <= tol * np.maximum(scale, 1e-300)).all(), (err / scale).max()


def test_eos_kernel(orc):
    rho = np.linspace(990.0, 1010.0, 1000)
    p = sph.tait_eos(tdev(rho, np.float64), 1000.0, 10.0, 7.0).cpu().numpy()
    assert np.allclose(p, orc.tait_eos(rho, 1000.0, 10.0, 7.0), rtol=1e-13, atol=1e-9)
    # the fused pressure term P = p / rho^2
    want = orc.tait_eos(rho, 1000.0, 10.0, 7.0) / rho**2
    P = sph.tait_pressure_term(tdev(rho, np.float64), 1000.0, 10.0, 7.0).cpu().numpy()
    assert np.allclose(P, want, rtol=1e-13, atol=1e-15)
    P32 = sph.tait_pressure_term(tdev(rho, np.float32), 1000.0, 10.0, 7.0).cpu().numpy()
    assert np.allclose(P32, want, rtol=0, atol=2e-4 * np.abs(want).max())


def test_pair_energy(orc):
    pts = synthetic.jittered_lattice((20, 21, 22), 1.0, 0.2, dtype=np.float64, seed=4)
    nhs, x, lists = setup(orc, pts, 2.5, np.float64, np.float64)
    e = torch.zeros(len(pts), dtype=torch.float64, device="cuda")
    pn.foreach_point_neighbor(sph.WendlandPairEnergy(e, 1.3, 2.5), x, x, nhs)
    oe = orc.pair_energy(pts, lists, 1.3, 2.5)
    assert np.allclose(e.cpu().numpy(), oe, rtol=1e-12, atol=0)


def rel_err(g, ref):
    scale = np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    return np.abs(g - ref) / scale


def lattice_mass(kernel, h, cutoff):
    """Mass so that rho = 1 on the unjittered unit lattice."""
    R = int(np.ceil(cutoff))
    k = np.arange(-R, R + 1, dtype=np.float64)
    off = np.stack(np.meshgrid(k, k, k, indexing="ij"), axis=-1).reshape(-1, 3)
    r = np.linalg.norm(off, axis=1)
    r = r[r <= cutoff]
    sigma = 21 / (16 * np.pi * h**3)
    q = r / h
    w = np.where(q < 2, sigma * (1 - q / 2) ** 4 * (2 * q + 1), 0.0)
    return 1.0 / w.sum()


def losses(which, x, nhs, m, kw):
    if which == 0:
        return sph.pair_energy(x, nhs, kw["eps"], kw["rc"]).sum()
    rho = sph.sph_density(x, m, nhs, kw["h"], kw["kernel"])
    if which == 1:
        return ((rho - kw["rho0"]) ** 2).sum()
    B = kw["rho0"] * kw["c"] ** 2 / kw["gamma"]
    p = B * ((rho / kw["rho0"]) ** kw["gamma"] - 1.0)
    P = p / rho**2
    a = sph.sph_pressure_acceleration(x, m, P, nhs, kw["h"], kw["kernel"])
    return (a * a).sum()


@pytest.mark.parametrize("shape,dt,tol", [((14, 15, 13), np.float64, 1e-9), ((40, 37), np.float64, 1e-9),
                                          ((14, 15, 13), np.float32, 5e-3)])
@pytest.mark.parametrize("which", [0, 1, 2])
def test_gradients_vs_dual_and_fd(orc, which, shape, dt, tol):
    pts = synthetic.jittered_lattice(shape, 1.0, 0.2, dtype=dt, seed=6)
    nd = len(shape)
    cutoff = 2.5
    if which == 2:
        kw = dict(kernel="cubic_spline", h=1.2, rc=2.4, eps=1.0, rho0=1.0, c=10.0, gamma=7.0)
        cutoff = 2.4
    else:
        kw = dict(kernel="wendland_c2", h=cutoff / 2, rc=cutoff, eps=1.3, rho0=1.0, c=10.0, gamma=7.0)
    nhs, x, lists = setup(orc, pts, dt(cutoff), dt, dt)
    m = torch.full((len(pts),), lattice_mass("wendland_c2", cutoff / 2, cutoff), dtype=x.dtype, device="cuda")
    xg = x.detach().clone().requires_grad_(True)
    # the differentiable functions need the grid built from xg's own storage
    pn.initialize(nhs, xg.detach(), xg.detach())
    L = losses(which, xg, nhs, m, kw)
    L.backward()
    g = xg.grad.t().cpu().numpy().astype(np.float64)
    assert np.abs(g.sum(axis=0)).max() <= (1e-10 if dt == np.float64 else 1e-4) * np.abs(g).sum(), \
        "translation invariance"
    n = len(pts)
    idx = (synthetic.uniform(100 + which, 12) * n).astype(int)
    ref = []
    got = []
    for a, k in enumerate(idx):
        dim = a % nd
        _, d = orc.loss_dual(which, pts.astype(np.float64), lists, k, dim, mass=m.cpu().numpy().astype(np.float64),
                             kernel=kw["kernel"], h=kw["h"], eps=kw["eps"], rc=kw["rc"], rho0=kw["rho0"],
                             c=kw["c"], gamma=kw["gamma"])
        ref.append(d)
        got.append(g[k, dim])
    ref, got = np.array(ref), np.array(got)
    assert rel_err(got, ref).max() < tol, (got, ref)


def test_fd_density_loss_large(orc):
    """64^3 = 262,144 Float64 points (jitter 0.2, cutoff 2.5): E and the density loss,
    gradients against central differences (step 1e-6) of the GPU forward pass on 64
    random coordinates; differences are taken per point so unchanged terms cancel."""
    pts = synthetic.jittered_lattice((64, 64, 64), 1.0, 0.2, dtype=np.float64, seed=1)
    cutoff = 2.5
    h = cutoff / 2
    lo, hi = pts.min(axis=0) - 3 * cutoff, pts.max(axis=0) + 3 * cutoff
    nhs = make_nhs(pts, cutoff, lo, hi)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    m = torch.full((len(pts),), lattice_mass("wendland_c2", h, cutoff), dtype=torch.float64, device="cuda")
    xg = x.detach().clone().requires_grad_(True)
    pn.initialize(nhs, xg.detach(), xg.detach())
    E = sph.pair_energy(xg, nhs, 1.0, cutoff).sum()
    gE = torch.autograd.grad(E, xg)[0].t().cpu().numpy()
    rho = sph.sph_density(xg, m, nhs, h, "wendland_c2")
    L = ((rho - 1.0) ** 2).sum()
    gL = torch.autograd.grad(L, xg)[0].t().cpu().numpy()
    step = 1e-6
    idx = (synthetic.uniform(77, 64) * len(pts)).astype(int)
    fdE, fdL, rE, rL = [], [], [], []
    xp = x.clone()
    for a, k in enumerate(idx):
        dim = a % 3
        vals = []
        for sgn in (1, -1):
            xp.copy_(x)
            xp[dim, k] += sgn * step
            pn.update(nhs, xp, xp)
            e = torch.zeros(len(pts), dtype=torch.float64, device="cuda")
            pn.foreach_point_neighbor(sph.WendlandPairEnergy(e, 1.0, cutoff, accumulate=False), xp, xp, nhs)
            r_ = torch.zeros(len(pts), dtype=torch.float64, device="cuda")
            pn.foreach_point_neighbor(sph.SPHDensity(r_, m, h, "wendland_c2", accumulate=False), xp, xp, nhs)
            vals.append((e, (r_ - 1.0) ** 2))
        fdE.append(float((vals[0][0] - vals[1][0]).sum()) / (2 * step))
        fdL.append(float((vals[0][1] - vals[1][1]).sum()) / (2 * step))
        rE.append(gE[k, dim])
        rL.append(gL[k, dim])
    pn.update(nhs, x, x)
    assert rel_err(np.array(rE), np.array(fdE)).max() < 1e-4
    assert rel_err(np.array(rL), np.array(fdL)).max() < 1e-4


def test_csr_lattice_and_list_density(orc):
    """Undeformed 24^3 lattice, r = 3: interior rows hold 123 neighbours, corner rows 29
    (integer points in a ball / closed octant of radius 3); the density over the stored
    lists equals the grid traversal's (summation order differs)."""
    pts = synthetic.jittered_lattice((24, 24, 24), 1.0, 0.0, dtype=np.float32)
    r = np.float32(3.0)
    nhs = make_nhs(pts, r, pts.min(axis=0) - r, pts.max(axis=0) + r)
    p = pn.PrecomputedNeighborhoodSearch(3, search_radius=r, n_points=len(pts), update_neighborhood_search=nhs)
    x = dev(pts)
    pn.initialize(p, x, x)
    lens = p.neighbor_lists.lengths().cpu().numpy()
    assert lens.max() == 123 and lens.min() == 29
    assert lens[12 + 24 * 12 + 24 * 24 * 12] == 123 and lens[0] == 29
    m = torch.full((len(pts),), 0.5, dtype=torch.float32, device="cuda")
    rl = torch.zeros(len(pts), dtype=torch.float32, device="cuda")
    rg = torch.zeros(len(pts), dtype=torch.float32, device="cuda")
    pn.foreach_point_neighbor(sph.SPHDensity(rl, m, 1.5, "wendland_c2"), x, x, pn.freeze_neighborhood_search(p))
    pn.foreach_point_neighbor(sph.SPHDensity(rg, m, 1.5, "wendland_c2"), x, x, nhs)
    assert torch.allclose(rl, rg, rtol=1e-5, atol=0)
