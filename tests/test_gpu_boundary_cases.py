"""Float32 clouds built to put many pairs on or near the boundary d2 = r^2: lattice ties,
near-ties a few ulps either side, coordinates far from the origin, 2D Pythagorean ties,
dense clustered cells (more than 64 queries per cell).  Counts, index sums and CSR lists
must equal the oracle's bit for bit with both traversal algorithms, and the SPH sums and
adjoints must repeat bit for bit."""

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import sph, synthetic
from test_gpu_parity import dev, gpu_counts, gpu_sums, make_nhs, oracle_like

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

F32 = np.float32


def clouds():
    rng = np.random.default_rng(11)
    lat = synthetic.jittered_lattice((24, 25, 26), 1.0, 0.0, dtype=F32)
    yield "lattice ties d2 = 9", lat, F32(3.0)
    # a few ulps (ulp(26) ~ 1.9e-6) around the ties
    near = (lat + rng.integers(-3, 4, size=lat.shape) * F32(2.0**-19)).astype(F32)
    yield "near ties", near, F32(3.0)
    far = (synthetic.jittered_lattice((20, 21, 22), 0.5, 0.0, dtype=F32) + F32(3000.0)).astype(F32)
    yield "far from the origin, ties", far, F32(1.5)
    farj = (far + rng.uniform(-0.01, 0.01, size=far.shape)).astype(F32)
    yield "far from the origin, jittered", farj, F32(1.5)
    yield "uniform", rng.uniform(0, 20, size=(9000, 3)).astype(F32), F32(1.7)
    centres = rng.uniform(0, 10, size=(6, 3))
    blobs = np.concatenate([c + rng.normal(0, 0.15, size=(700, 3)) for c in centres]).astype(F32)
    yield "dense clusters (over 64 queries per cell)", blobs, F32(0.45)
    yield "2D lattice, 3-4-5 ties", synthetic.jittered_lattice((90, 80), 1.0, 0.0, dtype=F32), F32(5.0)
    yield "2D uniform", rng.uniform(-5, 5, size=(6000, 2)).astype(F32), F32(0.3)
    sp = F32(0.01)
    yield ("WCSPH-like jittered lattice", synthetic.jittered_lattice((30, 31, 32), 0.01, 0.001, dtype=F32),
           F32(2) * (F32(1.2) * sp))
    # extreme scales: lattices with near-ties scaled to tiny and huge coordinates and radii
    # (r^2 near 1e-31 and terms near 1e30 in Float32)
    base = synthetic.jittered_lattice((12, 13, 14), 1.0, 0.0, dtype=np.float64)
    jit = rng.uniform(-0.02, 0.02, size=base.shape)
    for scale, label in ((1e-16, "tiny scale, r^2 below 2^-100"), (1e-14, "small scale, band test"),
                         (1e13, "large scale, band test"), (1e14, "huge scale, terms above 2^100")):
        yield label, ((base + jit) * scale).astype(F32), F32(3.0 * scale)


CASES = list(clouds())


@pytest.mark.parametrize("case", range(len(CASES)), ids=[c[0] for c in CASES])
def test_sets_match_oracle(orc, case):
    name, pts, r = CASES[case]
    nhs = make_nhs(pts, r, pts.min(axis=0) - r, pts.max(axis=0) + r, cap=4000)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    og = oracle_like(orc, nhs, F32)
    assert og.build(pts)[0] == orc.OK
    st, _, ocnt, osum, _ = og.traverse(pts, pts)
    assert st == orc.OK
    for algorithm in (0, 1):
        nhs.set_tuning("algorithm", algorithm)
        assert np.array_equal(gpu_counts(nhs, x, x), ocnt), (name, algorithm)
        assert np.array_equal(gpu_sums(nhs, x, x), osum), (name, algorithm)
    nhs.set_tuning("algorithm", 0)
    p = pn.PrecomputedNeighborhoodSearch(pts.shape[1], search_radius=r, n_points=len(pts),
                                         update_neighborhood_search=nhs, max_neighbors=4096)
    pn.initialize(p, x, x)
    rp, cols = p.neighbor_lists.to_host()
    st, _, (orp, ocols) = og.precomputed(pts, pts, 4096)
    assert st == orc.OK
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(cols, ocols), name


def physics(nhs, x, h, grad):
    n = x.shape[1]
    m = torch.full((n,), 1e-3, device="cuda")
    if grad:
        # x is the search's own (leaf) tensor
        x.requires_grad_(True)
        rho = sph.sph_density(x, m, nhs, h, "wendland_c2")
        P = (rho - 1.0) / rho**2
        a = sph.sph_pressure_acceleration(x, m, P, nhs, h, "wendland_c2")
        loss = (a * a).sum() + sph.pair_energy(x, nhs, 1e-3, 2 * h).sum()
        (g,) = torch.autograd.grad(loss, x)
        return [rho.detach(), a.detach(), g]
    rho = torch.zeros(n, device="cuda")
    pn.foreach_point_neighbor(sph.SPHDensity(rho, m, h), x, x, nhs)
    P = torch.rand(n, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    dv = sph._vec_like(x).zero_()
    pn.foreach_point_neighbor(sph.SPHPressureAcceleration(dv, m, P, h), x, x, nhs)
    return [rho, dv]


@pytest.mark.parametrize("grad", [False, True])
@pytest.mark.parametrize("case", [0, 1, 5, 8], ids=[CASES[i][0] for i in (0, 1, 5, 8)])
def test_functors_repeat_bit_identical(case, grad):
    name, pts, r = CASES[case]
    h = float(r) / 2
    nhs = make_nhs(pts, r, pts.min(axis=0) - r, pts.max(axis=0) + r, cap=4000)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    first = physics(nhs, x, h, grad)
    again = physics(nhs, x, h, grad)
    for a, b in zip(first, again):
        assert torch.equal(a, b), name
