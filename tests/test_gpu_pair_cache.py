"""Pair cache of the cell traversal: the first SPH traversal after a build records each
cell's accepted pairs, later ones replay them.  Results are identical to a traversal that
tests every candidate, and every build or in-place coordinate change invalidates the cache."""

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import sph, synthetic
from test_gpu_parity import dev, make_nhs

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def setup(seed=3):
    pts = synthetic.jittered_lattice((36, 37, 38), 0.01, 0.001, dtype=np.float32, seed=seed)
    h = np.float32(1.2) * np.float32(0.01)
    r = np.float32(2) * h
    nhs = make_nhs(pts, r, pts.min(axis=0) - 3 * r, pts.max(axis=0) + 3 * r)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    return nhs, x, float(h)


def step(nhs, x, h):
    n = x.shape[1]
    m = torch.full((n,), 1e-3, device="cuda")
    rho = torch.zeros(n, device="cuda")
    pn.foreach_point_neighbor(sph.SPHDensity(rho, m, h), x, x, nhs)
    P = (rho - rho.mean()) / (rho * rho)
    dv = torch.zeros((n, 3), device="cuda").t()
    pn.foreach_point_neighbor(sph.SPHPressureAcceleration(dv, m, P, h), x, x, nhs)
    e = torch.zeros(n, device="cuda")
    pn.foreach_point_neighbor(sph.WendlandPairEnergy(e, 1.0, float(nhs.search_radius)), x, x, nhs)
    return rho, dv.clone(), e


def both(nhs, x, h):
    nhs.set_tuning("pair_cache", 0)
    ref = step(nhs, x, h)
    nhs.set_tuning("pair_cache", 1)
    first = step(nhs, x, h)    # density records, pressure and energy replay
    second = step(nhs, x, h)   # everything replays
    return ref, first, second


def test_replay_identical_to_full_traversal():
    nhs, x, h = setup()
    ref, first, second = both(nhs, x, h)
    for a, b, c in zip(ref, first, second):
        assert torch.equal(a, b) and torch.equal(a, c)


def test_update_invalidates():
    nhs, x, h = setup()
    step(nhs, x, h)
    x.add_(torch.empty_like(x).uniform_(-5e-4, 5e-4))
    pn.update(nhs, x, x)
    ref, first, second = both(nhs, x, h)
    for a, b, c in zip(ref, first, second):
        assert torch.equal(a, b) and torch.equal(a, c)


def density_energy(nhs, x, h):
    n = x.shape[1]
    m = torch.full((n,), 1e-3, device="cuda")
    rho = torch.zeros(n, device="cuda")
    pn.foreach_point_neighbor(sph.SPHDensity(rho, m, h), x, x, nhs)
    e = torch.zeros(n, device="cuda")
    pn.foreach_point_neighbor(sph.WendlandPairEnergy(e, 1.0, float(nhs.search_radius)), x, x, nhs)
    return rho, e


def test_inplace_change_invalidates():
    """Coordinates changed without update!: the reference reads them at traversal time, so
    recorded pairs must not be replayed."""
    nhs, x, h = setup()
    step(nhs, x, h)
    x.add_(torch.empty_like(x).uniform_(-3e-3, 3e-3))   # changes accepted pairs
    got = density_energy(nhs, x, h)
    again = density_energy(nhs, x, h)
    nhs.set_tuning("pair_cache", 0)
    ref = density_energy(nhs, x, h)
    for a, b, c in zip(ref, got, again):
        assert torch.equal(a, b) and torch.equal(a, c)


def test_overflowing_cells_fall_back():
    nhs, x, h = setup()
    nhs.set_tuning("pair_cache_per_point", 12)   # far too small: most cells are not cached
    ref, first, second = both(nhs, x, h)
    for a, b, c in zip(ref, first, second):
        assert torch.equal(a, b) and torch.equal(a, c)


def test_knob_changes_keep_results():
    nhs, x, h = setup()
    ref = step(nhs, x, h)
    for k, v in (("chunk", 1), ("chunk", 0), ("algorithm", 1), ("algorithm", 0)):
        nhs.set_tuning(k, v)
        got = step(nhs, x, h)
        for a, b in zip(ref, got):
            assert torch.allclose(a, b, rtol=1e-5, atol=1e-6 * float(a.abs().max()))
