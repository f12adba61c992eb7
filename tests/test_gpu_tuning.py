"""Every autotuner choice gives the same results; the autotuner picks, applies and caches."""

import json

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import autotune, sph, synthetic
from test_gpu_parity import dev, make_nhs

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def setup():
    pts = synthetic.jittered_lattice((40, 41, 42), 0.01, 0.001, dtype=np.float32, seed=8)
    h = np.float32(1.2) * np.float32(0.01)
    r = np.float32(2) * h
    nhs = make_nhs(pts, r, pts.min(axis=0) - 3 * r, pts.max(axis=0) + 3 * r)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    return pts, h, nhs, x


def run(nhs, x, h):
    n = x.shape[1]
    c = torch.zeros(n, dtype=torch.int32, device="cuda")
    pn.foreach_point_neighbor(pn.CountNeighbors(c), x, x, nhs)
    m = torch.full((n,), 1e-3, device="cuda")
    rho = torch.zeros(n, device="cuda")
    pn.foreach_point_neighbor(sph.SPHDensity(rho, m, float(h)), x, x, nhs)
    return c.cpu().numpy(), rho.cpu().numpy()


@pytest.mark.parametrize("algorithm", [0, 1])
@pytest.mark.parametrize("blocks", [8, 1024, 16384])
@pytest.mark.parametrize("chunk", [-1, 0, 1])
def test_knobs_do_not_change_results(algorithm, blocks, chunk):
    pts, h, nhs, x = setup()
    c0, r0 = run(nhs, x, h)
    nhs.set_tuning("algorithm", algorithm)
    nhs.set_tuning("persist_blocks", blocks)
    nhs.set_tuning("chunk", chunk)
    c1, r1 = run(nhs, x, h)
    c2, r2 = run(nhs, x, h)
    assert np.array_equal(c0, c1)
    # the summation order depends on the algorithm and chunk, not on the run
    assert np.allclose(r0, r1, rtol=1e-5, atol=0)
    assert np.array_equal(r1, r2)


def test_atomic_order_same_sets():
    pts, h, nhs, x = setup()
    s0, i0 = nhs.cell_table()
    nhs.set_tuning("canonical_order", 0)
    pn.update(nhs, x, x)
    s1, i1 = nhs.cell_table()
    assert np.array_equal(s0, s1)
    for c in np.nonzero(np.diff(s0))[0][:3000]:
        assert sorted(i0[s0[c]:s0[c + 1]]) == sorted(i1[s1[c]:s1[c + 1]])
    c1, _ = run(nhs, x, h)
    nhs.set_tuning("canonical_order", 1)
    pn.update(nhs, x, x)
    c0, _ = run(nhs, x, h)
    assert np.array_equal(c0, c1)


def test_autotune_picks_and_caches(tmp_path):
    pts, h, nhs, x = setup()
    n = x.shape[1]
    c = torch.zeros(n, dtype=torch.int32, device="cuda")
    f = pn.CountNeighbors(c, accumulate=False)

    def step():
        pn.update(nhs, x, x, check_errors=False)
        pn.foreach_point_neighbor(f, x, x, nhs)

    cache = str(tmp_path / "tuning.json")
    key = autotune.signature(nhs, x, "count")
    space = {"persist_blocks": [1024, 4096], "chunk": [0, 1]}
    best, table = autotune.autotune(nhs, step, space=space, reps=2, warmup=1, cache=cache, key=key)
    assert len(table) == 4 and best in [e["params"] for e in table]
    data = json.load(open(cache))
    assert data[key]["params"] == best
    other = make_nhs(pts, nhs.search_radius, pts.min(axis=0) - 3 * nhs.search_radius,
                     pts.max(axis=0) + 3 * nhs.search_radius)
    assert autotune.load_tuning(other, cache, key)
    assert other._tuning["persist_blocks"] == best["persist_blocks"]


def test_launch_timing_hook():
    """pn_grid_launch_times: one positive duration per traversal kernel launched while
    "time_launches" is on, in launch order, cleared on read; none when it is off; results
    unchanged."""
    pts, h, nhs, x = setup()
    c0, rho0 = run(nhs, x, h)
    nhs.set_tuning("time_launches", 1)
    c1, rho1 = run(nhs, x, h)  # a count and a density traversal
    c2, rho2 = run(nhs, x, h)
    t = nhs.launch_times()
    assert len(t) == 4 and all(v > 0 for v in t)
    assert nhs.launch_times() == []
    nhs.set_tuning("time_launches", 0)
    run(nhs, x, h)
    assert nhs.launch_times() == []
    assert np.array_equal(c0, c1) and np.array_equal(c1, c2)
    assert np.array_equal(rho0, rho1) and np.array_equal(rho1, rho2)
    with pytest.raises(pn.ArgumentError):
        nhs.set_tuning("time_launches", 2)
