"""The benchmark configurations at their configured sizes, checked against the oracle and
closed forms, with the method of the reference's own all-variants test
(test/neighborhood_search.jl:186-343: every search against a brute force on the same cloud).

* 1M-point WCSPH cloud (100^3, Float32): density and pressure sums of every point against
  the oracle's double sums over the same neighbour sets, |gpu - oracle| <= 1e-5 * sum|term|.
* 512 uneven clouds (1,000-65,536 points, Float32): the batched count equals a per-cloud
  search for every cloud and the oracle on 9 clouds (the largest included); 100 replays of
  the captured displace + update! + count step equal eager steps.
* 256^3 = 16.8M points (Float32): counts of every point equal the threaded oracle; density and
  pressure sums on sampled rows at 1e-5; the adjoint of sum |a|^2 on 64 coordinates against
  forward-mode duals and central differences (step 1e-6) through the oracle in double,
  evaluated on the ball of radius 4r around the moved point (which holds every term the
  coordinate enters) at 1e-4.
* 128^3 undeformed lattice, r = 3: every CSR row length equals the closed form (123 interior),
  sampled rows are sorted and equal the oracle's.
"""

import sys

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import batch, sph, synthetic
from test_gpu_batch import batched
from test_gpu_parity import dev, make_nhs, oracle_like
from test_gpu_physics import rel_err

import bench

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def wcsph_sums(orc, pts, h, r, rows, mass, threads):
    """GPU density / pressure of the cloud and the oracle's sums on `rows`."""
    lo, hi = pts.min(axis=0) - np.float32(3) * r, pts.max(axis=0) + np.float32(3) * r
    nhs = make_nhs(pts, r, lo, hi)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    n = len(pts)
    og = oracle_like(orc, nhs, np.float32)
    st, counts, _, _ = og.threaded_count_f32(pts, threads)
    assert st == orc.OK
    c = torch.zeros(n, dtype=torch.int32, device="cuda")
    pn.foreach_point_neighbor(pn.CountNeighbors(c), x, x, nhs)
    assert np.array_equal(c.cpu().numpy(), counts)
    m = np.float32(mass)
    mt = torch.full((n,), float(m), dtype=torch.float32, device="cuda")
    rho = torch.zeros(n, dtype=torch.float32, device="cuda")
    pn.foreach_point_neighbor(sph.SPHDensity(rho, mt, float(h), "cubic_spline"), x, x, nhs)
    st, orho, rs, _, _ = og.threaded_sph_rows_f32(pts, rows, threads, float(m), float(h), "cubic_spline")
    assert st == orc.OK
    err = np.abs(rho.cpu().numpy()[rows].astype(np.float64) - orho)
    assert (err <= 1e-5 * rs).all(), (err / rs).max()
    # pressure: P from the GPU density, the same Float32 values on both sides
    P = sph.tait_pressure_term(rho, bench.RHO0, bench.SOUND, bench.GAMMA)
    Pn = P.cpu().numpy().astype(np.float64)
    dv = sph._vec_like(x).zero_()
    pn.foreach_point_neighbor(sph.SPHPressureAcceleration(dv, mt, P, float(h), "cubic_spline"), x, x, nhs)
    st, _, _, odv, ds = og.threaded_sph_rows_f32(pts, rows, threads, float(m), float(h), "cubic_spline", P=Pn)
    assert st == orc.OK
    err = np.abs(dv.t().cpu().numpy()[rows].astype(np.float64) - odv).max(axis=1)
    assert (err <= 1e-5 * np.maximum(ds, 1e-300)).all(), (err / ds).max()
    return nhs, x, og, counts


def test_wcsph_1m_sums_all_points(orc):
    """1,000,000 points (100^3, spacing 0.01, jitter 0.001, Float32), h = 1.2 spacing, r = 2h:
    density and pressure sums of every point."""
    say("C2 1M WCSPH sums")
    pts = synthetic.jittered_lattice(bench.SHAPE, bench.SPACING, bench.JITTER, dtype=np.float32, seed=1)
    h = np.float32(1.2) * np.float32(bench.SPACING)
    r = np.float32(2.0) * h
    rows = np.arange(len(pts))
    wcsph_sums(orc, pts, h, r, rows, bench.RHO0 * bench.SPACING**3, orc.host_threads())


def test_batch_512_clouds_configured_size(orc):
    """All 512 clouds of the batching configuration (sizes log-uniform 1,000-65,536, r = 2.4)."""
    say("C4 512 clouds")
    clouds = bench.batch_clouds()
    assert len(clouds) == 512
    sizes = np.array([len(p) for p in clouds])
    r = np.float32(bench.BATCH_R)
    margin = np.float32(3 * bench.BATCH_R)
    nb = batched(clouds, r, margin)
    allpts = np.concatenate(clouds)
    y = dev(allpts)
    nb.update(y)
    counts = torch.zeros(len(allpts), dtype=torch.int32, device="cuda")
    nb.count_neighbors(counts)
    got = counts.cpu().numpy()
    offs = nb.point_offsets
    largest = int(np.argmax(sizes))
    order = np.argsort(sizes)
    checked = sorted({largest, int(order[0]), int(order[1]), int(order[len(order) // 2]),
                      *range(0, 512, 100)})
    assert len(checked) >= 8 and sizes[largest] > 60000
    for c, p in enumerate(clouds):
        lo, hi = p.min(axis=0) - margin, p.max(axis=0) + margin
        nhs = make_nhs(p, r, lo, hi)
        x = dev(p)
        pn.initialize(nhs, x, x)
        cc = torch.zeros(len(p), dtype=torch.int32, device="cuda")
        pn.foreach_point_neighbor(pn.CountNeighbors(cc), x, x, nhs)
        assert np.array_equal(got[offs[c]:offs[c + 1]], cc.cpu().numpy()), c
        if c in checked:
            og = oracle_like(orc, nhs, np.float32)
            st, want, _, _ = og.threaded_count_f32(p, orc.host_threads())
            assert st == orc.OK
            assert np.array_equal(got[offs[c]:offs[c + 1]], want), c
    say(f"  batched == per-cloud on 512 clouds, == oracle on {len(checked)} (largest {sizes[largest]})")


def test_batch_512_clouds_100_graph_replays():
    """The captured step (displace U[+-0.05], update!, count) replayed 100 times equals the
    same step run eagerly: coordinates bit-identical, counts equal a fresh search's."""
    say("C4 100 graph replays")
    clouds = bench.batch_clouds()
    r = np.float32(bench.BATCH_R)
    margin = np.float32(3 * bench.BATCH_R)
    nb = batched(clouds, r, margin)
    allpts = np.concatenate(clouds)
    n = len(allpts)
    y = dev(allpts)
    ye = dev(allpts)
    counts = torch.zeros(n, dtype=torch.int32, device="cuda")
    step = torch.zeros(1, dtype=torch.int64, device="cuda")
    step_e = torch.zeros(1, dtype=torch.int64, device="cuda")

    def do_step():
        batch.displace_uniform(y, 0.05, 11, step)
        nb.update(y, check_errors=False)
        nb.count_neighbors(counts)

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):  # sizes every buffer before capture
            do_step()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    for _ in range(2):
        batch.displace_uniform(ye, 0.05, 11, step_e)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        do_step()
    torch.cuda.synchronize()
    ref = batched(clouds, r, margin)
    eager = torch.zeros(n, dtype=torch.int32, device="cuda")
    for rep in range(100):
        g.replay()
        batch.displace_uniform(ye, 0.05, 11, step_e)
        ref.update(ye)
        ref.count_neighbors(eager)
        torch.cuda.synchronize()
        nb.check()
        assert int(step.item()) == int(step_e.item()), rep
        assert torch.equal(y, ye), rep
        assert torch.equal(counts, eager), rep
    say("  100 replays equal eager steps")


def large_cloud():
    return synthetic.jittered_lattice((256, 256, 256), 1.0, 0.1, dtype=np.float32, seed=1)


def test_large_16m_counts_and_sampled_sums(orc):
    """256^3 = 16,777,216 points (jitter 0.1, Float32), h = 1.2, r = 2h: every count equals the
    threaded oracle; density and pressure sums on 20,000 sampled rows."""
    say("C5 16.8M counts and sums")
    pts = large_cloud()
    h = np.float32(1.2)
    r = np.float32(2) * h
    rows = np.unique((synthetic.uniform(91, 20000) * len(pts)).astype(np.int64))
    wcsph_sums(orc, pts, h, r, rows, 1000.0, orc.host_threads())


def local_patch(pts, shape, k, radius):
    """Indices of the lattice points within `radius` of point k (x-fastest lattice index)."""
    nx, ny, nz = shape
    ix, iy, iz = k % nx, (k // nx) % ny, k // (nx * ny)
    R = int(np.ceil(radius)) + 1
    xs = np.arange(max(0, ix - R), min(nx, ix + R + 1))
    ys = np.arange(max(0, iy - R), min(ny, iy + R + 1))
    zs = np.arange(max(0, iz - R), min(nz, iz + R + 1))
    idx = (xs[None, None, :] + nx * ys[None, :, None] + nx * ny * zs[:, None, None]).reshape(-1)
    d = np.linalg.norm(pts[idx].astype(np.float64) - pts[k].astype(np.float64), axis=1)
    return idx[d <= radius]


def test_large_16m_adjoint_local_checks(orc):
    """Adjoint of L = sum |a|^2 (density -> Tait EOS -> pressure acceleration) on the 16.8M cloud,
    recomputing the traversal: 64 coordinates against forward-mode duals and central
    differences through the oracle in double.  Perturbing x_p changes rho only within r of p, a
    only within 2r, so the terms of L that depend on x_p are those of points within 2r, whose
    densities need points within 4r: the oracle evaluates L on that ball with the Float32
    neighbour sets."""
    say("C5 16.8M adjoint")
    shape = (256, 256, 256)
    pts = large_cloud()
    n = len(pts)
    h = np.float32(1.2)
    r = np.float32(2) * h
    lo, hi = pts.min(axis=0) - np.float32(3) * r, pts.max(axis=0) + np.float32(3) * r
    nhs = make_nhs(pts, r, lo, hi)
    x = dev(pts)
    pn.initialize(nhs, x, x)
    mass = torch.full((n,), 1000.0, dtype=torch.float32, device="cuda")
    B = bench.RHO0 * bench.SOUND**2 / bench.GAMMA
    xg = x.detach().requires_grad_(True)
    rho = sph.sph_density(xg, mass, nhs, float(h))
    p = B * ((rho / bench.RHO0) ** bench.GAMMA - 1.0)
    a = sph.sph_pressure_acceleration(xg, mass, p / (rho * rho), nhs, float(h))
    (g,) = torch.autograd.grad((a * a).sum(), xg)
    g = g.t().cpu().numpy().astype(np.float64)
    assert np.isfinite(g).all()
    idx = (synthetic.uniform(123, 64) * n).astype(np.int64)
    got, dual, fd = [], [], []
    step = 1e-6
    for t, k in enumerate(idx):
        dim = t % 3
        patch = local_patch(pts, shape, int(k), 4 * float(r) + 0.5)
        sub = pts[patch]
        kk = int(np.nonzero(patch == k)[0][0])
        pl, ph = sub.min(axis=0) - np.float32(3) * r, sub.max(axis=0) + np.float32(3) * r
        sg = make_nhs(sub, r, pl, ph)
        og = oracle_like(orc, sg, np.float32)
        assert og.build(sub)[0] == orc.OK
        st, _, lists = og.precomputed(sub, sub, 100000)
        assert st == orc.OK
        kw = dict(mass=np.full(len(sub), 1000.0), kernel="cubic_spline", h=float(h), rho0=bench.RHO0,
                  c=bench.SOUND, gamma=bench.GAMMA)
        x64 = sub.astype(np.float64)
        _, d = orc.loss_dual(2, x64, lists, kk, dim, **kw)
        vals = []
        for sgn in (1, -1):
            xp = x64.copy()
            xp[kk, dim] += sgn * step
            vals.append(orc.loss_dual(2, xp, lists, kk, dim, **kw)[0])
        got.append(g[k, dim])
        dual.append(d)
        fd.append((vals[0] - vals[1]) / (2 * step))
    got, dual, fd = map(np.array, (got, dual, fd))
    assert rel_err(dual, fd).max() < 1e-4, "oracle duals vs central differences"
    assert rel_err(got, dual).max() < 1e-4, (got, dual)
    assert rel_err(got, fd).max() < 1e-4
    say(f"  64 coordinates: max rel err {rel_err(got, dual).max():.2e}")


def test_csr_128_lattice_closed_form_and_oracle_rows(orc):
    """PrecomputedNeighborhoodSearch on the undeformed 128^3 lattice (spacing 1), r = 3: every
    row length is the number of lattice offsets of norm <= 3 that stay in the box (123 in the
    interior); 4,000 sampled rows are ascending and equal the oracle's lists."""
    say("C5 128^3 CSR")
    side = 128
    pts = synthetic.jittered_lattice((side, side, side), 1.0, 0.0, dtype=np.float32, seed=1)
    n = len(pts)
    r = np.float32(3.0)
    nhs = make_nhs(pts, r, pts.min(axis=0) - r, pts.max(axis=0) + r)
    pre = pn.PrecomputedNeighborhoodSearch(3, search_radius=r, n_points=n, update_neighborhood_search=nhs)
    x = dev(pts)
    pn.initialize(pre, x, x)
    nl = pre.neighbor_lists
    rp = nl.row_ptr.cpu().numpy().astype(np.int64)
    cols = nl.col_idx.cpu().numpy()
    lens = np.diff(rp)
    # closed form: offsets (a, b, c) with a^2 + b^2 + c^2 <= 9 that keep the point in the box
    k = np.arange(-3, 4)
    off = np.stack(np.meshgrid(k, k, k, indexing="ij"), axis=-1).reshape(-1, 3)
    off = off[(off**2).sum(axis=1) <= 9]
    assert len(off) == 123
    i = np.arange(side)
    want = np.zeros((side, side, side), dtype=np.int64)  # [z, y, x]
    for a, b, c in off:
        fx = ((i + a >= 0) & (i + a < side)).astype(np.int64)
        fy = ((i + b >= 0) & (i + b < side)).astype(np.int64)
        fz = ((i + c >= 0) & (i + c < side)).astype(np.int64)
        want += fz[:, None, None] * fy[None, :, None] * fx[None, None, :]
    assert np.array_equal(lens, want.reshape(-1))
    assert lens.max() == 123 and lens.min() == 29
    rows = np.unique((synthetic.uniform(55, 4000) * n).astype(np.int64))
    og = oracle_like(orc, nhs, np.float32)
    assert og.build(pts)[0] == orc.OK
    st, _, ocnt, _, (orp, ocols) = og.traverse(pts, pts, points=rows, lists=True)
    assert st == orc.OK
    for t, row in enumerate(rows):
        mine = cols[rp[row]:rp[row + 1]]
        assert (np.diff(mine) > 0).all(), row
        assert np.array_equal(mine, np.sort(ocols[orp[t]:orp[t + 1]])), row
