"""Batched search over many uneven clouds: equals per-cloud searches (and the oracle),
and an update + count step captured in a graph equals the eager step on every replay."""

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import batch, synthetic
from test_gpu_parity import dev, make_nhs, oracle_like

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def clouds(n_clouds, lo, hi, seed, jitter=0.25, dtype=np.float32):
    """Jittered lattices with log-uniform sizes in [lo, hi]; each in its own box."""
    u = synthetic.uniform(seed, n_clouds)
    sizes = np.round(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(int)
    out = []
    for c, s in enumerate(sizes):
        side = max(2, int(round(s ** (1 / 3))))
        shape = (side, side, max(1, int(round(s / side / side))))
        out.append(synthetic.jittered_lattice(shape, 1.0, jitter, dtype=dtype, seed=seed * 1000 + c))
    return out


def batched(pts_list, r, margin):
    cls = [pn.FullGridCellList(min_corner=p.min(axis=0) - margin, max_corner=p.max(axis=0) + margin,
                               search_radius=r) for p in pts_list]
    return batch.BatchedGridNeighborhoodSearch(3, search_radius=r, cell_lists=cls,
                                               n_points=[len(p) for p in pts_list])


def test_batch_equals_per_cloud_and_oracle(orc):
    r = np.float32(2.4)
    pts_list = clouds(40, 50, 3000, seed=3)
    nb = batched(pts_list, r, np.float32(3 * 2.4))
    allpts = np.concatenate(pts_list)
    y = dev(allpts)
    nb.update(y)
    counts = torch.zeros(len(allpts), dtype=torch.int32, device="cuda")
    nb.count_neighbors(counts)
    got = counts.cpu().numpy()
    offs = nb.point_offsets
    for c, p in enumerate(pts_list):
        nhs = make_nhs(p, r, p.min(axis=0) - np.float32(3 * 2.4), p.max(axis=0) + np.float32(3 * 2.4))
        x = dev(p)
        pn.initialize(nhs, x, x)
        cc = torch.zeros(len(p), dtype=torch.int32, device="cuda")
        pn.foreach_point_neighbor(pn.CountNeighbors(cc), x, x, nhs)
        assert np.array_equal(got[offs[c]:offs[c + 1]], cc.cpu().numpy()), c
        if c < 5:
            og = oracle_like(orc, nhs, np.float32)
            assert og.build(p)[0] == orc.OK
            assert np.array_equal(got[offs[c]:offs[c + 1]], og.traverse(p, p)[2])


def test_batch_bounds_error():
    r = np.float32(2.4)
    pts_list = clouds(6, 50, 500, seed=4)
    nb = batched(pts_list, r, np.float32(3 * 2.4))
    allpts = np.concatenate(pts_list)
    allpts[nb.point_offsets[3] + 2, 1] += np.float32(1000.0)
    with pytest.raises(pn.PointNeighborsError, match="outside the domain bounds"):
        nb.update(dev(allpts))


def test_batch_graph_replays_match_eager():
    r = np.float32(2.4)
    pts_list = clouds(30, 100, 4000, seed=5)
    nb = batched(pts_list, r, np.float32(3 * 2.4))
    allpts = np.concatenate(pts_list)
    n = len(allpts)
    y = dev(allpts)
    counts = torch.zeros(n, dtype=torch.int32, device="cuda")
    step = torch.zeros(1, dtype=torch.int64, device="cuda")

    def do_step():
        batch.displace_uniform(y, 0.05, 42, step)
        nb.update(y, check_errors=False)
        nb.count_neighbors(counts)

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):  # warm up: sizes every buffer before capture
            do_step()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        do_step()
    torch.cuda.synchronize()
    ref = batched(pts_list, r, np.float32(3 * 2.4))
    for rep in range(4):
        before = int(step.item())
        g.replay()
        torch.cuda.synchronize()
        assert int(step.item()) == before + 1
        nb.check()
        # eager recount of the same coordinates with a fresh batch search
        ref.update(y)
        eager = torch.zeros(n, dtype=torch.int32, device="cuda")
        ref.count_neighbors(eager)
        _ = (y * 2).sum().item()  # unrelated eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(counts, eager), rep
