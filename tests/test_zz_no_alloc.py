"""pn_count_neighbors_noalloc (the allocation-free build + count, scratch in a per-GPU static
device array; no counterpart in the reference) against the allocating entry points pn_grid_create / pn_grid_update /
pn_count_neighbors through the C-ABI: two calls in a row, the second with a smaller cloud
(the scratch keeps the first call's contents), plus the size limit and the reference's
errors."""

import ctypes

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import synthetic
from pointneighbors._lib import load
from pointneighbors.util import dtype_code, promote, scalar_dtype

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def noalloc_fn():
    return load().pn_count_neighbors_noalloc


def geometry(pts, r):
    cl = pn.FullGridCellList(min_corner=pts.min(axis=0) - 3 * r, max_corner=pts.max(axis=0) + 3 * r,
                             search_radius=r)
    return cl


def call_noalloc(pts, r, cl, max_ppc=100):
    fn = noalloc_fn()
    nd = pts.shape[1]
    x = synthetic.to_device_coords(pts)
    counts = torch.full((len(pts),), -1, dtype=torch.int32, device="cuda")
    # the types pn_grid_create receives from the host mirror
    code = dtype_code(pts.dtype)
    rcode = dtype_code(scalar_dtype(r))
    cell = dtype_code(promote(pts.dtype, cl.corner_dtype))
    mc = (ctypes.c_double * nd)(*[float(v) for v in cl.min_corner])
    nc = (ctypes.c_int64 * nd)(*[int(v) for v in cl.n_cells])
    st = fn(nd, code, rcode, cell, mc, nc, float(r), int(max_ppc), ctypes.c_void_p(x.data_ptr()), len(pts),
            ctypes.c_void_p(counts.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return st, counts.cpu().numpy()


def reference_counts(pts, r, cl):
    nhs = pn.GridNeighborhoodSearch(pts.shape[1], search_radius=r, n_points=len(pts), cell_list=cl)
    x = synthetic.to_device_coords(pts)
    pn.initialize(nhs, x, x)
    c = torch.zeros(len(pts), dtype=torch.int32, device="cuda")
    pn.foreach_point_neighbor(pn.CountNeighbors(c), x, x, nhs)
    return c.cpu().numpy()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_noalloc_equals_allocating_path_twice(dt):
    r = dt(2.0) * (dt(1.2) * dt(0.01))
    big = synthetic.jittered_lattice((60, 62, 58), 0.01, 0.001, dtype=dt, seed=21)
    small = synthetic.jittered_lattice((17, 11, 23), 0.01, 0.002, dtype=dt, seed=22)
    for pts in (big, small):
        cl = geometry(pts, r)
        st, got = call_noalloc(pts, r, cl)
        assert st == 0, pn.load().pn_last_error()
        assert np.array_equal(got, reference_counts(pts, r, cl))


def test_noalloc_2d_and_errors():
    r = 2.5
    pts = synthetic.jittered_lattice((90, 70), 1.0, 0.3, dtype=np.float64, seed=23)
    cl = geometry(pts, r)
    st, got = call_noalloc(pts, r, cl)
    assert st == 0
    assert np.array_equal(got, reference_counts(pts, r, cl))
    # a point outside the cell list: the reference's bounds error
    bad = pts.copy()
    bad[7] = cl.max_corner + 50.0
    st, _ = call_noalloc(bad, r, cl)
    assert st == 2 and "outside the domain bounds" in pn.load().pn_last_error().decode()
    # more points in a cell than max_points_per_cell
    st, _ = call_noalloc(pts, r, cl, max_ppc=3)
    assert st == 3 and "cell list is full" in pn.load().pn_last_error().decode()
    # beyond the scratch array: a clear error, nothing launched
    huge = synthetic.jittered_lattice((160, 160, 160), 1.0, 0.1, dtype=np.float32, seed=1)
    st, _ = call_noalloc(huge, np.float32(2.4), geometry(huge, np.float32(2.4)))
    assert st == 8 and "scratch" in pn.load().pn_last_error().decode()
    assert load().pn_noalloc_scratch_bytes() == 64 << 20
