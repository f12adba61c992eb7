"""TLSPH deformation gradient on the GPU (pn_list_tlsph_deformation_grad over the stored lists of
a PrecomputedNeighborhoodSearch of the initial coordinates): equal to the oracle's sum within
the summation-order tolerance (Float64 1e-11, Float32 1e-5 of the largest entry), and the
closed forms hold on the device: the correction matrix of the undeformed body is symmetric, and
with it an affine map gives its matrix at every point."""

import math

import numpy as np
import pytest

import pointneighbors as pn
from pointneighbors import sph, synthetic
from test_gpu_parity import dev

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def setup(shape, dt, h, seed=9):
    X = synthetic.jittered_lattice(shape, 1.0, 0.15, dtype=dt, seed=seed)
    nd = len(shape)
    p = pn.PrecomputedNeighborhoodSearch(nd, search_radius=dt(2 * h), n_points=len(X), max_neighbors=4096)
    Xd = dev(X)
    pn.initialize(p, Xd, Xd)
    rp, cols = p.neighbor_lists.to_host()
    V = (1.0 + 0.1 * synthetic.uniform(5, len(X))).astype(dt)
    return X, Xd, p, (rp.astype(np.int64), cols.astype(np.int32)), V


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
@pytest.mark.parametrize("kernel", ["wendland_c2", "cubic_spline"])
@pytest.mark.parametrize("shape", [(11, 10, 9), (30, 27)])
def test_matches_oracle(orc, shape, kernel, dt, tol):
    h = 1.3
    X, Xd, p, lists, V = setup(shape, dt, h)
    rng = np.random.default_rng(1)
    x = (X + 0.1 * rng.standard_normal(X.shape)).astype(dt)
    Vd = torch.as_tensor(V, device="cuda")
    M = sph.tlsph_deformation_gradient(p, Xd, Xd, Vd, h, kernel)
    oM = orc.tlsph_grad(X, X, lists, V, h, kernel)
    assert np.abs(M.cpu().numpy() - oM).max() <= tol * np.abs(oM).max()
    L = torch.linalg.inv(M)
    # the kernel's closed-form inverse equals the batched LU inverse
    Lk = sph.tlsph_correction_matrix(p, Xd, Vd, h, kernel)
    assert np.abs(Lk.cpu().numpy() - L.cpu().numpy()).max() <= 100 * tol * np.abs(L.cpu().numpy()).max()
    F = sph.tlsph_deformation_gradient(p, Xd, dev(x), Vd, h, kernel, correction=L)
    oF = orc.tlsph_grad(X, x, lists, V, h, kernel, L=L.cpu().numpy().astype(np.float64))
    assert np.abs(F.cpu().numpy() - oF).max() <= 10 * tol * np.abs(oF).max()


@pytest.mark.parametrize("dt,atol", [(np.float64, 1e-9), (np.float32, 2e-4)])
@pytest.mark.parametrize("shape", [(9, 8, 7), (16, 15)])
def test_affine_closed_form(shape, dt, atol):
    h = 1.2
    X, Xd, p, _, V = setup(shape, dt, h, seed=3)
    nd = len(shape)
    Vd = torch.as_tensor(V, device="cuda")
    L = sph.tlsph_correction_matrix(p, Xd, Vd, h)
    M = torch.linalg.inv(L).cpu().numpy()
    assert np.allclose(M, M.transpose(0, 2, 1), rtol=0, atol=10 * atol * np.abs(M).max())
    rng = np.random.default_rng(7)
    A = np.eye(nd) + 0.25 * rng.standard_normal((nd, nd))
    th = 0.4
    R = np.eye(nd)
    R[:2, :2] = [[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]]
    for T in (A, R, np.eye(nd)):
        x = (X @ T.T + 1.5).astype(dt)
        F = sph.tlsph_deformation_gradient(p, Xd, dev(x), Vd, h, correction=L).cpu().numpy()
        assert np.allclose(F, T, rtol=0, atol=atol), np.abs(F - T).max()


def test_row_subset():
    h = 1.2
    X, Xd, p, _, V = setup((8, 9, 10), np.float64, h)
    Vd = torch.as_tensor(V, device="cuda")
    x = dev(X * 1.1)
    full = sph.tlsph_deformation_gradient(p, Xd, x, Vd, h)
    sub = sph.tlsph_deformation_gradient(p, Xd, x, Vd, h, points=range(100, 300))
    assert torch.equal(sub[100:300], full[100:300])
