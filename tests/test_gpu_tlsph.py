"""TLSPH on the GPU over the stored lists of a PrecomputedNeighborhoodSearch of the initial
coordinates.  Deformation gradient (pn_list_tlsph_deformation_grad): equal to the oracle's sum
within the summation-order tolerance (Float64 1e-11, Float32 1e-5 of the largest entry), and the
closed forms hold on the device: the correction matrix of the undeformed body is symmetric, and
with it an affine map gives its matrix at every point.  Interaction (pn_list_tlsph_interaction)
and corrected stress (pn_tlsph_pk1_rho2): equal to the oracle on the same inputs (Float64 1e-10,
Float32 2e-5 of the largest entry), momentum conserved, rows and accumulation as selected."""

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


def materials(n, dt, seed=4):
    rng = np.random.default_rng(seed)
    m = (0.1 * (1.0 + 0.3 * rng.random(n))).astype(dt)
    rho = (1000.0 * (1.0 + 0.05 * rng.random(n))).astype(dt)
    return m, rho


def cuda(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 1e-6)])
@pytest.mark.parametrize("nd", [2, 3])
def test_pk1_matches_oracle(orc, nd, dt, tol):
    n = 777
    rng = np.random.default_rng(nd)
    F = (np.eye(nd) + 0.3 * rng.standard_normal((n, nd, nd))).astype(dt)
    L = (np.eye(nd) + 0.1 * rng.standard_normal((n, nd, nd))).astype(dt)
    rho = (1000.0 * (1.0 + 0.1 * rng.random(n))).astype(dt)
    E, nu = 1.4e6, 0.4
    lam, mu = sph.lame_parameters(E, nu)
    for corr in (None, L):
        got = sph.tlsph_pk1_corrected(cuda(F), cuda(rho), E, nu, correction=None if corr is None else cuda(corr))
        want = orc.tlsph_pk1_rho2(F, rho, lam, mu, L=corr)
        assert np.abs(got.cpu().numpy() - want).max() <= tol * np.abs(want).max()


@pytest.mark.parametrize("penalty", [0.0, 0.1])
@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-10), (np.float32, 2e-5)])
@pytest.mark.parametrize("kernel", ["wendland_c2", "cubic_spline"])
@pytest.mark.parametrize("shape", [(11, 10, 9), (30, 27)])
def test_interaction_matches_oracle(orc, shape, kernel, dt, tol, penalty):
    """Random per-point stresses and deformation gradients (no cancellation hides a term)."""
    h = 1.3
    X, Xd, p, lists, _ = setup(shape, dt, h)
    n, nd = X.shape
    rng = np.random.default_rng(12)
    x = (X + 0.1 * rng.standard_normal(X.shape)).astype(dt)
    m, rho = materials(n, dt)
    Pr = (1e-3 * rng.standard_normal((n, nd, nd))).astype(dt)
    F = (np.eye(nd) + 0.2 * rng.standard_normal((n, nd, nd))).astype(dt)
    E = 1e4  # penalty and stress terms of one order
    dv = sph.tlsph_interaction(p, Xd, dev(x), cuda(m), cuda(rho), cuda(Pr), h, kernel, F=cuda(F),
                               young_modulus=E, penalty_alpha=penalty)
    want = orc.tlsph_force(X, x, lists, m, rho, Pr, h, kernel, F=F, young_modulus=E, penalty_alpha=penalty)
    got = dv.t().cpu().numpy()
    assert np.abs(got - want).max() <= tol * np.abs(want).max()


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-9), (np.float32, 1e-4)])
def test_tlsph_step_matches_oracle(orc, dt, tol):
    """Correction matrix -> F -> corrected stress -> interaction with the penalty force, the
    benchmark's material (rho 1000, E 1.4e6, nu 0.4, alpha 0.1, m 0.1), all on the device,
    against the oracle's pipeline in Float64."""
    h = 1.5
    X, Xd, p, lists, _ = setup((12, 11, 10), dt, h, seed=2)
    n, nd = X.shape
    x = (X + 0.05 * np.sin(0.7 * X[:, ::-1])).astype(dt)
    rho = np.full(n, 1000.0, dtype=dt)
    m = np.full(n, 0.1, dtype=dt)
    V = (m / rho).astype(dt)
    E, nu, alpha = 1.4e6, 0.4, 0.1
    lam, mu = sph.lame_parameters(E, nu)
    Vd, xd = cuda(V), dev(x)
    L = sph.tlsph_correction_matrix(p, Xd, Vd, h)
    F = sph.tlsph_deformation_gradient(p, Xd, xd, Vd, h, correction=L)
    Pr = sph.tlsph_pk1_corrected(F, cuda(rho), E, nu, correction=L)
    dv = sph.tlsph_interaction(p, Xd, xd, cuda(m), cuda(rho), Pr, h, F=F, young_modulus=E, penalty_alpha=alpha)
    X64, x64 = X.astype(np.float64), x.astype(np.float64)
    oL = np.linalg.inv(orc.tlsph_grad(X64, X64, lists, V, h))
    oF = orc.tlsph_grad(X64, x64, lists, V, h, L=oL)
    oP = orc.tlsph_pk1_rho2(oF, rho, lam, mu, L=oL)
    want = orc.tlsph_force(X64, x64, lists, m, rho, oP, h, F=oF, young_modulus=E, penalty_alpha=alpha)
    got = dv.t().cpu().numpy()
    assert np.abs(got - want).max() <= tol * np.abs(want).max()


def test_interaction_momentum_rows_and_accumulate():
    h = 1.2
    X, Xd, p, _, _ = setup((20, 19, 18), np.float64, h, seed=6)
    n, nd = X.shape
    rng = np.random.default_rng(3)
    x = dev(X + 0.1 * rng.standard_normal(X.shape))
    m, rho = materials(n, np.float64)
    Pr = cuda(rng.standard_normal((n, nd, nd)))
    F = cuda(np.eye(nd) + 0.2 * rng.standard_normal((n, nd, nd)))
    md, rd = cuda(m), cuda(rho)
    kw = dict(F=F, young_modulus=50.0, penalty_alpha=0.1)
    dv = sph.tlsph_interaction(p, Xd, x, md, rd, Pr, h, **kw)
    mom = (torch.as_tensor(m, device="cuda")[None, :] * dv).sum(dim=1)
    assert mom.abs().max().item() <= 1e-11 * (md[None, :] * dv).abs().sum().item()
    base = torch.randn_like(dv.t()).t()
    acc = sph.tlsph_interaction(p, Xd, x, md, rd, Pr, h, dv=base.clone(), **kw)
    assert torch.allclose(acc, base + dv, rtol=1e-14, atol=1e-14 * dv.abs().max().item())
    sub = sph.tlsph_interaction(p, Xd, x, md, rd, Pr, h, points=range(500, 1500), **kw)
    assert torch.equal(sub[:, 500:1500], dv[:, 500:1500])
    assert not sub[:, :500].any() and not sub[:, 1500:].any()
    # repeat calls give the same bits
    assert torch.equal(sph.tlsph_interaction(p, Xd, x, md, rd, Pr, h, **kw), dv)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(10, 9, 8), (25, 24)])
def test_packed_records_same_bits(shape, dt):
    """The penalty interaction over packed per-point records equals the field-by-field gathers
    bit for bit, also for a row subset and accumulation."""
    h = 1.2
    X, Xd, p, _, _ = setup(shape, dt, h, seed=8)
    n, nd = X.shape
    rng = np.random.default_rng(5)
    x = dev((X + 0.1 * rng.standard_normal(X.shape)).astype(dt))
    m, rho = materials(n, dt)
    Pr = cuda((1e-3 * rng.standard_normal((n, nd, nd))).astype(dt))
    F = cuda((np.eye(nd) + 0.2 * rng.standard_normal((n, nd, nd))).astype(dt))
    kw = dict(F=F, young_modulus=1e4, penalty_alpha=0.1)
    args = (p, Xd, x, cuda(m), cuda(rho), Pr, h)
    a = sph.tlsph_interaction(*args, **kw)
    b = sph.tlsph_interaction(*args, _packed=False, **kw)
    assert torch.equal(a, b)
    base = torch.randn_like(a.t()).t()
    sa = sph.tlsph_interaction(*args, points=[3, 1, 200, 7], dv=base.clone(), **kw)
    sb = sph.tlsph_interaction(*args, points=[3, 1, 200, 7], dv=base.clone(), _packed=False, **kw)
    assert torch.equal(sa, sb)


def test_interaction_argument_errors():
    h = 1.2
    X, Xd, p, _, _ = setup((5, 5, 5), np.float32, h)
    n, nd = X.shape
    m, rho = materials(n, np.float32)
    Pr = torch.zeros((n, nd, nd), dtype=torch.float32, device="cuda")
    with pytest.raises(pn.ArgumentError):
        sph.tlsph_interaction(p, Xd, Xd, cuda(m), cuda(rho), Pr, h, penalty_alpha=0.1)
    with pytest.raises(pn.ArgumentError):
        sph.tlsph_interaction(p, Xd, Xd, cuda(m), cuda(rho), Pr[:, :2], h)
    with pytest.raises(pn.ArgumentError):
        sph.tlsph_interaction(p, Xd, Xd, cuda(m), cuda(rho), Pr.double(), h)
    # a workspace off 16-byte alignment is refused before any launch (C-ABI status)
    import ctypes
    from pointneighbors._lib import PN_ERR_ARGUMENT, load
    nl = p.neighbor_lists
    ws = torch.empty(n * 32 + 8, dtype=torch.int32, device="cuda")
    md, rd, dv = cuda(m), cuda(rho), torch.zeros((n, nd), dtype=torch.float32, device="cuda")
    Pc = Pr.contiguous()
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    st = load().pn_list_tlsph_interaction(vp(nl.row_ptr), vp(nl.col_idx), nl.n, 0, nd, 0, 1, float(h), vp(Xd), vp(Xd),
                                          vp(md), vp(rd), vp(Pc), vp(Pc), 1.0, 0.1, 0, n, None, vp(dv), 0,
                                          ctypes.c_void_p(ws.data_ptr() + 4), None)
    assert st == PN_ERR_ARGUMENT
    with pytest.raises(pn.ArgumentError):  # lists of another cloud
        sph.tlsph_interaction(p, Xd[:, :-1], Xd[:, :-1], cuda(m[:-1]), cuda(rho[:-1]), Pr[:-1], h)
