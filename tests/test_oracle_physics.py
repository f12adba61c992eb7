"""Pins the oracle's interaction physics (parity unpinned against the reference: the
physics lives in TrixiParticles.jl, absent here).  Checks that a dropped term, a wrong
sign, normalization or index fails: kernel normalization by quadrature, closed-form
single-pair values, Newton's third law, translation invariance, and dual-number
derivatives against central finite differences.  CPU only."""

import math

import numpy as np
import pytest

from pointneighbors import synthetic


def neighbor_lists(orc, pts, r, cap=1000):
    lo, hi = pts.min(axis=0) - r, pts.max(axis=0) + r
    pmin, _, ncells, _ = orc.fullgrid_geometry(lo, hi, r)
    g = orc.Grid(pts.shape[1], pmin, ncells, r, max_points_per_cell=cap)
    assert g.build(pts)[0] == orc.OK
    st, _, lists = g.precomputed(pts, pts, 100000)
    assert st == orc.OK
    return lists


@pytest.mark.parametrize("kernel", ["cubic_spline", "wendland_c2"])
@pytest.mark.parametrize("nd", [2, 3])
def test_kernel_normalization(orc, kernel, nd):
    """sum_j dV W(|x_c - x_j|) over a fine lattice approximates the integral of W = 1."""
    h = 1.0
    n = 49 if nd == 2 else 33
    dx = 4.4 * h / (n - 1)
    axes = [np.arange(n) * dx - 2.2 * h for _ in range(nd)]
    pts = np.stack([g.reshape(-1) for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    c = int(np.argmin(np.linalg.norm(pts, axis=1)))
    # only the centre row: every lattice point within the support
    nb = np.nonzero(np.linalg.norm(pts - pts[c], axis=1) <= 2.0 * h)[0].astype(np.int32)
    rowptr = np.zeros(len(pts) + 1, dtype=np.int64)
    rowptr[c + 1:] = len(nb)
    rho, _ = orc.sph_density(pts, (rowptr, nb), np.full(len(pts), dx**nd), h, kernel)
    assert abs(rho[c] - 1.0) < 2e-3


@pytest.mark.parametrize("kernel,w0,w1", [("cubic_spline", 1.0, 0.25), ("wendland_c2", 1.0, 3 / 16)])
def test_single_pair_values(orc, kernel, w0, w1):
    h = 0.7
    sigma = (1 / (math.pi * h**3)) if kernel == "cubic_spline" else 21 / (16 * math.pi * h**3)
    pts = np.array([[0.0, 0.0, 0.0], [h, 0.0, 0.0]])
    lists = (np.array([0, 2, 4]), np.array([0, 1, 0, 1], dtype=np.int32))
    rho, _ = orc.sph_density(pts, lists, np.array([1.0, 2.0]), h, kernel)
    # rho_0 = m_0 W(0) + m_1 W(h); W(0) = sigma, W(h) = sigma / 4 (cubic) or 3 sigma / 16 (Wendland)
    assert rho[0] == pytest.approx(sigma * w0 + 2.0 * sigma * w1, rel=1e-14)
    assert rho[1] == pytest.approx(2.0 * sigma * w0 + sigma * w1, rel=1e-14)
    far = np.array([[0.0, 0.0, 0.0], [2 * h, 0.0, 0.0]])
    rho, _ = orc.sph_density(far, lists, np.array([1.0, 1.0]), h, kernel)
    assert rho[0] == pytest.approx(sigma, rel=1e-14)


def test_pair_energy_values(orc):
    eps, rc = 1.7, 2.5
    for r, want in [(0.5 * rc, eps * 0.5**4 * 3.0), (rc, 0.0), (0.2 * rc, eps * 0.8**4 * 1.8)]:
        pts = np.array([[0.0, 0.0], [r, 0.0]])
        lists = (np.array([0, 2, 4]), np.array([0, 1, 0, 1], dtype=np.int32))
        e = orc.pair_energy(pts, lists, eps, rc)
        assert e.sum() == pytest.approx(want, rel=1e-14, abs=1e-300)
        assert e[0] == pytest.approx(e[1], rel=1e-15, abs=1e-300)


def cloud(shape, seed, jitter=0.2):
    return synthetic.jittered_lattice(shape, 1.0, jitter, dtype=np.float64, seed=seed)


def test_pressure_newton_third_law(orc):
    pts = cloud((9, 10, 8), 3)
    h = 1.2
    lists = neighbor_lists(orc, pts, 2 * h)
    m = np.full(len(pts), 0.8)
    rho, _ = orc.sph_density(pts, lists, m, h)
    P = orc.tait_eos(rho, 1.0, 10.0) / rho**2
    dv, scale = orc.sph_pressure_accel(pts, lists, m, P, h)
    total = (m[:, None] * dv).sum(axis=0)
    assert np.abs(total).max() < 1e-10 * np.abs(m[:, None] * dv).sum()
    # two points: a_1 = -m (P_1 + P_2) W'(r) (x_1 - x_2) / r, opposite to a_2
    two = np.array([[0.0, 0.0, 0.0], [0.9, 0.3, -0.2]])
    lists2 = (np.array([0, 2, 4]), np.array([0, 1, 0, 1], dtype=np.int32))
    dv2, _ = orc.sph_pressure_accel(two, lists2, np.array([1.0, 1.0]), np.array([0.5, 1.5]), h)
    assert np.allclose(dv2[0], -dv2[1], rtol=1e-14)
    r = np.linalg.norm(two[0] - two[1])
    # cubic spline W'(r) for q = r/h < 1: sigma/h (-3q + 9/4 q^2); the pair repels (P > 0)
    q = r / h
    dw = 1 / (math.pi * h**3) / h * (-3 * q + 2.25 * q * q)
    want = -(0.5 + 1.5) * dw * (two[0] - two[1]) / r
    assert np.allclose(dv2[0], want, rtol=1e-13)


@pytest.mark.parametrize("which", [0, 1, 2])
def test_dual_matches_central_differences(orc, which):
    pts = cloud((7, 8, 6), 5)
    kw = dict(kernel="wendland_c2", h=1.25, eps=1.3, rc=2.5, rho0=1.0, c=1.0, gamma=7.0)
    lists = neighbor_lists(orc, pts, 2.5)
    m = np.full(len(pts), 0.9)
    step = 1e-6
    idx = (synthetic.uniform(17 + which, 6) * len(pts)).astype(int)
    for a, k in enumerate(idx):
        dim = a % 3
        _, d = orc.loss_dual(which, pts, lists, k, dim, mass=m, **kw)
        xp, xm = pts.copy(), pts.copy()
        xp[k, dim] += step
        xm[k, dim] -= step
        lp, _ = orc.loss_dual(which, xp, lists, k, dim, mass=m, **kw)
        lm, _ = orc.loss_dual(which, xm, lists, k, dim, mass=m, **kw)
        fd = (lp - lm) / (2 * step)
        assert abs(fd - d) <= 1e-6 * max(abs(d), 1e-3), (which, k, dim, d, fd)


def test_energy_gradient_translation_invariance(orc):
    pts = cloud((5, 6, 5), 7)
    lists = neighbor_lists(orc, pts, 2.5)
    g = np.zeros(3)
    for k in range(len(pts)):
        for dim in range(3):
            g[dim] += orc.loss_dual(0, pts, lists, k, dim, eps=1.0, rc=2.5)[1]
    assert np.abs(g).max() < 1e-12 * len(pts)


# ---- TLSPH deformation gradient (TrixiParticles 0.5 calc_deformation_grad!) ----------------

@pytest.mark.parametrize("kernel", ["wendland_c2", "cubic_spline"])
@pytest.mark.parametrize("nd", [2, 3])
def test_tlsph_grad_closed_forms(orc, kernel, nd):
    """The correction matrix M (x = X) is symmetric; with L = M^-1 an affine map x = A X + c
    gives F = A at every point (boundary points included), a rotation gives R, and the
    undeformed body gives I.  A transposed outer product or a wrong sign fails these."""
    shape = (9, 8, 7) if nd == 3 else (14, 13)
    X = synthetic.jittered_lattice(shape, 1.0, 0.15, dtype=np.float64, seed=21)
    h = 1.3
    lists = neighbor_lists(orc, X, 2 * h)
    V = 1.0 + 0.1 * synthetic.uniform(5, len(X))
    M = orc.tlsph_grad(X, X, lists, V, h, kernel)
    assert np.allclose(M, M.transpose(0, 2, 1), rtol=0, atol=1e-12 * np.abs(M).max())
    L = np.linalg.inv(M)
    rng = np.random.default_rng(3)
    A = np.eye(nd) + 0.3 * rng.standard_normal((nd, nd))
    c = rng.standard_normal(nd)
    F = orc.tlsph_grad(X, X @ A.T + c, lists, V, h, kernel, L=L)
    assert np.allclose(F, A, rtol=0, atol=1e-10)
    th = 0.7
    R = np.eye(nd)
    R[:2, :2] = [[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]]
    F = orc.tlsph_grad(X, X @ R.T, lists, V, h, kernel, L=L)
    assert np.allclose(F, R, rtol=0, atol=1e-10)
    assert np.allclose(orc.tlsph_grad(X, X, lists, V, h, kernel, L=L), np.eye(nd), rtol=0, atol=1e-10)
    # uncorrected, in the interior of a regular lattice M is close to the identity
    lat = synthetic.jittered_lattice(shape, 1.0, 0.0, dtype=np.float64)
    Ml = orc.tlsph_grad(lat, lat, neighbor_lists(orc, lat, 2 * h), np.ones(len(lat)), h, kernel)
    centre = int(np.argmin(np.linalg.norm(lat - lat.mean(axis=0), axis=1)))
    assert np.allclose(Ml[centre], np.eye(nd), atol=0.05)


def test_tlsph_grad_translation_and_locality(orc):
    """F depends on x only through differences (a translation changes nothing), and a point
    displaced outside every neighbour's support changes only rows that list it."""
    X = synthetic.jittered_lattice((8, 8, 8), 1.0, 0.2, dtype=np.float64, seed=4)
    h = 1.2
    lists = neighbor_lists(orc, X, 2 * h)
    V = np.ones(len(X))
    x = X + 0.05 * np.sin(X)
    F = orc.tlsph_grad(X, x, lists, V, h)
    assert np.allclose(orc.tlsph_grad(X, x + np.array([3.0, -1.0, 2.5]), lists, V, h), F, rtol=0, atol=1e-12)
    rp, cols = lists
    k = 100
    x2 = x.copy()
    x2[k] += 0.3
    F2 = orc.tlsph_grad(X, x2, lists, V, h)
    rows = {a for a in range(len(X)) if k in cols[rp[a]:rp[a + 1]]}
    changed = {a for a in range(len(X)) if not np.allclose(F2[a], F[a], rtol=0, atol=1e-14)}
    assert changed <= rows and k in changed


# ---- TLSPH interaction (TrixiParticles interact_structure_structure!) ------------------------

def _svk_energy(F, V, lam, mu):
    """sum_c V_c Psi(F_c), Psi = mu E:E + lambda/2 tr(E)^2, E = (F^T F - I)/2 (numpy)."""
    nd = F.shape[1]
    E = 0.5 * (np.einsum("nki,nkj->nij", F, F) - np.eye(nd))
    tr = np.trace(E, axis1=1, axis2=2)
    return float(np.sum(V * (mu * np.sum(E * E, axis=(1, 2)) + 0.5 * lam * tr * tr)))


@pytest.mark.parametrize("kernel", ["wendland_c2", "cubic_spline"])
@pytest.mark.parametrize("nd", [2, 3])
def test_tlsph_stress_force_is_energy_gradient(orc, kernel, nd):
    """With one material density and a symmetric correction L, the stress force is exactly the
    negative gradient of the discrete strain energy: m_a dv_a = -dE/dx_a, E = sum_c V_c Psi(F_c),
    F from orc_tlsph_grad, Psi the Saint Venant-Kirchhoff energy (numpy), checked by central
    differences.  L applied on the wrong side or left out, the second Piola-Kirchhoff stress in
    place of the first, a wrong sign or a wrong rho power fails it."""
    shape = (6, 5, 5) if nd == 3 else (9, 8)
    X = synthetic.jittered_lattice(shape, 1.0, 0.15, dtype=np.float64, seed=11)
    h = 1.2
    lists = neighbor_lists(orc, X, 2 * h)
    n = len(X)
    rho = np.full(n, 1000.0)
    V = 1.0 + 0.2 * synthetic.uniform(8, n)
    m = V * rho
    M = orc.tlsph_grad(X, X, lists, V, h, kernel)
    L = np.linalg.inv(M)
    L = 0.5 * (L + L.transpose(0, 2, 1))
    lam, mu = 0.7, 0.45
    x = X + 0.06 * np.sin(1.3 * X[:, ::-1]) + 0.02 * X
    F = orc.tlsph_grad(X, x, lists, V, h, kernel, L=L)
    Pr = orc.tlsph_pk1_rho2(F, rho, lam, mu, L=L)
    dv = orc.tlsph_force(X, x, lists, m, rho, Pr, h, kernel)
    energy = lambda y: _svk_energy(orc.tlsph_grad(X, y, lists, V, h, kernel, L=L), V, lam, mu)
    eps = 1e-6
    scale = np.abs(m[:, None] * dv).max()
    for a in [0, n // 3, n // 2, n - 1]:
        for d in range(nd):
            xp, xm = x.copy(), x.copy()
            xp[a, d] += eps
            xm[a, d] -= eps
            g = (energy(xp) - energy(xm)) / (2 * eps)
            assert -m[a] * dv[a, d] == pytest.approx(g, rel=1e-5, abs=1e-6 * scale)


def _two_points(nd, d=1.0, squeeze=0.0):
    X = np.zeros((2, nd))
    X[1, 0] = d
    x = X.copy()
    x[1, 0] = d - squeeze
    lists = (np.array([0, 2, 4]), np.array([0, 1, 0, 1], dtype=np.int32))
    return X, x, lists


@pytest.mark.parametrize("nd", [2, 3])
def test_tlsph_pair_signs_and_penalty_value(orc, nd):
    """Tension (Pr = p I, p > 0) pulls a pair together, compression pushes it apart; a pair
    squeezed against F = I is pushed apart by the penalty with Ganzenmueller's magnitude
    (alpha/2) V_a V_b W(|X_ab|) / |X_ab|^2 E delta / m_a, here delta = 2 * squeeze."""
    h = 0.8
    X, x, lists = _two_points(nd, 1.0, 0.1)
    m, rho = np.array([2.0, 3.0]), np.array([4.0, 5.0])
    eye = np.broadcast_to(np.eye(nd), (2, nd, nd))
    for p, sign in [(1.0, +1), (-1.0, -1)]:
        dv = orc.tlsph_force(X, X, lists, m, rho, p * eye, h)
        assert sign * dv[0, 0] > 0 and sign * dv[1, 0] < 0
        assert m[0] * dv[0] == pytest.approx(-m[1] * dv[1], rel=1e-14)
    zero = np.zeros((2, nd, nd))
    alpha, E = 0.1, 1.4e6
    dv = orc.tlsph_force(X, x, lists, m, rho, zero, h, F=eye, young_modulus=E, penalty_alpha=alpha)
    Va, Vb = m / rho
    # W(|X_ab| = 1) of Wendland C2 through the density sum rho_0 = 0 W(0) + 1 W(1)
    w1 = orc.sph_density(X, lists, np.array([0.0, 1.0]), h, "wendland_c2")[0][0]
    assert w1 > 0
    want = 0.5 * alpha * Va * Vb * w1 / 1.0**2 * E * (2 * 0.1) / m[0]
    assert dv[0, 0] == pytest.approx(-want, rel=1e-12)
    assert np.all(dv[0, 1:] == 0) and m[0] * dv[0, 0] == pytest.approx(-m[1] * dv[1, 0], rel=1e-12)


@pytest.mark.parametrize("nd", [2, 3])
def test_tlsph_momentum_and_affine_penalty(orc, nd):
    """Every pair term is antisymmetric: sum_a m_a dv_a = 0 with arbitrary stresses and
    deformation gradients, penalty on.  Under an affine map x = A X + c with F_a = A the
    penalty adds exactly nothing (its error vector vanishes)."""
    shape = (7, 6, 5) if nd == 3 else (12, 11)
    X = synthetic.jittered_lattice(shape, 1.0, 0.2, dtype=np.float64, seed=5)
    h = 1.1
    lists = neighbor_lists(orc, X, 2 * h)
    n = len(X)
    rng = np.random.default_rng(2)
    m = 1.0 + 0.3 * rng.random(n)
    rho = 1.0 + 0.3 * rng.random(n)
    Pr = rng.standard_normal((n, nd, nd))
    F = np.eye(nd) + 0.2 * rng.standard_normal((n, nd, nd))
    x = X + 0.1 * rng.standard_normal(X.shape)
    dv = orc.tlsph_force(X, x, lists, m, rho, Pr, h, F=F, young_modulus=3.0, penalty_alpha=0.1)
    scale = np.abs(m[:, None] * dv).max()
    assert np.abs((m[:, None] * dv).sum(axis=0)).max() <= 1e-12 * n * scale
    A = np.eye(nd) + 0.25 * rng.standard_normal((nd, nd))
    xa = X @ A.T + 0.7
    FA = np.broadcast_to(A, (n, nd, nd))
    base = orc.tlsph_force(X, xa, lists, m, rho, Pr, h)
    pen = orc.tlsph_force(X, xa, lists, m, rho, Pr, h, F=FA, young_modulus=3.0, penalty_alpha=0.1)
    assert np.allclose(pen, base, rtol=0, atol=1e-12 * np.abs(base).max())
    assert not np.allclose(dv, orc.tlsph_force(X, x, lists, m, rho, Pr, h), rtol=0, atol=1e-6)
