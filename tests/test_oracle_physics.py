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


@pytest.mark.parametrize("kernel", ["cubic_spline", "wendland_c2"])
def test_threaded_rows_equal_list_sums(orc, kernel):
    """The threaded grid-based density / pressure sums (used at 1M and 16.8M points) equal the
    list-based sums on the same Float32 cloud, for all rows and for a scattered subset."""
    sp = 0.01
    pts = synthetic.jittered_lattice((17, 18, 19), sp, 0.1 * sp, dtype=np.float32, seed=8)
    h = np.float32(1.2) * np.float32(sp)
    r = np.float32(2) * h
    lo, hi = pts.min(axis=0) - 3 * r, pts.max(axis=0) + 3 * r
    pmin, _, ncells, _ = orc.fullgrid_geometry(lo, hi, r, np.float32, np.float32)
    g = orc.Grid(3, pmin, ncells, float(r), coord_dtype=np.float32, radius_dtype=np.float32,
                 cell_dtype=np.float32)
    st, counts, _, _ = g.threaded_count_f32(pts, 4)
    assert st == orc.OK
    st, _, lists = g.precomputed(pts, pts, 100000)
    assert st == orc.OK
    assert np.array_equal(np.diff(lists[0]), counts)
    n = len(pts)
    m = synthetic.uniform(3, n) + 0.5
    rho, rs = orc.sph_density(pts, lists, m, float(h), kernel)
    P = synthetic.uniform(4, n) - 0.3
    dv, ds = orc.sph_pressure_accel(pts, lists, m, P, float(h), kernel)
    for rows in (np.arange(n), np.array([0, 5, n - 1, 777, 4000, 12])):
        st, rho2, rs2, dv2, ds2 = g.threaded_sph_rows_f32(pts, rows, 3, m, float(h), kernel, P=P)
        assert st == orc.OK
        np.testing.assert_allclose(rho2, rho[rows], rtol=1e-13, atol=0)
        np.testing.assert_allclose(rs2, rs[rows], rtol=1e-13, atol=0)
        assert (np.abs(dv2 - dv[rows]).max(axis=1) <= 1e-13 * ds[rows]).all()
        np.testing.assert_allclose(ds2, ds[rows], rtol=1e-13, atol=0)


# ---- n-body closure (benchmarks/n_body.jl:34-51) -------------------------------------------

def two_point_lists():
    return (np.array([0, 2, 4]), np.array([0, 1, 0, 1], dtype=np.int32))


def test_nbody_two_bodies_and_conservation(orc):
    """Two bodies attract along their axis with G m_other / d^2; the self pair and coincident
    bodies contribute nothing; sum_i m_i dv_i = 0 over symmetric lists."""
    G, d = 6.6743e-11, 0.37
    x = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
    m = np.array([2e10, 3e10])
    dv, _ = orc.nbody(x, two_point_lists(), m, G)
    assert dv[0] == pytest.approx([G * m[1] / d**2, 0, 0], rel=1e-14)
    assert dv[1] == pytest.approx([-G * m[0] / d**2, 0, 0], rel=1e-14)
    same = np.zeros((2, 3))
    dv, _ = orc.nbody(same, two_point_lists(), m, G)
    assert not dv.any()
    pts = synthetic.jittered_lattice((9, 10, 8), 1.0, 0.3, dtype=np.float64, seed=5)
    lists = neighbor_lists(orc, pts, 2.5)
    m = 1e10 * (synthetic.uniform(5, len(pts)) + 1)
    dv, scale = orc.nbody(pts, lists, m, G)
    assert np.abs((m[:, None] * dv).sum(axis=0)).max() <= 1e-13 * (m * scale).sum()


# ---- WCSPH interaction (TrixiParticles interact!, benchmark_wcsph) ------------------------------

WC = dict(h=0.6, c=10.0, kernel="wendland_c2", alpha=0.02, beta=0.0, epsilon=0.01, delta=0.1)


def wendland_deriv_fd(orc, d, h):
    """W'(d) by central differences of the oracle's density of a two-point system (m = 1)."""
    s = 1e-6
    vals = []
    for dd in (d + s, d - s):
        x = np.array([[0.0, 0.0, 0.0], [dd, 0.0, 0.0]])
        rho, _ = orc.sph_density(x, two_point_lists(), np.ones(2), h, "wendland_c2")
        vals.append(rho[0])
    return (vals[0] - vals[1]) / (2 * s)


def test_wcsph_two_particle_terms(orc):
    """Closed forms of each term for two particles on the x axis (a at 0, b at d > 0), with W'
    from finite differences of the density sum: pressure repels, viscosity opposes an
    approach, an approach raises the density, diffusion lowers the denser particle's density."""
    h, c, d, m = WC["h"], WC["c"], 0.7, 0.3
    x = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
    L = two_point_lists()
    dW = wendland_deriv_fd(orc, d, h)
    assert dW < 0
    # pressure only (at rest, one density)
    v = np.array([[0, 0, 0, 1000.0], [0, 0, 0, 1000.0]])
    p = np.array([40.0, 60.0])
    dv, _ = orc.wcsph_interact(x, L, np.full(2, m), v, p, h, c, **{k: WC[k] for k in ("kernel", "alpha", "beta",
                                                                                   "epsilon", "delta")})
    # grad_a W = W'(d) / d * (x_a - x_b) = -W'(d) e_x; dv_a = -m (p_a + p_b) / rho^2 grad_a W
    want = -m * 100.0 / 1000.0**2 * (-dW)
    assert dv[0, 0] == pytest.approx(want, rel=1e-7) and dv[0, 0] < 0
    assert dv[1, 0] == pytest.approx(-want, rel=1e-7)
    assert dv[0, 3] == 0.0 and not dv[:, 1:3].any()
    # approach at speed 2u: viscosity and continuity, no pressure
    u = 0.5
    v = np.array([[u, 0, 0, 1000.0], [-u, 0, 0, 1000.0]])
    dv, _ = orc.wcsph_interact(x, L, np.full(2, m), v, np.zeros(2), h, c, **{k: WC[k] for k in (
        "kernel", "alpha", "beta", "epsilon", "delta")})
    vr = 2 * u * -d
    mu = h * vr / (d**2 + WC["epsilon"] * h**2)
    pi = -WC["alpha"] * c * mu / 1000.0
    assert dv[0, 0] == pytest.approx(-m * pi * (-dW), rel=1e-7) and dv[0, 0] < 0
    assert dv[0, 3] == pytest.approx(m * 2 * u * (-dW), rel=1e-7) and dv[0, 3] > 0
    # receding pair: no viscosity, the density falls
    dv2, _ = orc.wcsph_interact(x, L, np.full(2, m), -v * np.array([1, 1, 1, -1]), np.zeros(2), h, c,
                                **{k: WC[k] for k in ("kernel", "alpha", "beta", "epsilon", "delta")})
    assert dv2[0, 0] == 0.0 and dv2[0, 3] == pytest.approx(-dv[0, 3], rel=1e-12)
    # diffusion at rest: rho_a > rho_b
    v = np.array([[0, 0, 0, 1010.0], [0, 0, 0, 1000.0]])
    dv, _ = orc.wcsph_interact(x, L, np.full(2, m), v, np.zeros(2), h, c, **{k: WC[k] for k in (
        "kernel", "alpha", "beta", "epsilon", "delta")})
    # psi . grad W = 2 (rho_a - rho_b) (x_ab . grad W) / d^2 = 2 * 10 * (d W'(d)) / d^2
    want = WC["delta"] * h * c * 2 * 10.0 * dW / d * m / 1000.0
    assert dv[0, 3] == pytest.approx(want, rel=1e-7) and dv[0, 3] < 0


def test_wcsph_conservation_and_fields(orc):
    """Random velocities, densities, pressures and masses on a jittered cloud: total momentum
    sum_a m_a dv_a = 0 and, at rest, the diffused volume sum_a (m_a / rho_a) drho_a = 0; a
    uniform velocity field changes nothing; with one density the continuity term is the rate of
    change of the summation density along a linear velocity field, and with one pressure and
    density the pressure term is -2p/rho^2 times its gradient (finite differences)."""
    pts = synthetic.jittered_lattice((8, 9, 7), 0.3, 0.06, dtype=np.float64, seed=3)
    n = len(pts)
    h = 0.3
    r = 2 * h
    lists = neighbor_lists(orc, pts, r)
    kw = {k: WC[k] for k in ("kernel", "alpha", "beta", "epsilon", "delta")}
    m = 0.02 * (1 + synthetic.uniform(7, n))
    rho = 1000 + 5 * synthetic.uniform(8, n)
    vel = synthetic.uniform(9, 3 * n).reshape(n, 3) - 0.5
    p = 10 + 90 * synthetic.uniform(10, n)
    v = np.concatenate([vel, rho[:, None]], axis=1)
    dv, scale = orc.wcsph_interact(pts, lists, m, v, p, h, 10.0, **kw)
    assert np.abs((m[:, None] * dv[:, :3]).sum(axis=0)).max() <= 1e-13 * (m * scale[:, 0]).sum()
    rest = np.concatenate([np.zeros((n, 3)), rho[:, None]], axis=1)
    dr, sr = orc.wcsph_interact(pts, lists, m, rest, p, h, 10.0, **kw)
    assert abs((m / rho * dr[:, 3]).sum()) <= 1e-13 * (m / rho * sr[:, 1]).sum()
    shifted = rest.copy()
    shifted[:, :3] = [0.3, -0.2, 0.7]
    ds, _ = orc.wcsph_interact(pts, lists, m, shifted, p, h, 10.0, **kw)
    np.testing.assert_allclose(ds, dr, rtol=0, atol=1e-12 * np.abs(dr).max())
    # continuity along v = A x, one density and mass, no diffusion or viscosity contribution
    A = np.array([[0.2, -0.1, 0.05], [0.3, 0.1, 0.0], [-0.2, 0.4, -0.3]])
    lin = np.concatenate([pts @ A.T, np.full((n, 1), 1000.0)], axis=1)
    mm = np.full(n, 0.02)
    dl, _ = orc.wcsph_interact(pts, lists, mm, lin, np.zeros(n), h, 10.0, **dict(kw, alpha=0.0))
    s = 1e-6
    rp_, rm_ = (orc.sph_density(pts + sg * s * (pts @ A.T), lists, mm, h, "wendland_c2")[0] for sg in (1, -1))
    np.testing.assert_allclose(dl[:, 3], (rp_ - rm_) / (2 * s), rtol=1e-6, atol=1e-6 * np.abs(dl[:, 3]).max())
    # pressure: -2p/rho^2 grad_a rho_a (the gradient of a point's summation density w.r.t. itself)
    one = np.concatenate([np.zeros((n, 3)), np.full((n, 1), 1000.0)], axis=1)
    dp, _ = orc.wcsph_interact(pts, lists, mm, one, np.full(n, 50.0), h, 10.0, **kw)
    for a in (0, 17, n // 2, n - 5):
        g = []
        for dim in range(3):
            vals = []
            for sg in (1, -1):
                xx = pts.copy()
                xx[a, dim] += sg * s
                vals.append(orc.sph_density(xx, lists, mm, h, "wendland_c2")[0][a])
            g.append((vals[0] - vals[1]) / (2 * s))
        np.testing.assert_allclose(dp[a, :3], -2 * 50.0 / 1000.0**2 * np.array(g), rtol=1e-6,
                                   atol=1e-9 * np.abs(dp[:, :3]).max())
