"""Interaction functors that run as dedicated HIP kernels, and their differentiable
wrappers.

Functors (the `f` of foreach_point_neighbor; outputs indexed by original point index):
  SPHDensity(rho, mass, h)                rho[i] += m[j] W(d)            (self included)
  SPHPressureAcceleration(dv, mass, P, h) dv[:, i] -= m[j] (P[i] + P[j]) grad W_ij, j != i
  WendlandPairEnergy(e, eps, cutoff)      e[i] += 1/2 phi(d), j != i

Differentiable functions (torch.autograd.Function: the counterpart of the
ChainRulesCore.rrule a Julia binding would define).  Their backward passes call
hand-written adjoint kernels that recompute the traversal; neighbour sets are held fixed
(piecewise constant in the coordinates), so gradients are those of the fixed-pair sums:
  sph_density(x, mass, nhs, h)             -> rho
  sph_pressure_acceleration(x, mass, P, nhs, h) -> dv (D x N)
  pair_energy(x, nhs, eps, cutoff)          -> e (per point; E = e.sum())
`nhs` must have been updated with x (x is the neighbour set and the query set).

The physics is not part of the reference (it lives in TrixiParticles.jl); the formulas
are the textbook ones listed in pointneighbors.jl_amd/csrc/pn_sph.h.
"""

from __future__ import annotations

import ctypes

from . import _device
from ._lib import (PN_KERNEL_CUBIC_SPLINE, PN_KERNEL_WENDLAND_C2, PN_QUERIES_GRID, ArgumentError,
                   PointNeighborsError, check, load)

KERNELS = {"cubic_spline": PN_KERNEL_CUBIC_SPLINE, "wendland_c2": PN_KERNEL_WENDLAND_C2}


def _kernel_code(kernel):
    if kernel not in KERNELS:
        raise ArgumentError(f"unknown smoothing kernel {kernel!r}; use one of {sorted(KERNELS)}")
    return KERNELS[kernel]


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _field(t, n, dtype, name, ndims=None):
    import torch
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise ArgumentError(f"{name} must be a GPU tensor")
    if t.dtype != dtype:
        raise ArgumentError(f"{name} must have the radius type {dtype}")
    if ndims is None:
        if t.dim() != 1 or t.shape[0] != n or not t.is_contiguous():
            raise ArgumentError(f"{name} must be a contiguous vector of {n} entries")
    else:
        if tuple(t.shape) != (ndims, n) or (n > 1 and t.stride() != (1, ndims)):
            raise ArgumentError(f"{name} must be a {ndims} x {n} matrix with strides (1, {ndims})")


def _radius_torch_dtype(nhs):
    import torch
    return torch.float32 if nhs.radius_dtype == "float32" else torch.float64


class KernelFunctor:
    """Base of the interaction functors that run as dedicated HIP kernels."""

    def run_grid(self, nhs, x, y, points):  # pragma: no cover - abstract
        raise NotImplementedError

    def run_lists(self, nhs, x, y, points):
        raise ArgumentError(f"{type(self).__name__} is not available on precomputed lists")


class SPHDensity(KernelFunctor):
    def __init__(self, rho, mass, smoothing_length, kernel="cubic_spline", accumulate=True):
        self.rho, self.mass, self.h = rho, mass, float(smoothing_length)
        self.kernel = _kernel_code(kernel)
        self.accumulate = accumulate

    def run_grid(self, nhs, x, y, points):
        mode, x = nhs._prepare(x, y, points)
        td = _radius_torch_dtype(nhs)
        _field(self.rho, x.shape[1], td, "rho")
        _field(self.mass, y.shape[1], td, "mass")
        check(load().pn_sph_density(nhs._handle, mode, self.kernel, self.h, _vp(self.mass), _vp(self.rho),
                                    1 if self.accumulate else 0), "SPHDensity")

    def run_lists(self, nhs, x, y, points):
        """Over the stored lists (no radius test), src/nhs_precomputed.jl:210-247."""
        from .util import dtype_code
        nl = nhs.neighbor_lists
        x = _device.as_coordinates(x, nhs.ndims, "system coordinates")
        y = _device.as_coordinates(y, nhs.ndims, "neighbor coordinates")
        td = _radius_torch_dtype(nhs)
        _field(self.rho, x.shape[1], td, "rho")
        _field(self.mass, y.shape[1], td, "mass")
        first, count, ptr, keep = _device.selection(points, x.shape[1], x.device)
        args = (_vp(nl.row_ptr), _vp(nl.col_idx), nl.n, 0, nhs.ndims, dtype_code(_device.np_dtype(x)),
                dtype_code(nhs.radius_dtype), self.kernel, self.h, _vp(x), _vp(y), _vp(self.mass), first, count,
                ctypes.c_void_p(ptr) if ptr else None, _vp(self.rho), 1 if self.accumulate else 0)
        stream = ctypes.c_void_p(_device.current_stream_handle(x.device))
        if nhs.periodic_box is not None:
            pb = nhs.periodic_box
            box = (ctypes.c_double * nhs.ndims)(*[float(v) for v in pb.size])
            check(load().pn_list_sph_density_periodic(*args, float(nhs.search_radius), box, stream), "SPHDensity")
        else:
            check(load().pn_list_sph_density(*args, stream), "SPHDensity")
        del keep


def _require_self(nhs, x, y, points, what):
    if not _device.same_storage(x, y) or not _device.is_full(points, x.shape[1]):
        raise ArgumentError(f"{what} needs x == y (the neighborhood search's own points)")
    mode, _ = nhs._prepare(x, y, points)
    if mode != PN_QUERIES_GRID:
        raise ArgumentError(f"{what} needs the neighborhood search updated with x and all points active")


class SPHPressureAcceleration(KernelFunctor):
    """P is p / rho^2 per point."""

    def __init__(self, dv, mass, P, smoothing_length, kernel="cubic_spline", accumulate=True):
        self.dv, self.mass, self.P, self.h = dv, mass, P, float(smoothing_length)
        self.kernel = _kernel_code(kernel)
        self.accumulate = accumulate

    def run_grid(self, nhs, x, y, points):
        _require_self(nhs, x, y, points, "SPHPressureAcceleration")
        td = _radius_torch_dtype(nhs)
        n = x.shape[1]
        _field(self.mass, n, td, "mass")
        _field(self.P, n, td, "P")
        _field(self.dv, n, td, "dv", nhs.ndims)
        check(load().pn_sph_pressure_accel(nhs._handle, self.kernel, self.h, _vp(self.mass), _vp(self.P),
                                           _vp(self.dv), 1 if self.accumulate else 0), "SPHPressureAcceleration")


class WendlandPairEnergy(KernelFunctor):
    def __init__(self, e, eps, cutoff, accumulate=True):
        self.e, self.eps, self.cutoff, self.accumulate = e, float(eps), float(cutoff), accumulate

    def run_grid(self, nhs, x, y, points):
        mode, x = nhs._prepare(x, y, points)
        _field(self.e, x.shape[1], _radius_torch_dtype(nhs), "e")
        check(load().pn_pair_energy(nhs._handle, mode, self.eps, self.cutoff, _vp(self.e),
                                    1 if self.accumulate else 0), "WendlandPairEnergy")


def tait_eos(rho, rho0, c, gamma=7.0):
    """p = B ((rho / rho0)^gamma - 1), B = rho0 c^2 / gamma, as a HIP kernel."""
    import torch
    p = torch.empty_like(rho)
    code = 0 if rho.dtype == torch.float32 else 1
    check(load().pn_tait_eos(_vp(rho), _vp(p), rho.numel(), code, float(rho0), float(c), float(gamma),
                             ctypes.c_void_p(_device.current_stream_handle(rho.device))), "tait_eos")
    return p


def tlsph_deformation_gradient(nhs, X, x, volume, smoothing_length, kernel="wendland_c2", correction=None,
                               points=None, _invert=False):
    """TLSPH deformation gradient over the stored lists of `nhs`, a PrecomputedNeighborhoodSearch
    initialized with the initial coordinates X (TrixiParticles 0.5 calc_deformation_grad!, as run by
    benchmarks/smoothed_particle_hydrodynamics.jl:127-146):

        F_a = sum_b V_b (x_b - x_a) (grad_a W(X_a - X_b))^T L_a,   |X_a - X_b| >= sqrt(eps)

    X, x: D x N coordinates; volume: V = m / rho0 per point; correction: L as an (N, D, D)
    tensor (None: identity).  Returns F as an (N, D, D) tensor (F[a] is F_a), one HIP kernel."""
    import torch
    from .util import dtype_code
    nl = getattr(nhs, "neighbor_lists", None)
    if nl is None:
        raise ArgumentError("tlsph_deformation_gradient needs a PrecomputedNeighborhoodSearch")
    if nhs.periodic_box is not None:
        raise ArgumentError("tlsph_deformation_gradient does not support a PeriodicBox")
    D = nhs.ndims
    X = _device.as_coordinates(X, D, "initial coordinates")
    x = _device.as_coordinates(x, D, "current coordinates")
    if X.dtype != x.dtype or X.shape != x.shape:
        raise ArgumentError("X and x must have the same dtype and shape")
    n = X.shape[1]
    if nl.n != n:
        raise ArgumentError(f"the neighbor lists have {nl.n} rows for {n} points: build them on X")
    _field(volume, n, X.dtype, "volume")
    out = torch.empty((n, D, D), dtype=X.dtype, device=X.device)  # out[a] holds F_a column-major
    L = None
    if correction is not None:
        if tuple(correction.shape) != (n, D, D) or correction.dtype != X.dtype:
            raise ArgumentError(f"correction must be an ({n}, {D}, {D}) tensor of {X.dtype}")
        L = correction.transpose(1, 2).contiguous()
    first, count, ptr, keep = _device.selection(points, n, X.device)
    check(load().pn_list_tlsph_deformation_grad(
        _vp(nl.row_ptr), _vp(nl.col_idx), nl.n, 0, D, dtype_code(_device.np_dtype(X)), _kernel_code(kernel),
        float(smoothing_length), _vp(X), _vp(x), _vp(volume), _vp(L) if L is not None else None,
        1 if _invert else 0, first, count, ctypes.c_void_p(ptr) if ptr else None, _vp(out),
        ctypes.c_void_p(_device.current_stream_handle(X.device))), "tlsph_deformation_gradient")
    del keep
    return out.transpose(1, 2)


def tlsph_correction_matrix(nhs, X, volume, smoothing_length, kernel="wendland_c2"):
    """Kernel-gradient correction L_a = M_a^-1, M_a = sum_b V_b (X_b - X_a) (grad_a W(X_ab))^T
    (the deformation gradient of the undeformed body), one HIP list traversal with a closed-form
    D x D inverse per point."""
    return tlsph_deformation_gradient(nhs, X, X, volume, smoothing_length, kernel, _invert=True)


def _matrices(t, n, D, dtype, name):
    """An (N, D, D) tensor -> its column-major per-point storage (a no-op for the views that
    tlsph_deformation_gradient and tlsph_pk1_corrected return)."""
    import torch
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise ArgumentError(f"{name} must be a GPU tensor")
    if tuple(t.shape) != (n, D, D) or t.dtype != dtype:
        raise ArgumentError(f"{name} must be an ({n}, {D}, {D}) tensor of {dtype}")
    return t.transpose(1, 2).contiguous()


def lame_parameters(young_modulus, poisson_ratio):
    """(lambda, mu) of an isotropic linear-elastic material."""
    E, nu = float(young_modulus), float(poisson_ratio)
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def tlsph_pk1_corrected(F, rho, young_modulus, poisson_ratio, correction=None):
    """Corrected first Piola-Kirchhoff stress over rho^2 of a Saint Venant-Kirchhoff material
    (TrixiParticles compute_pk1_corrected!, which the reference's benchmark_tlsph runs once at
    benchmarks/smoothed_particle_hydrodynamics.jl:186):

        E = (F^T F - I) / 2,  S = lambda tr(E) I + 2 mu E,  out = F S L / rho^2

    F, correction (L, None: identity): (N, D, D); rho: material density per point.  Returns an
    (N, D, D) tensor, one HIP kernel."""
    import torch
    from .util import dtype_code
    if not isinstance(F, torch.Tensor) or F.dim() != 3:
        raise ArgumentError("F must be an (N, D, D) GPU tensor")
    n, D = F.shape[0], F.shape[1]
    Fc = _matrices(F, n, D, F.dtype, "F")
    Lc = None if correction is None else _matrices(correction, n, D, F.dtype, "correction")
    _field(rho, n, F.dtype, "rho")
    lam, mu = lame_parameters(young_modulus, poisson_ratio)
    out = torch.empty((n, D, D), dtype=F.dtype, device=F.device)  # out[a] holds the matrix column-major
    check(load().pn_tlsph_pk1_rho2(_vp(Fc), _vp(Lc) if Lc is not None else None, _vp(rho), n, D,
                                   dtype_code(_device.np_dtype(F)), lam, mu, _vp(out),
                                   ctypes.c_void_p(_device.current_stream_handle(F.device))), "tlsph_pk1_corrected")
    return out.transpose(1, 2)


def tlsph_interaction(nhs, X, x, mass, rho, pk1_rho2, smoothing_length, kernel="wendland_c2", F=None,
                      young_modulus=0.0, penalty_alpha=0.0, dv=None, points=None, _packed=True):
    """TLSPH interaction forces over the stored lists of `nhs`, a PrecomputedNeighborhoodSearch of
    the initial coordinates X (TrixiParticles interact_structure_structure!, the loop the
    reference times in benchmark_tlsph, benchmarks/smoothed_particle_hydrodynamics.jl:115-121):

        dv_a += sum_b m_b (Pr_a + Pr_b) grad_a W(X_ab)                       (stress)
              + (alpha/2) V_a V_b W / |X_ab|^2 E delta_ab x_ab / |x_ab| / m_a  (penalty)

    with delta_ab = ((F_a + F_b) X_ab - 2 x_ab) . x_ab / |x_ab| (Ganzenmueller's hourglass
    control, PenaltyForceGanzenmueller; on when penalty_alpha > 0, then F is required).
    X, x: D x N; mass, rho: per point; pk1_rho2, F: (N, D, D).  dv: D x N with the coordinate
    layout (None: a new zeroed one; given: accumulated into).  One HIP kernel; with the penalty
    on, a pack kernel first writes each point's fields into one record of a workspace (fewer,
    wider gathers; the same bits)."""
    import torch
    from .util import dtype_code
    nl = getattr(nhs, "neighbor_lists", None)
    if nl is None:
        raise ArgumentError("tlsph_interaction needs a PrecomputedNeighborhoodSearch")
    if nhs.periodic_box is not None:
        raise ArgumentError("tlsph_interaction does not support a PeriodicBox")
    D = nhs.ndims
    X = _device.as_coordinates(X, D, "initial coordinates")
    x = _device.as_coordinates(x, D, "current coordinates")
    if X.dtype != x.dtype or X.shape != x.shape:
        raise ArgumentError("X and x must have the same dtype and shape")
    n = X.shape[1]
    if nl.n != n:
        raise ArgumentError(f"the neighbor lists have {nl.n} rows for {n} points: build them on X")
    _field(mass, n, X.dtype, "mass")
    _field(rho, n, X.dtype, "rho")
    Pc = _matrices(pk1_rho2, n, D, X.dtype, "pk1_rho2")
    alpha = float(penalty_alpha)
    if not alpha >= 0.0:
        raise ArgumentError("penalty_alpha must be non-negative")
    Fc = None
    if alpha > 0.0:
        if F is None:
            raise ArgumentError("the penalty force needs the deformation gradient F")
        Fc = _matrices(F, n, D, X.dtype, "F")
    if dv is None:
        dv, acc = _vec_like(X), 0
    else:
        _field(dv, n, X.dtype, "dv", ndims=D)
        acc = 1
    ws = None
    if Fc is not None and _packed:
        nbytes = load().pn_tlsph_interaction_workspace_bytes(D, dtype_code(_device.np_dtype(X)), nl.n)
        ws = torch.empty(((nbytes + 15) // 16, 4), dtype=torch.int32, device=X.device)
    first, count, ptr, keep = _device.selection(points, n, X.device)
    check(load().pn_list_tlsph_interaction(
        _vp(nl.row_ptr), _vp(nl.col_idx), nl.n, 0, D, dtype_code(_device.np_dtype(X)), _kernel_code(kernel),
        float(smoothing_length), _vp(X), _vp(x), _vp(mass), _vp(rho), _vp(Pc), _vp(Fc) if Fc is not None else None,
        float(young_modulus), alpha, first, count, ctypes.c_void_p(ptr) if ptr else None, _vp(dv), acc,
        _vp(ws) if ws is not None else None, ctypes.c_void_p(_device.current_stream_handle(X.device))),
        "tlsph_interaction")
    del keep
    return dv


def tait_pressure_term(rho, rho0, c, gamma=7.0, out=None):
    """P = p / rho^2 with p = tait_eos(rho, ...): the per-point term that
    SPHPressureAcceleration takes, in one HIP kernel."""
    import torch
    P = torch.empty_like(rho) if out is None else out
    code = 0 if rho.dtype == torch.float32 else 1
    check(load().pn_tait_pressure_term(_vp(rho), _vp(P), rho.numel(), code, float(rho0), float(c), float(gamma),
                                       ctypes.c_void_p(_device.current_stream_handle(rho.device))),
          "tait_pressure_term")
    return P


def _vec_like(x):
    """A zeroed D x N tensor with the coordinate layout (strides (1, D))."""
    import torch
    return torch.zeros((x.shape[1], x.shape[0]), dtype=x.dtype, device=x.device).t()


def _grid_key(nhs):
    return (nhs._y_key, id(nhs))


def _make_functions():
    import torch

    class _Density(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, mass, nhs, h, kernel):
            _require_self(nhs, x, x, None, "sph_density")
            rho = torch.zeros_like(mass)
            check(load().pn_sph_density(nhs._handle, PN_QUERIES_GRID, kernel, h, _vp(mass), _vp(rho), 0),
                  "sph_density")
            ctx.nhs, ctx.h, ctx.kernel, ctx.key = nhs, h, kernel, _grid_key(nhs)
            ctx.save_for_backward(x, mass)
            return rho

        @staticmethod
        def backward(ctx, rho_bar):
            x, mass = ctx.saved_tensors
            nhs = ctx.nhs
            if _grid_key(nhs) != ctx.key:
                raise PointNeighborsError("the neighborhood search changed between forward and backward")
            rho_bar = rho_bar.contiguous()
            xb = _vec_like(x).to(mass.dtype)
            mb = torch.zeros_like(mass)
            check(load().pn_sph_density_vjp(nhs._handle, ctx.kernel, ctx.h, _vp(mass), _vp(rho_bar), _vp(xb),
                                            _vp(mb), 0), "sph_density backward")
            return xb.to(x.dtype), mb, None, None, None

    class _Pressure(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, mass, P, nhs, h, kernel):
            _require_self(nhs, x, x, None, "sph_pressure_acceleration")
            dv = _vec_like(x).to(mass.dtype)
            check(load().pn_sph_pressure_accel(nhs._handle, kernel, h, _vp(mass), _vp(P), _vp(dv), 0),
                  "sph_pressure_acceleration")
            ctx.nhs, ctx.h, ctx.kernel, ctx.key = nhs, h, kernel, _grid_key(nhs)
            ctx.save_for_backward(x, mass, P)
            return dv

        @staticmethod
        def backward(ctx, dv_bar):
            x, mass, P = ctx.saved_tensors
            nhs = ctx.nhs
            if _grid_key(nhs) != ctx.key:
                raise PointNeighborsError("the neighborhood search changed between forward and backward")
            ab = dv_bar.t().contiguous().t()
            xb = _vec_like(x).to(mass.dtype)
            Pb = torch.zeros_like(P)
            mb = torch.zeros_like(mass)
            check(load().pn_sph_pressure_accel_vjp(nhs._handle, ctx.kernel, ctx.h, _vp(mass), _vp(P), _vp(ab),
                                                   _vp(xb), _vp(Pb), _vp(mb), 0),
                  "sph_pressure_acceleration backward")
            return xb.to(x.dtype), mb, Pb, None, None, None

    class _Energy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, nhs, eps, cutoff):
            _require_self(nhs, x, x, None, "pair_energy")
            td = _radius_torch_dtype(nhs)
            e = torch.zeros(x.shape[1], dtype=td, device=x.device)
            check(load().pn_pair_energy(nhs._handle, PN_QUERIES_GRID, eps, cutoff, _vp(e), 0), "pair_energy")
            ctx.nhs, ctx.eps, ctx.cutoff, ctx.key = nhs, eps, cutoff, _grid_key(nhs)
            ctx.save_for_backward(x)
            return e

        @staticmethod
        def backward(ctx, e_bar):
            (x,) = ctx.saved_tensors
            nhs = ctx.nhs
            if _grid_key(nhs) != ctx.key:
                raise PointNeighborsError("the neighborhood search changed between forward and backward")
            e_bar = e_bar.contiguous()
            xb = _vec_like(x).to(e_bar.dtype)
            check(load().pn_pair_energy_vjp(nhs._handle, ctx.eps, ctx.cutoff, _vp(e_bar), _vp(xb), 0),
                  "pair_energy backward")
            return xb.to(x.dtype), None, None, None

    return _Density, _Pressure, _Energy


_FUNCS = None


def _funcs():
    global _FUNCS
    if _FUNCS is None:
        _FUNCS = _make_functions()
    return _FUNCS


def sph_density(x, mass, nhs, smoothing_length, kernel="cubic_spline"):
    return _funcs()[0].apply(x, mass, nhs, float(smoothing_length), _kernel_code(kernel))


def sph_pressure_acceleration(x, mass, P, nhs, smoothing_length, kernel="cubic_spline"):
    return _funcs()[1].apply(x, mass, P, nhs, float(smoothing_length), _kernel_code(kernel))


def pair_energy(x, nhs, eps, cutoff):
    return _funcs()[2].apply(x, nhs, float(eps), float(cutoff))
