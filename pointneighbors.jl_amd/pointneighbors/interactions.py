"""Interaction functors of the reference's benchmark closures, each a dedicated HIP kernel.

  NBodyAcceleration(dv, mass, G)
      benchmarks/n_body.jl:34-51 (benchmark_n_body): for every neighbour j of i at distance
      d >= sqrt(eps), dv[:, i] += -G m[j] pos_diff / d^3.  Grid and precomputed searches.

  WCSPHInteraction(dv, v, pressure, mass, smoothing_length, sound_speed; ...)
      The fluid-fluid `TrixiParticles.interact!` that the reference's benchmark_wcsph times
      (benchmarks/smoothed_particle_hydrodynamics.jl:44-102): weakly compressible SPH with
      ContinuityDensity, a precomputed pressure (StateEquationCole), Monaghan's artificial
      viscosity and Molteni-Colagrossi density diffusion.  v and dv are (D+1) x N matrices
      (velocity rows, then the density row), TrixiParticles' layout with ContinuityDensity.
      The formulas are in pointneighbors.jl_amd/csrc/pn_interact.h.  TrixiParticles itself is
      not part of the reference; its published formulas are restated.

Both take the place of the reference's closure `f` in foreach_point_neighbor.
"""

from __future__ import annotations

import ctypes

from . import _device
from ._lib import ArgumentError, check, load
from .sph import KernelFunctor, _field, _kernel_code, _radius_torch_dtype, _require_self, _vp


def _same_type(nhs, x):
    td = _radius_torch_dtype(nhs)
    if x.dtype != td:
        raise ArgumentError("this functor needs coordinates and search radius of one floating type")
    return td


class NBodyAcceleration(KernelFunctor):
    """benchmarks/n_body.jl:34-51.  dv: D x N (strides (1, D)); mass: one entry per
    neighbour point; G converted to the coordinate type as the reference does."""

    def __init__(self, dv, mass, G=6.6743e-11, accumulate=True):
        self.dv, self.mass, self.G, self.accumulate = dv, mass, float(G), accumulate

    def run_grid(self, nhs, x, y, points):
        mode, x = nhs._prepare(x, y, points)
        td = _same_type(nhs, x)
        _field(self.dv, x.shape[1], td, "dv", nhs.ndims)
        _field(self.mass, y.shape[1], td, "mass")
        check(load().pn_nbody_accel(nhs._handle, mode, self.G, _vp(self.mass), _vp(self.dv),
                                    1 if self.accumulate else 0), "NBodyAcceleration")

    def run_lists(self, nhs, x, y, points):
        """Over the stored lists (src/nhs_precomputed.jl:210-247)."""
        from .util import dtype_code
        nl = nhs.neighbor_lists
        x = _device.as_coordinates(x, nhs.ndims, "system coordinates")
        y = _device.as_coordinates(y, nhs.ndims, "neighbor coordinates")
        td = _same_type(nhs, x)
        if nhs.periodic_box is not None:
            raise ArgumentError("NBodyAcceleration over precomputed lists does not support a PeriodicBox")
        _field(self.dv, x.shape[1], td, "dv", nhs.ndims)
        _field(self.mass, y.shape[1], td, "mass")
        first, count, ptr, keep = _device.selection(points, x.shape[1], x.device)
        check(load().pn_list_nbody_accel(_vp(nl.row_ptr), _vp(nl.col_idx), nl.n, 0, nhs.ndims,
                                         dtype_code(_device.np_dtype(x)), self.G, _vp(x), _vp(y),
                                         _vp(self.mass), first, count, ctypes.c_void_p(ptr) if ptr else None,
                                         _vp(self.dv), 1 if self.accumulate else 0,
                                         ctypes.c_void_p(_device.current_stream_handle(x.device))),
              "NBodyAcceleration")
        del keep


class WCSPHInteraction(KernelFunctor):
    """TrixiParticles' fluid-fluid interact! with the benchmark_wcsph models.

    dv, v: (D+1) x N matrices (strides (1, D+1)); pressure, mass: N vectors; the defaults are
    the benchmark's ArtificialViscosityMonaghan(alpha = 0.02, beta = 0, epsilon = 0.01),
    DensityDiffusionMolteniColagrossi(delta = 0.1) and WendlandC2Kernel.  dv accumulates, as
    interact! does.  x must be the search's own points (a fluid interacting with itself)."""

    def __init__(self, dv, v, pressure, mass, smoothing_length, sound_speed, *, kernel="wendland_c2",
                 alpha=0.02, beta=0.0, epsilon=0.01, delta=0.1, accumulate=True):
        self.dv, self.v, self.pressure, self.mass = dv, v, pressure, mass
        self.h, self.c = float(smoothing_length), float(sound_speed)
        self.kernel = _kernel_code(kernel)
        self.alpha, self.beta, self.epsilon, self.delta = float(alpha), float(beta), float(epsilon), float(delta)
        self.accumulate = accumulate

    def run_grid(self, nhs, x, y, points):
        _require_self(nhs, x, y, points, "WCSPHInteraction")
        td = _same_type(nhs, x)
        n, D = x.shape[1], nhs.ndims
        _field(self.v, n, td, "v", D + 1)
        _field(self.dv, n, td, "dv", D + 1)
        _field(self.pressure, n, td, "pressure")
        _field(self.mass, n, td, "mass")
        check(load().pn_wcsph_interact(nhs._handle, self.kernel, self.h, self.c, self.alpha, self.beta,
                                       self.epsilon, self.delta, _vp(self.mass), _vp(self.v), _vp(self.pressure),
                                       _vp(self.dv), 1 if self.accumulate else 0), "WCSPHInteraction")


def state_equation_cole(rho, reference_density, sound_speed, exponent):
    """StateEquationCole (no background pressure): p = rho0 c^2 / gamma ((rho / rho0)^gamma - 1),
    the Tait form, as a HIP kernel (TrixiParticles' compute_pressure!)."""
    from .sph import tait_eos
    return tait_eos(rho, reference_density, sound_speed, exponent)
