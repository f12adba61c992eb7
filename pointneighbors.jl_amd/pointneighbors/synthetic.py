"""Seeded synthetic point clouds: jittered lattices.

SplitMix64 (S. Vigna, public domain) evaluated on a counter, so every value is a pure
function of (seed, index) and the stream is identical on any machine.  Uniform variates
are the top 53 bits scaled to [0, 1); jitter is (2u - 1) * a computed in Float64 and
then rounded to the coordinate type.  Lattice index order is x fastest, like Julia's
CartesianIndices (docs/literate/src/tut_basic_usage.jl:16-20); lattice point k along a
dimension sits at (k + 1) * spacing (1-based, as the reference tutorials place points).
"""

from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, counter: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (counter.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n Float64 variates in [0, 1) from stream `seed`, starting at counter `offset`."""
    z = splitmix64(seed, np.arange(offset, offset + n, dtype=np.uint64))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def jittered_lattice(shape, spacing: float, jitter: float, dtype=np.float32, seed: int = 1) -> np.ndarray:
    """(N, D) array of lattice points (x fastest) displaced by U[-jitter, jitter]."""
    shape = tuple(int(s) for s in shape)
    d = len(shape)
    n = int(np.prod(shape))
    axes = [np.arange(s, dtype=np.float64) for s in shape]
    grids = np.meshgrid(*axes, indexing="ij")
    base = np.stack([g.reshape(-1, order="F") for g in grids], axis=1)
    pts = (base + 1.0) * spacing
    if jitter:
        u = uniform(seed, n * d).reshape(n, d)
        pts = pts + (2.0 * u - 1.0) * jitter
    return pts.astype(dtype)


def displacement(n: int, d: int, amplitude: float, seed: int, step: int, dtype=np.float32) -> np.ndarray:
    """Per-step displacement U[-a, a] (N, D), stream keyed by (seed, step)."""
    u = uniform((seed * 1000003 + step * 7919 + 17) & 0xFFFFFFFFFFFFFFFF, n * d).reshape(n, d)
    return ((2.0 * u - 1.0) * amplitude).astype(dtype)


def lattice_box(shape, spacing: float, jitter: float):
    """Bounding box [lo, hi] per dimension that contains every jittered lattice point."""
    lo = np.array([spacing - jitter for _ in shape], dtype=np.float64)
    hi = np.array([s * spacing + jitter for s in shape], dtype=np.float64)
    return lo, hi


def to_device_coords(pts_nd: np.ndarray, device="cuda"):
    """(N, D) host array -> D x N torch tensor with strides (1, D) on the GPU."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(pts_nd)).to(device).t()
