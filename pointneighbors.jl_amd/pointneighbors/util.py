"""Host-side numeric helpers that mirror src/util.jl and the Julia type promotion rules
the reference relies on for its cell arithmetic."""

from __future__ import annotations

import math

import numpy as np

FLOAT64_EPS = float(np.finfo(np.float64).eps)  # Julia's eps() is eps(Float64)
_TYPEMAX = 2.0**63


def floor_to_int(x) -> int:
    """src/util.jl:19-34: floor with saturation; NaN and values >= typemax(Int) give
    typemax(Int), values <= typemin(Int) give typemin(Int)."""
    r = math.floor(x) if math.isfinite(x) else x
    if isinstance(r, float) and math.isnan(r):
        return 2**63 - 1
    if r >= _TYPEMAX:
        return 2**63 - 1
    if r <= -_TYPEMAX:
        return -(2**63)
    return int(r)


def wrap_int64(v: int) -> int:
    """Int64 arithmetic wraps in Julia (test/nhs_grid.jl:83-85)."""
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >= (1 << 63) else v


def scalar_dtype(v) -> np.dtype:
    """Julia element type of a user scalar: Python floats are Float64, numpy/torch
    scalars keep their type."""
    if isinstance(v, (np.floating,)):
        return v.dtype
    if isinstance(v, (bool, int, np.integer)):
        return np.dtype(np.int64)
    try:
        import torch
        if isinstance(v, torch.Tensor):
            return np.dtype(str(v.dtype).replace("torch.", ""))
    except ImportError:  # pragma: no cover
        pass
    return np.dtype(np.float64)


def array_dtype(a) -> np.dtype:
    """Element type of a corner given as tuple/list/ndarray/tensor."""
    try:
        import torch
        if isinstance(a, torch.Tensor):
            return np.dtype(str(a.dtype).replace("torch.", ""))
    except ImportError:  # pragma: no cover
        pass
    if isinstance(a, np.ndarray):
        return a.dtype
    dts = [scalar_dtype(v) for v in a]
    return np.result_type(*dts) if dts else np.dtype(np.float64)


def to_float_array(a, dtype) -> np.ndarray:
    try:
        import torch
        if isinstance(a, torch.Tensor):
            a = a.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.asarray(a, dtype=dtype).reshape(-1)


def promote(*dts) -> np.dtype:
    """Julia promote_type over Float32/Float64 (and Int64 with a float)."""
    floats = [np.dtype(d) for d in dts if np.dtype(d).kind == "f"]
    if not floats:
        return np.dtype(np.float64)
    return np.dtype(np.float64) if any(d == np.float64 for d in floats) else np.dtype(np.float32)


def dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float32:
        return 0
    if dt == np.float64:
        return 1
    raise TypeError(f"unsupported element type {dt}; use Float32 or Float64")
