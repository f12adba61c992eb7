"""Device-side plumbing: coordinate matrices as torch tensors, selections, streams.

Coordinates follow the reference's layout (src/neighborhood_search.jl:24-26): a D x N
matrix whose column i is point i.  In torch that is a (D, N) tensor with strides
(1, D) -- the transpose view of a contiguous (N, D) tensor -- which is the exact memory
of Julia's column-major D x N array.  Other layouts are copied to that layout.
"""

from __future__ import annotations

import numpy as np

from ._lib import ArgumentError, BoundsError


def _torch():
    import torch
    return torch


def as_coordinates(a, ndims: int, name: str = "coordinates"):
    """Validate a D x N device coordinate matrix; returns a point-major view."""
    torch = _torch()
    if not isinstance(a, torch.Tensor):
        raise ArgumentError(f"{name} must be a torch tensor on the GPU (D x N)")
    if a.device.type != "cuda":
        raise ArgumentError(
            f"{name} lives on {a.device}; the MI355X engine runs on GPU tensors only "
            "(no CPU fallback)")
    if a.dim() != 2 or a.shape[0] != ndims:
        raise BoundsError(f"{name} of size {tuple(a.shape)} is not a {ndims} x N matrix")
    if a.dtype not in (torch.float32, torch.float64):
        raise ArgumentError(f"{name} must be Float32 or Float64, got {a.dtype}")
    n = a.shape[1]
    if n > 0 and not (a.stride(0) == 1 and (a.stride(1) == ndims or n == 1)):
        a = a.t().contiguous().t()
    return a


_NP_DTYPES: dict = {}


def np_dtype(t) -> np.dtype:
    dt = _NP_DTYPES.get(t.dtype)
    if dt is None:
        dt = _NP_DTYPES[t.dtype] = np.dtype(str(t.dtype).replace("torch.", ""))
    return dt


def selection(sel, n: int, device):
    """A selection of column indices (eachindex_y / points) as (first, count, list).

    `None` is all columns; a Python `range` with step 1 stays a range; anything else
    becomes a device int32 index list (0-based)."""
    torch = _torch()
    if sel is None:
        return 0, n, None, None
    if isinstance(sel, range) and sel.step == 1:
        first, count = sel.start, max(0, sel.stop - sel.start)
        if count and (first < 0 or first + count > n):
            raise BoundsError(f"attempt to access {n}-element columns at index {sel}")
        return first, count, None, None
    if isinstance(sel, torch.Tensor):
        idx = sel.to(device=device, dtype=torch.int32).contiguous().reshape(-1)
    else:
        idx = torch.as_tensor(np.asarray(list(sel), dtype=np.int32), device=device)
    if idx.numel():
        lo, hi = int(idx.min()), int(idx.max())
        if lo < 0 or hi >= n:
            raise BoundsError(f"attempt to access {n}-element columns at index [{lo}, {hi}]")
    return 0, int(idx.numel()), idx.data_ptr(), idx


def is_full(sel, n: int) -> bool:
    if sel is None:
        return True
    return isinstance(sel, range) and sel.step == 1 and sel.start == 0 and sel.stop == n


def current_stream_handle(device) -> int:
    """The raw HIP stream of torch's current stream on `device`."""
    torch = _torch()
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        idx = device.index if device.index is not None else torch.cuda.current_device()
        return raw(idx)
    return torch.cuda.current_stream(device).cuda_stream


def same_storage(a, b) -> bool:
    return (a is b) or (a.data_ptr() == b.data_ptr() and tuple(a.shape) == tuple(b.shape)
                        and a.stride() == b.stride() and a.dtype == b.dtype)
