"""ctypes binding of libpointneighbors_hip.so (the C-ABI declared in
include/pointneighbors_hip.h) and the mapping of its status codes to the reference's
exceptions.

The library is the only compute path: if it is missing or cannot be loaded, every call
raises; there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; POINTNEIGHBORS_HIP_LIB names another build of the same library (A/B runs)
LIB_PATH = os.environ.get("POINTNEIGHBORS_HIP_LIB") or os.path.join(os.path.dirname(_HERE),
                                                                      "libpointneighbors_hip.so")

PN_OK = 0
PN_ERR_ARGUMENT = 1
PN_ERR_OUT_OF_BOUNDS = 2
PN_ERR_CELL_FULL = 3
PN_ERR_QUERY_BOUNDS = 4
PN_ERR_NOT_INITIALIZED = 5
PN_ERR_HIP = 6
PN_ERR_LIST_FULL = 7
PN_ERR_UNSUPPORTED = 8

PN_FLOAT32 = 0
PN_FLOAT64 = 1

PN_QUERIES_GRID = 0
PN_QUERIES_BINNED = 1


class ArgumentError(ValueError):
    """Julia's `ArgumentError`."""


class BoundsError(IndexError):
    """Julia's `BoundsError`."""


class PointNeighborsError(RuntimeError):
    """Julia's `ErrorException` raised by `error(...)` in the reference."""


_c_int = ctypes.c_int
_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_c_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_p_i64 = ctypes.POINTER(ctypes.c_int64)
_p_dbl = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); must match include/pointneighbors_hip.h
SIGNATURES = {
    "pn_abi_version": (_c_int, []),
    "pn_last_error": (ctypes.c_char_p, []),
    "pn_grid_create": (_c_int, [_c_int, _c_int, _c_int, _c_int, _p_dbl, _p_i64, _c_dbl, _c_i32,
                                ctypes.POINTER(_vp)]),
    "pn_grid_create_spatial_hash": (_c_int, [_c_int, _c_int, _c_int, _c_dbl, _c_i64, _c_i32, ctypes.POINTER(_vp)]),
    "pn_grid_destroy": (_c_int, [_vp]),
    "pn_grid_set_stream": (_c_int, [_vp, _vp]),
    "pn_grid_update": (_c_int, [_vp, _vp, _c_i64, _c_i64, _c_i64, _vp]),
    "pn_grid_update_async": (_c_int, [_vp, _vp, _c_i64, _c_i64, _c_i64, _vp]),
    "pn_grid_check": (_c_int, [_vp]),
    "pn_grid_error": (_c_int, [_vp, _p_i64, _p_i64]),
    "pn_grid_refresh": (_c_int, [_vp, _vp, _c_i64]),
    "pn_grid_bin_queries": (_c_int, [_vp, _vp, _c_i64, _c_i64, _c_i64, _vp]),
    "pn_grid_stats": (_c_int, [_vp, _p_i64, _p_i64, _p_i64]),
    "pn_grid_copy_cells": (_c_int, [_vp, _vp, _vp]),
    "pn_count_neighbors": (_c_int, [_vp, _c_int, _vp, _c_int]),
    "pn_sum_neighbor_indices": (_c_int, [_vp, _c_int, _vp, _c_int, _c_int]),
    "pn_csr_count": (_c_int, [_vp, _c_int, _c_i64, _vp, _c_i32, _p_i64, _p_i64]),
    "pn_csr_fill": (_c_int, [_vp, _c_int, _c_i64, _vp, _vp, _c_int]),
    "pn_list_count_neighbors": (_c_int, [_vp, _c_i64, _c_i64, _c_i64, _vp, _vp, _c_int, _vp]),
    "pn_list_sum_neighbor_indices": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_i64, _c_i64, _vp, _vp,
                                              _c_int, _c_int, _vp]),
    "pn_sph_density": (_c_int, [_vp, _c_int, _c_int, _c_dbl, _vp, _vp, _c_int]),
    "pn_sph_density_vjp": (_c_int, [_vp, _c_int, _c_dbl, _vp, _vp, _vp, _vp, _c_int]),
    "pn_sph_pressure_accel": (_c_int, [_vp, _c_int, _c_dbl, _vp, _vp, _vp, _c_int]),
    "pn_sph_pressure_accel_vjp": (_c_int, [_vp, _c_int, _c_dbl, _vp, _vp, _vp, _vp, _vp, _vp, _c_int]),
    "pn_pair_energy": (_c_int, [_vp, _c_int, _c_dbl, _c_dbl, _vp, _c_int]),
    "pn_pair_energy_vjp": (_c_int, [_vp, _c_dbl, _c_dbl, _vp, _vp, _c_int]),
    "pn_tait_eos": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_dbl, _c_dbl, _c_dbl, _vp]),
    "pn_tait_pressure_term": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_dbl, _c_dbl, _c_dbl, _vp]),
    "pn_batch_create": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_i64, _p_dbl, _p_i64, _c_dbl, _c_i32,
                                 _p_i64, ctypes.POINTER(_vp)]),
    "pn_set_tuning": (_c_int, [_vp, ctypes.c_char_p, _c_i64]),
    "pn_grid_launch_times": (_c_int, [_vp, ctypes.POINTER(ctypes.c_float), _c_i64, _p_i64]),
    "pn_grid_set_periodic": (_c_int, [_vp, _p_dbl, _p_i64, _p_dbl]),
    "pn_displace_uniform": (_c_int, [_vp, _c_i64, _c_int, _c_dbl, ctypes.c_uint64, _vp, _vp]),
    "pn_list_sph_density": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_int, _c_int, _c_int, _c_int, _c_dbl, _vp, _vp,
                                     _vp, _c_i64, _c_i64, _vp, _vp, _c_int, _vp]),
    "pn_csr_to_dvov": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_int, _c_i64, _vp, _vp, _vp]),
    "pn_list_tlsph_deformation_grad": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_int, _c_int, _c_int, _c_dbl, _vp, _vp,
                                                _vp, _vp, _c_int, _c_i64, _c_i64, _vp, _vp, _vp]),
    "pn_list_tlsph_interaction": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_int, _c_int, _c_int, _c_dbl, _vp, _vp, _vp,
                                           _vp, _vp, _vp, _c_dbl, _c_dbl, _c_i64, _c_i64, _vp, _vp, _c_int, _vp,
                                           _vp]),
    "pn_tlsph_interaction_workspace_bytes": (_c_i64, [_c_int, _c_int, _c_i64]),
    "pn_tlsph_pk1_rho2": (_c_int, [_vp, _vp, _vp, _c_i64, _c_int, _c_int, _c_dbl, _c_dbl, _vp, _vp]),
    "pn_nbody_accel": (_c_int, [_vp, _c_int, _c_dbl, _vp, _vp, _c_int]),
    "pn_list_nbody_accel": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_int, _c_int, _c_dbl, _vp, _vp, _vp, _c_i64,
                                     _c_i64, _vp, _vp, _c_int, _vp]),
    "pn_wcsph_interact": (_c_int, [_vp, _c_int, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _vp, _vp, _vp,
                                   _vp, _c_int]),
    "pn_mapreduce_neighbors": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp]),
    "pn_count_neighbors_noalloc": (_c_int, [_c_int, _c_int, _c_int, _c_int, _p_dbl, _p_i64, _c_dbl, _c_i32, _vp,
                                            _c_i64, _vp, _vp]),
    "pn_noalloc_scratch_bytes": (_c_i64, []),
    "pn_list_sph_density_periodic": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_int, _c_int, _c_int, _c_int, _c_dbl,
                                              _vp, _vp, _vp, _c_i64, _c_i64, _vp, _vp, _c_int, _c_dbl, _p_dbl,
                                              _vp]),
}

PN_KERNEL_CUBIC_SPLINE = 0
PN_KERNEL_WENDLAND_C2 = 1

_lib = None


def load(path: str | None = None):
    """Load the HIP library (once).  Raises OSError when it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(
            f"{p} is missing: build it with `python pointneighbors.jl_amd/build.py` "
            "(the MI355X path has no CPU fallback)")
    lib = ctypes.CDLL(p)
    # another build named by POINTNEIGHBORS_HIP_LIB may predate some entry points; the in-tree
    # build must export them all
    other = p != os.path.join(os.path.dirname(_HERE), "libpointneighbors_hip.so")
    for name, (res, args) in SIGNATURES.items():
        if other and not hasattr(lib, name):
            continue
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.pn_abi_version() != 1:
        raise OSError("ABI version mismatch of libpointneighbors_hip.so")
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    return load().pn_last_error().decode()


def check(status: int, context: str = "") -> None:
    """Raise the reference's exception type for a non-zero status."""
    if status == PN_OK:
        return
    msg = last_error()
    if context:
        msg = f"{msg} ({context})" if msg else context
    if status == PN_ERR_ARGUMENT:
        raise ArgumentError(msg)
    if status == PN_ERR_QUERY_BOUNDS:
        raise BoundsError(msg)
    if status in (PN_ERR_OUT_OF_BOUNDS, PN_ERR_CELL_FULL, PN_ERR_LIST_FULL,
                  PN_ERR_NOT_INITIALIZED):
        raise PointNeighborsError(msg)
    raise RuntimeError(f"pointneighbors HIP error {status}: {msg}")
