"""ctypes binding of libfvs2d_hip.so (the C-ABI declared in include/fvs2d.h).

The library is loaded from this package directory.  There is no fallback: if
the library is missing, or the machine has no HIP device, calls that need the
device raise :class:`FvsError`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfvs2d_hip.so")

BC_FREESTREAM, BC_SLIP_WALL, BC_DIRICHLET, BC_SOLID_WALL = 0, 1, 2, 3
RK_CLASSIC, RK_SSP42, RK_SSP3, RK_SSP42_SO = 0, 1, 2, 3

# every entry point of include/fvs2d.h
EXPORTS = (
    "fvs_last_error", "fvs_create", "fvs_create_batch", "fvs_batch_time_integration", "fvs_batch_set_active",
    "fvs_batch_wall_forces", "fvs_set_graph", "fvs_destroy", "fvs_sizes", "fvs_mesh_create", "fvs_mesh_destroy",
    "fvs_ctx_mesh", "fvs_mesh_sizes", "fvs_mesh_topology", "fvs_mesh_lists", "fvs_mesh_lsq_nnz", "fvs_mesh_lsq",
    "fvs_set_state", "fvs_get_state", "fvs_initialize", "fvs_primitives", "fvs_gradient", "fvs_gradient_1var",
    "fvs_limiter", "fvs_compute_residual", "fvs_roe_flux_batch", "fvs_local_dt", "fvs_time_integration",
    "fvs_time_integration_noalloc", "fvs_vortex_error", "fvs_wall_forces", "fvs_set_timestep", "fvs_kernel_times",
    "fvs_autotune", "fvs_freeze_limiter", "fvs_batch_freeze_limiter", "fvs_wall_cp", "fvs_last_elapsed_ms",
    "fvs_set_phase_timers", "fvs_phase_times", "fvs_set_persistent", "fvs_set_launch_choices",
    "fvs_mesh_create_batch", "fvs_mesh_cell_order",
)


class FvsError(RuntimeError):
    pass


class Options(C.Structure):
    """fvs_options (include/fvs2d.h)."""
    _fields_ = [
        ("grad_method", C.c_int), ("lsq_nn", C.c_int), ("lsq_p", C.c_double),
        ("limiter", C.c_int), ("recon", C.c_int), ("kappa", C.c_double),
        ("gamma", C.c_double), ("mach", C.c_double), ("aoa_deg", C.c_double),
        ("lvortex", C.c_int), ("vortex_x0", C.c_double), ("vortex_y0", C.c_double),
        ("vortex_K", C.c_double), ("vortex_inf", C.c_double * 4),
        ("rk_scheme", C.c_int), ("rk_order", C.c_int), ("steady", C.c_int),
        ("dt", C.c_double), ("cfl", C.c_double), ("reorder", C.c_int), ("kernel_path", C.c_int),
    ]


_lib = None


def load():
    """Load the library (once).  Raises FvsError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FvsError(f"{LIB_PATH} is missing: build it with `python -m fvs2d_amd.build`")
    L = C.CDLL(LIB_PATH)
    V, D, I = C.c_void_p, C.c_double, C.c_int
    L.fvs_last_error.restype = C.c_char_p
    L.fvs_create.argtypes = [C.POINTER(V), I, V, I, I, V, V, I, V, V, V, C.POINTER(Options)]
    L.fvs_create_batch.argtypes = [C.POINTER(V), I, V, V, I, V, I, I, V, I, V, V, V, C.POINTER(Options)]
    L.fvs_batch_time_integration.argtypes = [V, D, I, V]
    L.fvs_batch_set_active.argtypes = [V, V]
    L.fvs_batch_wall_forces.argtypes = [V, V]
    L.fvs_set_graph.argtypes = [V, I]
    L.fvs_set_phase_timers.argtypes = [V, I]
    L.fvs_set_persistent.argtypes = [V, I]
    L.fvs_set_launch_choices.argtypes = [V, V]
    L.fvs_phase_times.argtypes = [V, V]
    L.fvs_mesh_create.argtypes = [C.POINTER(V), I, V, I, I, V, V, I, V, V, V, I, I, D, I]
    L.fvs_mesh_create_batch.argtypes = [C.POINTER(V), I, V, I, V, I, I, V, I, V, V, V, I, I, D, I, I]
    for n in ("fvs_destroy", "fvs_mesh_destroy"):
        getattr(L, n).argtypes = [V]
    L.fvs_ctx_mesh.restype = V
    L.fvs_ctx_mesh.argtypes = [V]
    L.fvs_sizes.argtypes = [V, V]
    L.fvs_mesh_sizes.argtypes = [V, V]
    L.fvs_mesh_topology.argtypes = [V] * 9
    L.fvs_mesh_lists.argtypes = [V] * 5
    L.fvs_mesh_lsq_nnz.argtypes = [V]
    L.fvs_mesh_lsq.argtypes = [V] * 5
    L.fvs_mesh_cell_order.argtypes = [V, I, I, V, V]
    L.fvs_set_state.argtypes = [V, V]
    L.fvs_get_state.argtypes = [V, V]
    L.fvs_initialize.argtypes = [V, I, D]
    L.fvs_primitives.argtypes = [V, V]
    L.fvs_gradient.argtypes = [V, V]
    L.fvs_gradient_1var.argtypes = [V, V, V]
    L.fvs_limiter.argtypes = [V, V]
    L.fvs_compute_residual.argtypes = [V, D, V, V]
    L.fvs_roe_flux_batch.argtypes = [I, D, V, V, V, V, V]
    L.fvs_local_dt.argtypes = [V, V]
    L.fvs_time_integration.argtypes = [V, D, I, V, V]
    L.fvs_time_integration_noalloc.argtypes = [V, D, I, V, V]
    L.fvs_vortex_error.argtypes = [V, D, V]
    L.fvs_wall_forces.argtypes = [V, V]
    L.fvs_set_timestep.argtypes = [V, D, D]
    L.fvs_kernel_times.argtypes = [V, D, I, V, V, V]
    L.fvs_autotune.argtypes = [V, I, V, V]
    L.fvs_freeze_limiter.argtypes = [V, I]
    L.fvs_batch_freeze_limiter.argtypes = [V, V]
    L.fvs_wall_cp.argtypes = [V, V, V, V]
    L.fvs_last_elapsed_ms.restype = D
    L.fvs_last_elapsed_ms.argtypes = [V]
    _lib = L
    return L


def ptr(a):
    """data pointer of a numpy array (None passes NULL)."""
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def check(rc: int) -> None:
    if rc != 0:
        raise FvsError(f"fvs2d error {rc}: {load().fvs_last_error().decode()}")
