"""ctypes binding of the device library libanuga_hip.so (include/anuga_hip.h).

The library is built in-tree by anuga_core_amd.build (or __graft_entry__.build())
and loaded from this package directory. There is no fallback: if the library is
missing or cannot be loaded, `lib()` raises, and so does every Domain method that
needs the device.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
# ANUGA_HIP_LIBRARY: another build of the library (the autotuner's variants)
LIB_PATH = os.environ.get("ANUGA_HIP_LIBRARY") or os.path.join(HERE, "libanuga_hip.so")

ANUGA_OK = 0
ANUGA_E_ARG = -1
ANUGA_E_HIP = -2
ANUGA_E_SEMI_IMPLICIT = -3
ANUGA_E_NO_NEIGHBOUR = -4
ANUGA_E_STATE = -5
ANUGA_E_TIMESTEP = -6

STAGE, XMOMENTUM, YMOMENTUM, ELEVATION, FRICTION, HEIGHT = range(6)
PRECISION_FP64, PRECISION_MIXED = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)
_fp = C.POINTER(C.c_float)


class AnugaMesh(C.Structure):
    _fields_ = [("number_of_elements", C.c_int64), ("boundary_length", C.c_int64),
                ("neighbours", _ip), ("neighbour_edges", _ip),
                ("surrogate_neighbours", _ip), ("number_of_boundaries", _ip),
                ("normals", _dp), ("edgelengths", _dp), ("radii", _dp), ("areas", _dp),
                ("centroid_coordinates", _dp), ("edge_coordinates", _dp),
                ("vertex_coordinates", _dp)]


class AnugaParams(C.Structure):
    _fields_ = [("g", C.c_double), ("epsilon", C.c_double),
                ("minimum_allowed_height", C.c_double),
                ("beta_w", C.c_double), ("beta_w_dry", C.c_double),
                ("beta_uh", C.c_double), ("beta_uh_dry", C.c_double),
                ("beta_vh", C.c_double), ("beta_vh_dry", C.c_double),
                ("low_froude", C.c_int64), ("extrapolate_velocity_second_order", C.c_int64),
                ("timestep_fluxcalls", C.c_int64), ("precision", C.c_int64),
                ("reorder", C.c_int64), ("use_sloped_mannings", C.c_int64)]


class TimeControl(C.Structure):
    _fields_ = [("CFL", C.c_double), ("evolve_max_timestep", C.c_double),
                ("evolve_min_timestep", C.c_double), ("relative_finaltime", C.c_double),
                ("relative_yieldtime", C.c_double), ("fixed_flux_timestep", C.c_double),
                ("epsilon", C.c_double), ("max_smallsteps", C.c_int64),
                ("default_order", C.c_int64)]


class StepState(C.Structure):
    _fields_ = [("relative_time", C.c_double), ("timestep", C.c_double),
                ("flux_timestep", C.c_double), ("recorded_min_timestep", C.c_double),
                ("recorded_max_timestep", C.c_double), ("smallsteps", C.c_int64),
                ("order", C.c_int64), ("steps_taken", C.c_int64), ("reached", C.c_int64)]


# name -> (restype, argtypes); every symbol declared in include/anuga_hip.h
SIGNATURES = {
    "anuga_create": (C.c_int, [C.POINTER(AnugaMesh), C.POINTER(AnugaParams),
                               C.POINTER(C.c_void_p)]),
    "anuga_destroy": (None, [C.c_void_p]),
    "anuga_last_error": (C.c_char_p, [C.c_void_p]),
    "anuga_set_params": (C.c_int, [C.c_void_p, C.POINTER(AnugaParams)]),
    "anuga_set_centroid_values": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "anuga_get_centroid_values": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "anuga_get_edge_values": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "anuga_set_boundary": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), _dp]),
    "anuga_set_riverwalls": (C.c_int, [C.c_void_p, _ip, C.c_int64, _dp, _ip, C.c_int64,
                                       C.c_int64, _dp]),
    "anuga_set_boundary_values": (C.c_int, [C.c_void_p, C.c_int64, _dp]),
    "anuga_add_rate_operator": (C.c_int, [C.c_void_p, C.c_int64, _ip, _dp, C.c_double,
                                          C.c_double, _ip]),
    "anuga_add_set_stage_operator": (C.c_int, [C.c_void_p, C.c_int64, _ip, _dp, C.c_double,
                                               _ip]),
    "anuga_update_rate_operator": (C.c_int, [C.c_void_p, C.c_int64, _dp, C.c_double,
                                             C.c_double]),
    "anuga_apply_operators": (C.c_int, [C.c_void_p]),
    "anuga_add_inlet_operator": (C.c_int, [C.c_void_p, C.c_int64, _ip, _dp, C.c_double,
                                           C.c_int64, C.c_double, C.c_double, C.c_double,
                                           C.c_double, _ip]),
    "anuga_update_inlet_operator": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_double]),
    "anuga_get_inlet_operator_stats": (C.c_int, [C.c_void_p, C.c_int64, _dp]),
    "anuga_apply_operator": (C.c_int, [C.c_void_p, C.c_int64]),
    "anuga_get_centroid_subset": (C.c_int, [C.c_void_p, C.c_int64, _ip, _dp]),
    "anuga_set_centroid_subset": (C.c_int, [C.c_void_p, C.c_int64, _ip, _dp]),
    "anuga_get_operator_influx": (C.c_int, [C.c_void_p, C.c_int64, _dp]),
    "anuga_set_output_nodes": (C.c_int, [C.c_void_p, C.c_int64, _ip, _ip]),
    "anuga_get_sww_frame": (C.c_int, [C.c_void_p, C.c_double, _fp, _fp]),
    "anuga_get_static_output": (C.c_int, [C.c_void_p, C.c_int, _fp, _fp]),
    "anuga_set_gauges": (C.c_int, [C.c_void_p, C.c_int64, _ip, C.c_int64]),
    "anuga_sample_gauges": (C.c_int, [C.c_void_p, C.c_double]),
    "anuga_get_gauge_samples": (C.c_int, [C.c_void_p, C.c_int64, _ip, _dp, _dp]),
    "anuga_distribute_to_vertices_and_edges": (C.c_int, [C.c_void_p, _dp]),
    "anuga_compute_fluxes": (C.c_int, [C.c_void_p, _dp]),
    "anuga_update_conserved_quantities": (C.c_int, [C.c_void_p, C.c_double, _ip]),
    "anuga_rk2_second_substep": (C.c_int, [C.c_void_p, C.c_double, _ip]),
    "anuga_rk3_later_substeps": (C.c_int, [C.c_void_p, C.c_double, _ip]),
    "anuga_backup_conserved_quantities": (C.c_int, [C.c_void_p]),
    "anuga_saxpy_conserved_quantities": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "anuga_divide_conserved_quantities": (C.c_int, [C.c_void_p, C.c_double]),
    "anuga_update_boundary": (C.c_int, [C.c_void_p]),
    "anuga_compute_fluxes_ext_central": (C.c_int, [C.c_void_p, C.c_double, _dp]),
    "anuga_compute_forcing_terms": (C.c_int, [C.c_void_p]),
    "anuga_update_centroids": (C.c_int, [C.c_void_p, C.c_double]),
    "anuga_fix_negative_cells": (C.c_int, [C.c_void_p, _ip]),
    "anuga_boundary_flux_integral_operator": (C.c_int, [C.c_void_p, C.c_double]),
    "anuga_set_max_flux_update_frequency": (C.c_int, [C.c_void_p, C.c_int64]),
    "anuga_compute_flux_update_frequency": (C.c_int, [C.c_void_p, C.c_double]),
    "anuga_evolve_steps": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(TimeControl),
                                     C.POINTER(StepState), _dp]),
    "anuga_get_boundary_flux_sums": (C.c_int, [C.c_void_p, _dp]),
    "anuga_get_boundary_flux_integral": (C.c_int, [C.c_void_p, _dp]),
    "anuga_get_negative_cells": (C.c_int, [C.c_void_p, _ip]),
    "anuga_get_water_volume": (C.c_int, [C.c_void_p, _dp]),
    "anuga_water_volume_noalloc": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p]),
    "anuga_rcm_order": (C.c_int, [C.c_int64, _ip, _ip]),
    "anuga_build_config": (C.c_int, [_ip, _ip]),
    "anuga_tile_order": (C.c_int, [C.c_int64, _dp, _ip, _ip]),
    "anuga_tile_halos": (C.c_int, [C.c_int64, C.POINTER(C.c_int32), _ip, C.c_int64,
                                   C.POINTER(C.c_int32), C.POINTER(C.c_uint32), _ip, _ip]),
    "anuga_tile_edges": (C.c_int, [C.c_int64, C.POINTER(C.c_int32), _ip, _ip, C.c_int64,
                                   C.POINTER(C.c_uint32), _ip]),
    "anuga_synchronize": (C.c_int, [C.c_void_p]),
    "anuga_kernel_timing": (C.c_int, [C.c_void_p, C.c_char_p, _dp, _ip]),
    "anuga_reset_kernel_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "anuga_ensemble_create": (C.c_int, [C.c_int64, _ip, _ip, C.POINTER(AnugaParams),
                                        C.POINTER(C.c_void_p)]),
    "anuga_ensemble_set_mesh": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(AnugaMesh)]),
    "anuga_ensemble_set_centroid_values": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, _dp]),
    "anuga_ensemble_get_centroid_values": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, _dp]),
    "anuga_ensemble_set_boundary": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int32),
                                              _dp]),
    "anuga_ensemble_evolve_steps": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(TimeControl),
                                              C.POINTER(StepState)]),
    "anuga_ensemble_distribute_to_vertices_and_edges": (C.c_int, [C.c_void_p, _dp]),
    "anuga_ensemble_get_water_volumes": (C.c_int, [C.c_void_p, _dp]),
    "anuga_ensemble_get_boundary_flux_integrals": (C.c_int, [C.c_void_p, _dp]),
    "anuga_ensemble_size": (C.c_int, [C.c_void_p, _ip]),
    "anuga_adjoint_record": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_int64, C.c_int64,
                                       _ip, C.c_int64, _dp]),
    "anuga_adjoint_gradient": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
}

_lib = None
_lock = threading.Lock()


class DeviceError(RuntimeError):
    """A failure reported by the device library."""


def _one_hip_runtime():
    """PyTorch's ROCm wheel bundles its own libamdhip64 / libhsa-runtime64 with
    the same sonames as /opt/rocm's but loads them by file name. If this
    library were loaded first, a later `import torch` would map a second HIP
    and HSA runtime into the process (two clients of the GPU driver with their
    own queues and address apertures). Importing torch first, when it is
    installed, makes this library bind to torch's runtime: one per process."""
    import importlib.util
    import sys
    if "torch" in sys.modules:
        return
    try:
        if importlib.util.find_spec("torch") is not None:
            import torch  # noqa: F401
    except Exception:
        pass


def lib():
    """Load libanuga_hip.so (once). Raises if it is missing or does not load."""
    global _lib
    with _lock:
        if _lib is None:
            _one_hip_runtime()
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    "%s is not built: run `python -m anuga_core_amd.build` (hipcc, gfx950)"
                    % LIB_PATH)
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
        return _lib


def check(handle, rc, what):
    """Turn a library status into the reference's exception types."""
    if rc == ANUGA_OK:
        return
    msg = lib().anuga_last_error(handle)
    msg = msg.decode() if msg else ""
    if rc == ANUGA_E_SEMI_IMPLICIT:
        raise AssertionError(msg or "division by zero in semi implicit update")
    if rc == ANUGA_E_NO_NEIGHBOUR:
        raise Exception(msg or "Internal neighbour not found")
    if rc == ANUGA_E_TIMESTEP:
        raise Exception(msg)
    raise DeviceError("%s failed (%d): %s" % (what, rc, msg))
