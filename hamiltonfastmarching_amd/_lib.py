"""ctypes binding of libhfm_amd.so (the C ABI declared in include/hfm_amd.h).

The library is the only compute path: if it is missing, or no GPU is visible when a compute
entry point is called, this module raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libhfm_amd.so")

c_int, c_double, c_int64, c_size_t, c_void_p = (ctypes.c_int, ctypes.c_double, ctypes.c_int64,
                                                ctypes.c_size_t, ctypes.c_void_p)
P = ctypes.POINTER


class HfmPlanStats(ctypes.Structure):
    _fields_ = [("iterations", c_int64), ("tile_visits", c_int64), ("sweeps", c_int64),
                ("max_stencil_width", c_int64), ("n_points", c_int64), ("n_tiles", c_int64),
                ("n_deps", c_int64), ("stencil_ms", c_double), ("solve_ms", c_double),
                ("fim_ms", c_double), ("fim_launches", c_int64), ("tiles_carried", c_int64),
                ("delta", c_double), ("adjoint_rounds", c_int64),
                ("forward_rounds", c_int64), ("n_accepted", c_int64)]


class HfmRequest(ctypes.Structure):
    _fields_ = [("model", ctypes.c_char_p), ("dims", c_int64 * 3), ("gridScale", c_double),
                ("metric", c_void_p), ("metric_is_field", c_int), ("metric_on_device", c_int), ("dual", c_int),
                ("nseeds", c_int), ("seeds", P(c_double)), ("seed_values", P(c_double)),
                ("values_out", P(c_double))]


class HfmRequestResult(ctypes.Structure):
    _fields_ = [("status", c_int), ("latency_ms", c_double), ("queue_ms", c_double), ("setup_ms", c_double),
                ("rounds", c_int64), ("tile_visits", c_int64), ("error", ctypes.c_char * 256)]


class HfmServerStats(ctypes.Structure):
    _fields_ = [("submitted", c_int64), ("completed", c_int64), ("failed", c_int64), ("resident", c_int64),
                ("max_resident_seen", c_int64), ("batches", c_int64), ("rounds", c_int64),
                ("hbm_in_use", c_double), ("hbm_budget", c_double)]


# (name, restype, argtypes) of every exported symbol; tests check that each one is exported.
SIGNATURES = [
    ("hfm_io_create", c_void_p, []),
    ("hfm_io_destroy", None, [c_void_p]),
    ("hfm_io_set_string", c_int, [c_void_p, ctypes.c_char_p, ctypes.c_char_p]),
    ("hfm_io_set_scalar", c_int, [c_void_p, ctypes.c_char_p, c_double]),
    ("hfm_io_set_array", c_int, [c_void_p, ctypes.c_char_p, P(c_double), c_int, P(c_int64)]),
    ("hfm_io_set_array_f32", c_int, [c_void_p, ctypes.c_char_p, P(ctypes.c_float), c_int, P(c_int64)]),
    ("hfm_io_has_field", c_int, [c_void_p, ctypes.c_char_p]),
    ("hfm_io_erase_field", c_int, [c_void_p, ctypes.c_char_p]),
    ("hfm_io_field_kind", c_int, [c_void_p, ctypes.c_char_p]),
    ("hfm_io_get_string", ctypes.c_longlong, [c_void_p, ctypes.c_char_p, ctypes.c_char_p, c_size_t]),
    ("hfm_io_get_scalar", c_int, [c_void_p, ctypes.c_char_p, P(c_double)]),
    ("hfm_io_get_array_shape", c_int, [c_void_p, ctypes.c_char_p, P(c_int), P(c_int64), c_int]),
    ("hfm_io_get_array", c_int, [c_void_p, ctypes.c_char_p, P(c_double), c_size_t]),
    ("hfm_io_computed_keys", ctypes.c_longlong, [c_void_p, ctypes.c_char_p, c_size_t]),
    ("hfm_io_last_error", ctypes.c_char_p, [c_void_p]),
    ("hfm_io_messages", ctypes.c_char_p, [c_void_p]),
    ("hfm_io_read_files", c_int, [c_void_p, ctypes.c_char_p]),
    ("hfm_io_write_files", c_int, [c_void_p, ctypes.c_char_p]),
    ("hfm_run", c_int, [c_void_p]),
    ("hfm_plan_create", c_int, [P(c_void_p), ctypes.c_char_p, P(c_int64), P(c_double), P(c_double), c_int,
                                c_double, c_double]),
    ("hfm_plan_destroy", None, [c_void_p]),
    ("hfm_plan_set_metric", c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
    ("hfm_plan_set_cost", c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    ("hfm_plan_solve", c_int, [c_void_p, c_int, P(c_double), P(c_double), c_void_p]),
    ("hfm_plan_values", c_int, [c_void_p, c_void_p, c_void_p]),
    ("hfm_riemann_solve_noalloc", c_int, [ctypes.c_char_p, P(c_int64), P(c_double), P(c_double), c_void_p, c_int,
                                          c_int, c_int, P(c_double), P(c_double), c_void_p, c_void_p]),
    ("hfm_plan_load_values", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("hfm_plan_flow", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("hfm_plan_active_offsets", c_int, [c_void_p, c_void_p, c_void_p]),
    ("hfm_plan_geodesics", c_int, [c_void_p, c_int, P(c_double), c_double, c_double, c_double, P(c_double),
                                   c_int64, P(c_int64), P(c_int64), P(c_int), c_void_p]),
    ("hfm_plan_geodesics_ode", c_int, [c_void_p, c_int, P(c_double), c_double, c_double, P(c_double), c_int64,
                                       P(c_int64), P(c_int64), P(c_int), c_void_p]),
    ("hfm_plan_n_active_slots", c_int, [c_void_p]),
    ("hfm_plan_residual", c_int, [c_void_p, P(c_double), P(c_int64), c_void_p]),
    ("hfm_plan_adjoint", c_int, [c_void_p, c_int, P(c_double), P(c_double), c_void_p, P(c_int64), P(c_double),
                                 c_int, P(c_int), c_void_p]),
    ("hfm_plan_forward_variation", c_int, [c_void_p, c_int, c_void_p, c_int, P(c_double), c_void_p, c_void_p]),
    ("hfm_plan_get_stats", c_int, [c_void_p, P(HfmPlanStats)]),
    ("hfm_plan_set_timing", c_int, [c_void_p, c_int]),
    ("hfm_plan_set_graph", c_int, [c_void_p, c_int]),
    ("hfm_plan_set_fused_schedule", c_int, [c_void_p, c_int]),
    ("hfm_plan_set_order", c_int, [c_void_p, c_int]),
    ("hfm_plan_set_stopping", c_int, [c_void_p, c_int64, c_double, c_int, P(c_int64), c_int, P(c_int64), c_int,
                                      P(c_int64), c_int, P(c_int64)]),
    ("hfm_plan_set_factoring", c_int, [c_void_p, P(c_int64), P(c_int64), P(c_double), P(c_double),
                                       P(ctypes.c_uint8)]),
    ("hfm_plan_set_delta", c_int, [c_void_p, c_double]),
    ("hfm_plan_set_round_target", c_int, [c_void_p, c_int]),
    ("hfm_plan_set_grid", c_int, [c_void_p, c_int]),
    ("hfm_plan_set_max_sweeps", c_int, [c_void_p, c_int]),
    ("hfm_plan_set_walls", c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    ("hfm_plan_get_stencils", c_int, [c_void_p, P(ctypes.c_int8), P(c_double)]),
    ("hfm_server_create", c_int, [P(c_void_p), c_double, c_int]),
    ("hfm_server_submit", c_int, [c_void_p, P(HfmRequest), P(c_int64)]),
    ("hfm_server_poll", c_int, [c_void_p, c_int64, P(c_int)]),
    ("hfm_server_wait", c_int, [c_void_p, c_int64, c_double, P(HfmRequestResult)]),
    ("hfm_server_get_stats", c_int, [c_void_p, P(HfmServerStats)]),
    ("hfm_server_set_persistent", c_int, [c_void_p, c_int]),
    ("hfm_server_destroy", None, [c_void_p]),
    ("hfm_last_error", ctypes.c_char_p, []),
]

_lib = None


def load() -> ctypes.CDLL:
    """Load libhfm_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libhfm_amd.so is not built ({LIB_PATH}); run `python -m hamiltonfastmarching_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    return load().hfm_last_error().decode()
