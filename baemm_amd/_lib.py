"""ctypes binding of libbaemm_hip.so (the C ABI declared in include/baemm.h).

The library is built in-tree (``baemm_amd/libbaemm_hip.so``) by
``__graft_entry__.build()``.  There is no fallback: if the library is missing
or fails to load, every call raises ``BaemmError``.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libbaemm_hip.so"

OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_NOCONV, ERR_INTERNAL = range(6)
C128, C64 = 0, 1
PRECISION_FP32, PRECISION_MIXED, PRECISION_FP32_MFMA = 0, 1, 2
WAVE_PLANE, WAVE_RADIAL = 201, 202
SERVE_FIFO, SERVE_SHORTEST_FIRST = 0, 1

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_INT = ctypes.c_int
_D = ctypes.c_double

# name -> (restype, argtypes); mirrors include/baemm.h
SIGNATURES = {
    "baemm_version": (_INT, []),
    "baemm_create": (_INT, [_P, _I64, _P, _I64, _P, _I64, _INT, ctypes.POINTER(_P)]),
    "baemm_create_i32": (_INT, [_P, _I64, _P, _I64, _P, _I64, _INT, ctypes.POINTER(_P)]),
    "baemm_destroy": (_INT, [_P]),
    "baemm_last_error": (ctypes.c_char_p, [_P]),
    "baemm_vertex_count": (_I64, [_P]),
    "baemm_simplex_count": (_I64, [_P]),
    "baemm_meas_count": (_I64, [_P]),
    "baemm_use_diagonal": (_INT, [_P, _INT]),
    "baemm_set_block_size": (_INT, [_P, _INT]),
    "baemm_get_block_size": (_INT, [_P]),
    "baemm_set_wave_chunk_size": (_INT, [_P, _INT]),
    "baemm_get_wave_chunk_size": (_INT, [_P]),
    "baemm_set_precision": (_INT, [_P, _INT]),
    "baemm_get_precision": (_INT, [_P]),
    "baemm_set_segments": (_INT, [_P, _INT]),
    "baemm_set_mixed_kernel": (_INT, [_P, _INT]),
    "baemm_set_gmres_limits": (_INT, [_P, _INT, _INT]),
    "baemm_set_graphs": (_INT, [_P, _INT]),
    "baemm_get_graph_stats": (_INT, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "baemm_get_triangle_data": (_INT, [_P, _P, _P, _P, _P]),
    "baemm_get_coordinates": (_INT, [_P, _P, _P]),
    "baemm_device_info": (_INT, [_P, ctypes.c_char_p, _I64]),
    "baemm_get_flags": (_INT, [_P, _P]),
    "baemm_read_b": (_INT, [_P, _P, _I64, _I64]),
    "baemm_write_b": (_INT, [_P, _P, _I64]),
    "baemm_read_c": (_INT, [_P, _P, _I64, _I64]),
    "baemm_write_c": (_INT, [_P, _P, _I64]),
    "baemm_p0_buffer": (_INT, [_P, _INT, ctypes.POINTER(ctypes.POINTER(ctypes.c_float)), ctypes.POINTER(_I64)]),
    "baemm_mass_nnz": (_I64, [_P]),
    "baemm_get_mass_matrix": (_INT, [_P, _P, _P, _P]),
    "baemm_mass_dot": (_INT, [_P, _D, _P, _I64, _D, _P, _I64, _I64]),
    "baemm_load_parameters": (_INT, [_P, _P, _P, _I64, _I64, _D]),
    "baemm_get_parameters": (_INT, [_P, _P, _P, _P, _P, _P]),
    "baemm_load_boundary_operators": (_INT, [_P, _P, _P, _I64, _I64]),
    "baemm_unload_boundary_operators": (_INT, [_P]),
    "baemm_apply_boundary": (_INT, [_P, _P, _P, _I64, _P, _P, _I64]),
    "baemm_apply_far_field": (_INT, [_P, _P, _P, _I64, _P, _P, _I64]),
    "baemm_apply_near_field": (_INT, [_P, _P, _P, _I64, _P, _P, _I64, _P, _I64]),
    "baemm_create_herglotz_wave": (_INT, [_P, _P, _P, _I64, _P, _P, _I64]),
    "baemm_create_incident_wave": (_INT, [_P, _P, _P, _I64, _P, _P, _I64, _INT]),
    "baemm_autotune_segments": (_INT, [_P, _INT, ctypes.POINTER(_INT), ctypes.POINTER(_D)]),
    "baemm_apply_boundary_dev": (_INT, [_P, _P, _P, _I64, _INT, _P, _P, _I64, _INT, _P]),
    "baemm_apply_far_field_dev": (_INT, [_P, _P, _P, _I64, _INT, _P, _P, _I64, _INT, _P]),
    "baemm_apply_boundary_dev_noalloc": (_INT, [_P, _P, _P, _I64, _INT, _P, _P, _I64, _INT, _P]),
    "baemm_noalloc_scratch_bytes": (_I64, []),
    "baemm_apply_boundary_batch": (_INT, [_INT, _P, _P, _P, _P, _INT, _P, _P, _P, _INT, _P]),
    "baemm_batch_set_tuning": (_INT, [_INT, _INT, _INT, _INT]),
    "baemm_batch_get_tuning": (_INT, [ctypes.POINTER(_INT)] * 4),
    "baemm_batch_set_own_launch": (_INT, [ctypes.c_double]),
    "baemm_batch_get_own_launch": (_INT, [ctypes.POINTER(ctypes.c_double)]),
    "baemm_batch_get_stats": (_INT, [ctypes.POINTER(_I64)] * 4),
    "baemm_batch_autotune": (_INT, [_INT, _P, _P, _P, _INT, _P, _P, _INT, _INT, ctypes.POINTER(_D)]),
    "baemm_server_create": (_INT, [_INT, _INT, _I64, ctypes.POINTER(_P)]),
    "baemm_server_set_policy": (_INT, [_P, _INT, _INT]),
    "baemm_server_set_fusion": (_INT, [_P, _INT]),
    "baemm_server_destroy": (_INT, [_P]),
    "baemm_server_submit_gn": (_INT, [_P, _P, _P, _P, ctypes.POINTER(_I64)]),
    "baemm_server_wait": (_INT, [_P, _I64, ctypes.POINTER(_D), ctypes.POINTER(_I64)]),
    "baemm_server_stats": (_INT, [_P, _P]),
    "baemm_server_request_bytes": (_I64, [_P]),
    "baemm_apply_mass_inverse": (_INT, [_P, _P, _I64, _P, _I64, _I64, _D]),
    "baemm_apply_mass_inverse_complex": (_INT, [_P, _P, _I64, _P, _I64, _I64, _D]),
    "baemm_apply_lumped_mass_inverse": (_INT, [_P, _P, _I64, _P, _I64, _I64]),
    "baemm_boundary_potential": (_INT, [_P, _P, _I64, _P, _P, _P, _P, _I64, _D, _D]),
    "baemm_dirichlet_to_neumann": (_INT, [_P, _P, _I64, _P, _P, _I64, _D, _D]),
    "baemm_far_field": (_INT, [_P, _P, _I64, _P, _I64, _P, _INT, _D, _D]),
    "baemm_far_field_parameters": (_INT, [_P, _P, _I64, _P, _I64, _P, _INT, _P, _D, _D]),
    "baemm_derivative_ff_helper": (_INT, [_P, _P, _I64, _P, _I64, _P, _INT, _D, _D]),
    "baemm_derivative_ff": (_INT, [_P, _P, _I64, _P, _I64, _P, _P, ctypes.POINTER(_P), _INT, _D, _D]),
    "baemm_adjoint_derivative_ff": (_INT, [_P, _P, _I64, _P, _I64, _P, _P, ctypes.POINTER(_P), _INT, _D, _D]),
    "baemm_gauss_newton_solve": (_INT, [_P, _P, _I64, _P, _I64, _P, _P, _P, _P, _D, _P, _I64, _D, _P, _I64,
                                        ctypes.POINTER(_P), _INT, _D, _D, _D, ctypes.POINTER(_INT)]),
    "baemm_gauss_newton_solve_dev": (_INT, [_P, _P, _I64, _P, _I64, _P, _P, _P, _P, _D, _P, _I64, _D, _P, _I64,
                                            _P, _INT, _INT, _D, _D, _D, ctypes.POINTER(_INT), _P]),
    "baemm_mass_metric_apply": (None, [_P, _P, _P, _P]),
    "baemm_mass_metric_jacobi": (None, [_P, _P, _P, _P]),
    "baemm_tp_energy": (_INT, [_P, _D, _D, ctypes.POINTER(_D), _P]),
    "baemm_tp_metric_prepare": (_INT, [_P, _D, _D]),
    "baemm_tp_metric_apply": (None, [_P, _P, _P, _P]),
    "baemm_tp_metric_jacobi": (None, [_P, _P, _P, _P]),
    "baemm_tp_metric_apply_host": (_INT, [_P, _D, _D, _D, _P, _P]),
    "baemm_tp_metric_jacobi_host": (_INT, [_P, _D, _D, _D, _P, _P]),
    "baemm_gmres_dev": (_INT, [_P, _P, _P, _D, _D, ctypes.POINTER(_I64), _P]),
    "baemm_gmres_batch": (_INT, [_INT, _P, _P, _P, _D, _D, _P, _P, _P]),
    "baemm_get_solver_stats": (_INT, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_D),
                                      ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "baemm_reset_solver_stats": (_INT, [_P]),
    "baemm_set_timing": (_INT, [_P, _INT]),
    "baemm_get_timing": (_INT, [_P, ctypes.POINTER(_D), ctypes.POINTER(_I64)]),
    "baemm_hw_sincos": (_INT, [_INT, _P, _I64, _P, _P]),
    "baemm_dev_alloc": (_INT, [ctypes.POINTER(_P), ctypes.c_size_t]),
    "baemm_dev_free": (_INT, [_P]),
    "baemm_dev_upload": (_INT, [_P, _P, ctypes.c_size_t]),
    "baemm_dev_download": (_INT, [_P, _P, ctypes.c_size_t]),
    "baemm_dev_synchronize": (_INT, []),
}


class BaemmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"baemm status {status}: {msg}")
        self.status = status


REAL_OP = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double))


class MassMetric(ctypes.Structure):
    """baemm_mass_metric: the device metric c I + d (Mass (x) I_3)."""
    _fields_ = [("h", ctypes.c_void_p), ("c", ctypes.c_double), ("d", ctypes.c_double)]


class TPMetric(ctypes.Structure):
    """baemm_tp_metric: scale times the tangent-point metric of exponents (q, p)."""
    _fields_ = [("h", ctypes.c_void_p), ("q", ctypes.c_double), ("p", ctypes.c_double), ("scale", ctypes.c_double)]

_lib = None


def load(path: str | os.PathLike | None = None):
    """Load libbaemm_hip.so and attach the ctypes signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise BaemmError(ERR_STATE, f"{p} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(str(p), mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = L
    return L


def check(status, handle=None):
    if status != OK:
        msg = ""
        if handle is not None:
            raw = load().baemm_last_error(handle)
            msg = raw.decode() if raw else ""
        raise BaemmError(status, msg)
