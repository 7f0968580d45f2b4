"""Python mirror of the request loop in libbaemm_hip (baemm_server_* in
include/baemm.h): Gauss-Newton-step requests (forward solve + far field,
Neumann data, adjoint derivative of the far-field misfit) admitted by device
memory and run concurrently on their own handles and streams.

No reference counterpart; the numerics are those of FarField,
Derivative_FF_Helper and AdjointDerivative_FF (FarField.hpp:28-428).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import BaemmError

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double


class GNRequest(ctypes.Structure):
    _fields_ = [
        ("coords", _P), ("vertex_count", _I64),
        ("triangles", _P), ("simplex_count", _I64),
        ("mesh_id", _I64),
        ("meas_directions", _P), ("meas_count", _I64),
        ("kappa", _P), ("kappa_count", _I64),
        ("inc_directions", _P), ("wave_chunk_size", _I64),
        ("far_field_data", _P),
        ("wave_type", ctypes.c_int), ("precision", ctypes.c_int),
        ("cg_tol", _D), ("gmres_tol", _D),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("completed", _I64), ("failed", _I64), ("queued", _I64), ("running", _I64), ("max_in_flight", _I64),
        ("resident_bytes", _I64), ("peak_bytes", _I64), ("gmres_iterations", _I64),
        ("p50_ms", _D), ("p99_ms", _D), ("mean_ms", _D), ("solves_per_s", _D),
    ]


def _L():
    return _lib.load()


def _err(L, st):
    raw = L.baemm_last_error(None)
    return BaemmError(st, raw.decode() if raw else "")


class GNServer:
    """Request loop on one device: ``slots`` concurrent requests, admitted while
    their estimated device memory fits into ``hbm_budget_bytes``."""

    def __init__(self, slots=4, hbm_budget_bytes=64 << 30, device=0, policy="fifo", max_skips=8, fusion=False):
        self._L = _L()
        self._s = _P()
        st = self._L.baemm_server_create(int(device), int(slots), int(hbm_budget_bytes), ctypes.byref(self._s))
        if st != _lib.OK:
            raise _err(self._L, st)
        self._pending = {}
        if policy != "fifo":
            self.set_policy(policy, max_skips)
        if fusion:
            self.set_fusion(True)

    def set_fusion(self, on=True):
        """Fused mode: batches of up to ``slots`` queued requests run phase by
        phase, each of their solves one batched GMRES (baemm_server_set_fusion)."""
        st = self._L.baemm_server_set_fusion(self._s, 1 if on else 0)
        if st != _lib.OK:
            raise _err(self._L, st)

    def set_policy(self, policy, max_skips=8):
        """Admission order: "fifo" (oldest first) or "sjf" (least estimated work
        N_t^2 * W first; a request passed over max_skips times goes first)."""
        code = {"fifo": _lib.SERVE_FIFO, "sjf": _lib.SERVE_SHORTEST_FIRST}[policy]
        st = self._L.baemm_server_set_policy(self._s, code, int(max_skips))
        if st != _lib.OK:
            raise _err(self._L, st)

    @staticmethod
    def request_bytes(nv, nt, m, K, wcs, precision="fp32"):
        """Device-memory estimate that admission uses for a request on a fresh handle."""
        prec = {"fp32": _lib.PRECISION_FP32, "mixed": _lib.PRECISION_MIXED,
                "fp32_mfma": _lib.PRECISION_FP32_MFMA}.get(precision, precision)
        rq = GNRequest(None, int(nv), None, int(nt), -1, None, int(m), None, int(K), None, int(wcs), None, 0, int(prec),
                       0.0, 0.0)
        return int(_L().baemm_server_request_bytes(ctypes.byref(rq)))

    def submit(self, V, T, meas, kappa, inc_directions, far_field_data=None, mesh_id=-1, precision="fp32",
               cg_tol=1e-10, gmres_tol=1e-6, wave_type=_lib.WAVE_PLANE):
        """Queues one request; returns its id.  Inputs are copied."""
        V = np.ascontiguousarray(V, np.float64)
        T = np.ascontiguousarray(T, np.int64)
        meas = np.ascontiguousarray(meas, np.float64)
        kappa = np.ascontiguousarray(np.atleast_1d(kappa), np.float64)
        dirs = np.ascontiguousarray(inc_directions, np.float64).reshape(-1, 3)
        K, wcs = kappa.shape[0], dirs.shape[0]
        W = K * wcs
        F = np.zeros((meas.shape[0], W), np.complex128)
        G = np.zeros((V.shape[0], 3), np.float64)
        fobs = None
        if far_field_data is not None:
            fobs = np.ascontiguousarray(far_field_data, np.complex128)
            if fobs.shape != F.shape:
                raise ValueError(f"far_field_data must be {F.shape}")
        prec = {"fp32": _lib.PRECISION_FP32, "mixed": _lib.PRECISION_MIXED,
                "fp32_mfma": _lib.PRECISION_FP32_MFMA}.get(precision, precision)
        rq = GNRequest(V.ctypes.data, V.shape[0], T.ctypes.data, T.shape[0], int(mesh_id), meas.ctypes.data,
                       meas.shape[0], kappa.ctypes.data, K, dirs.ctypes.data, wcs,
                       fobs.ctypes.data if fobs is not None else None, int(wave_type), int(prec), float(cg_tol),
                       float(gmres_tol))
        rid = _I64()
        st = self._L.baemm_server_submit_gn(self._s, ctypes.byref(rq), F.ctypes.data, G.ctypes.data, ctypes.byref(rid))
        if st != _lib.OK:
            raise _err(self._L, st)
        self._pending[rid.value] = (F, G)
        return rid.value

    def wait(self, rid):
        """Blocks until request ``rid`` is done; returns (status, far field
        (m x W), gradient (V x 3), latency ms, GMRES iterations)."""
        lat = _D()
        its = _I64()
        st = self._L.baemm_server_wait(self._s, int(rid), ctypes.byref(lat), ctypes.byref(its))
        F, G = self._pending.pop(rid)
        if st not in (_lib.OK, _lib.ERR_NOCONV):
            raise _err(self._L, st)
        return st, F, G, lat.value, its.value

    def stats(self):
        s = ServerStats()
        st = self._L.baemm_server_stats(self._s, ctypes.byref(s))
        if st != _lib.OK:
            raise _err(self._L, st)
        return {name: getattr(s, name) for name, _ in ServerStats._fields_}

    def close(self):
        if self._s:
            self._L.baemm_server_destroy(self._s)
            self._s = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
