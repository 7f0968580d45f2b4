"""Python mirror of the ``BAEMM::Helmholtz_OpenCL`` class surface over the C ABI.

Method names, argument meaning and order follow the reference
(``Helmholtz_OpenCL.hpp``, ``src/Helmholtz_Common/*.hpp``).  Arrays are numpy:
complex data complex128 ``(rows, W)`` with the leading dimension taken from
the array's row stride, real data float64.  Output arrays are updated in place
(BLAS-style ``alpha``/``beta``) and also returned.

All numerics run in ``libbaemm_hip.so`` on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import BaemmError, check

WaveType = type("WaveType", (), {"Plane": _lib.WAVE_PLANE, "Radial": _lib.WAVE_RADIAL})


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ld(a, what):
    if a.ndim != 2:
        raise ValueError(f"{what} must be 2-D")
    if a.strides[1] != a.itemsize or a.strides[0] % a.itemsize:
        raise ValueError(f"{what} rows must be contiguous")
    return a.strides[0] // a.itemsize


def _cplx(z):
    z = complex(z)
    return (ctypes.c_double * 2)(z.real, z.imag)


def _c128_in(a, what, rows=None):
    a = np.asarray(a)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    a = np.require(a, dtype=np.complex128, requirements=["C"]) if (a.dtype != np.complex128 or a.strides[1] != 16) else a
    if rows is not None and a.shape[0] != rows:
        raise ValueError(f"{what} must have {rows} rows, got {a.shape[0]}")
    return a


def _c128_out(a, what, rows):
    if not isinstance(a, np.ndarray) or a.dtype != np.complex128:
        raise TypeError(f"{what} must be a complex128 numpy array")
    if a.shape[0] != rows:
        raise ValueError(f"{what} must have {rows} rows")
    return a


class Helmholtz_HIP:
    """Drop-in for ``BAEMM::Helmholtz_OpenCL`` (Helmholtz_OpenCL.hpp:34-399)."""

    def __init__(self, vertex_coords, triangles, meas_directions=None, device=0, CPU_thread_count=1):
        self._L = _lib.load()
        V = np.ascontiguousarray(vertex_coords, dtype=np.float64)
        T = np.ascontiguousarray(triangles, dtype=np.int64)
        M = np.zeros((0, 3)) if meas_directions is None else np.ascontiguousarray(meas_directions, dtype=np.float64)
        h = ctypes.c_void_p()
        st = self._L.baemm_create(_ptr(V), V.shape[0], _ptr(T), T.shape[0], _ptr(M), M.shape[0], int(device),
                                  ctypes.byref(h))
        if st != _lib.OK:
            raise BaemmError(st, "baemm_create failed")
        self._h = h
        self.CPU_thread_count = int(CPU_thread_count)
        self._V = V
        self._T = T

    def close(self):
        if getattr(self, "_h", None):
            self._L.baemm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        check(st, self._h)

    # ---- getters / setters (GetSetters.hpp) ----
    def VertexCount(self):
        return int(self._L.baemm_vertex_count(self._h))

    def SimplexCount(self):
        return int(self._L.baemm_simplex_count(self._h))

    def GetMeasCount(self):
        return int(self._L.baemm_meas_count(self._h))

    def ThreadCountCPU(self):
        return self.CPU_thread_count

    def UseDiagonal(self, use_it: bool):
        self._check(self._L.baemm_use_diagonal(self._h, int(bool(use_it))))

    def SetBlockSize(self, n):
        self._check(self._L.baemm_set_block_size(self._h, int(n)))

    def GetBlockSize(self):
        return int(self._L.baemm_get_block_size(self._h))

    def SetWaveChunkSize(self, n):
        self._check(self._L.baemm_set_wave_chunk_size(self._h, int(n)))

    def GetWaveChunkSize(self):
        return int(self._L.baemm_get_wave_chunk_size(self._h))

    def SetPrecision(self, mode):
        m = {"fp32": _lib.PRECISION_FP32, "mixed": _lib.PRECISION_MIXED}.get(mode, mode)
        self._check(self._L.baemm_set_precision(self._h, int(m)))

    def SetSegments(self, n):
        self._check(self._L.baemm_set_segments(self._h, int(n)))

    def _tri_data(self):
        nt = self.SimplexCount()
        mid = np.zeros((nt, 4), np.float32)
        nrm = np.zeros((nt, 4), np.float32)
        area = np.zeros(nt, np.float32)
        sd = np.zeros(nt, np.float32)
        self._check(self._L.baemm_get_triangle_data(self._h, _ptr(mid), _ptr(nrm), _ptr(area), _ptr(sd)))
        return mid, nrm, area, sd

    def TriangleMidpoints(self):
        return self._tri_data()[0]

    def TriangleNormals(self):
        return self._tri_data()[1]

    def TriangleAreas(self):
        return self._tri_data()[2]

    Areas = TriangleAreas

    def SingleLayerDiagonal(self):
        return self._tri_data()[3]

    def MassMatrix(self):
        """CSR (indptr, indices, data) of the P1 mass matrix."""
        nnz = int(self._L.baemm_mass_nnz(self._h))
        nv = self.VertexCount()
        ptr = np.zeros(nv + 1, np.int32)
        idx = np.zeros(nnz, np.int32)
        val = np.zeros(nnz, np.float64)
        self._check(self._L.baemm_get_mass_matrix(self._h, _ptr(ptr), _ptr(idx), _ptr(val)))
        return ptr, idx, val

    def MassDot(self, alpha, X, beta, Y):
        """MassMatrix().Dot(alpha, X, ldX, beta, Y, ldY, nrhs) on real columns."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        X2 = X.reshape(X.shape[0], -1)
        Y2 = Y.reshape(Y.shape[0], -1)
        self._check(self._L.baemm_mass_dot(self._h, float(alpha), _ptr(X2), X2.shape[1], float(beta), _ptr(Y2),
                                           _ld(Y2, "Y"), X2.shape[1]))
        return Y

    # ---- parameters (LoadParameters.hpp) ----
    def LoadParameters(self, kappa_list, coeff_list, wave_count, wave_chunk_size, factor=1.0):
        k = np.ascontiguousarray(kappa_list, dtype=np.float64).ravel()
        c = np.ascontiguousarray(coeff_list, dtype=np.complex128).ravel()
        K = -(-int(wave_count) // int(wave_chunk_size))
        if k.size < K or c.size < 4 * K:
            raise ValueError("kappa_list needs K entries and coeff_list 4K entries")
        self._check(self._L.baemm_load_parameters(self._h, _ptr(k), _ptr(c), int(wave_count), int(wave_chunk_size),
                                                  float(factor)))

    def GetWaveCount(self):
        W = ctypes.c_int64()
        self._check(self._L.baemm_get_parameters(self._h, None, None, ctypes.byref(W), None, None))
        return W.value

    def GetWaveNumbers(self):
        K = ctypes.c_int64()
        self._check(self._L.baemm_get_parameters(self._h, None, None, None, None, ctypes.byref(K)))
        k = np.zeros(K.value, np.float32)
        self._check(self._L.baemm_get_parameters(self._h, _ptr(k), None, None, None, None))
        return k

    def GetCoefficients(self):
        K = ctypes.c_int64()
        self._check(self._L.baemm_get_parameters(self._h, None, None, None, None, ctypes.byref(K)))
        c = np.zeros((K.value, 4), np.complex64)
        self._check(self._L.baemm_get_parameters(self._h, None, _ptr(c), None, None, None))
        return c

    def LoadBoundaryOperators_PL(self, kappa_list, coeff_list, wave_count, wave_chunk_size):
        self.LoadParameters(kappa_list, coeff_list, wave_count, wave_chunk_size, 1.0)

    def UnloadBoundaryOperators_PL(self):
        self._check(self._L.baemm_unload_boundary_operators(self._h))

    # ---- applies ----
    def ApplyBoundaryOperators_PL(self, alpha, B_in, beta, C_out, kappa_list=None, coeff_list=None,
                                  wave_count=None, wave_chunk_size=None):
        """ApplyBoundaryOperators.hpp:26-43 (with parameters) and :77-178 (pre-loaded)."""
        if kappa_list is not None:
            self.LoadBoundaryOperators_PL(kappa_list, coeff_list, wave_count, wave_chunk_size)
        nv = self.VertexCount()
        B = _c128_in(B_in, "B_in", nv)
        C = _c128_out(C_out, "C_out", nv)
        self._check(self._L.baemm_apply_boundary(self._h, _cplx(alpha), _ptr(B), _ld(B, "B_in"), _cplx(beta),
                                                 _ptr(C), _ld(C, "C_out")))
        if kappa_list is not None:
            self.UnloadBoundaryOperators_PL()
        return C

    def ApplyFarFieldOperators_PL(self, alpha, B_in, beta, C_out, kappa_list=None, coeff_list=None,
                                  wave_count=None, wave_chunk_size=None):
        """ApplyFarFieldOperators.hpp:27-104."""
        if kappa_list is not None:
            self.LoadParameters(kappa_list, coeff_list, wave_count, wave_chunk_size)
        B = _c128_in(B_in, "B_in", self.VertexCount())
        C = _c128_out(C_out, "C_out", self.GetMeasCount())
        self._check(self._L.baemm_apply_far_field(self._h, _cplx(alpha), _ptr(B), _ld(B, "B_in"), _cplx(beta),
                                                  _ptr(C), _ld(C, "C_out")))
        return C

    def ApplyNearFieldOperators_PL(self, alpha, B_in, beta, C_out, kappa_list=None, coeff_list=None,
                                   wave_count=None, wave_chunk_size=None, evaluation_points=None):
        """ApplyNearFieldOperators.hpp:132-226."""
        if kappa_list is not None:
            self.LoadParameters(kappa_list, coeff_list, wave_count, wave_chunk_size)
        pts = np.ascontiguousarray(evaluation_points, dtype=np.float64).reshape(-1, 3)
        B = _c128_in(B_in, "B_in", self.VertexCount())
        C = _c128_out(C_out, "C_out", pts.shape[0])
        self._check(self._L.baemm_apply_near_field(self._h, _cplx(alpha), _ptr(B), _ld(B, "B_in"), _cplx(beta),
                                                   _ptr(C), _ld(C, "C_out"), _ptr(pts), pts.shape[0]))
        return C

    def CreateHerglotzWave_PL(self, alpha, B_in, beta, C_out, kappa_list=None, coeff_list=None,
                              wave_count=None, wave_chunk_size=None):
        """CreateHerglotzWave.hpp:152-222 (factor 4 pi / m)."""
        if kappa_list is not None:
            self.LoadParameters(kappa_list, coeff_list, wave_count, wave_chunk_size,
                                4.0 * np.pi / self.GetMeasCount())
        g = _c128_in(B_in, "B_in", self.GetMeasCount())
        C = _c128_out(C_out, "C_out", self.VertexCount())
        self._check(self._L.baemm_create_herglotz_wave(self._h, _cplx(alpha), _ptr(g), _ld(g, "B_in"), _cplx(beta),
                                                       _ptr(C), _ld(C, "C_out")))
        return C

    def CreateIncidentWave_PL(self, alpha, incident_directions, beta, C_out, kappa_list=None, coeff_list=None,
                              wave_count=None, wave_chunk_size=None, type=WaveType.Plane, inc_count=None):
        """CreateIncidentWave.hpp:22-127 (factor 4 pi)."""
        if kappa_list is not None:
            self.LoadParameters(kappa_list, coeff_list, wave_count, wave_chunk_size, 4.0 * np.pi)
        d = np.ascontiguousarray(incident_directions, dtype=np.float64).reshape(-1, 3)
        n = d.shape[0] if inc_count is None else int(inc_count)
        C = _c128_out(C_out, "C_out", self.VertexCount())
        self._check(self._L.baemm_create_incident_wave(self._h, _cplx(alpha), _ptr(d), n, _cplx(beta), _ptr(C),
                                                       _ld(C, "C_out"), int(type)))
        return C

    # ---- timing of the dominant pair kernel ----
    def SetTiming(self, on: bool):
        self._check(self._L.baemm_set_timing(self._h, int(bool(on))))

    def GetTiming(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self._L.baemm_get_timing(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
