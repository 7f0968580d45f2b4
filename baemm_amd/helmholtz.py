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


class _HelmholtzBase:
    """Construction, getters/setters and operator applies."""

    def __init__(self, vertex_coords, triangles, meas_directions=None, device=0, CPU_thread_count=1):
        self._L = _lib.load()
        V = np.ascontiguousarray(vertex_coords, dtype=np.float64)
        T = np.ascontiguousarray(triangles, dtype=np.int64)
        M = np.zeros((0, 3)) if meas_directions is None else np.ascontiguousarray(meas_directions, dtype=np.float64)
        h = ctypes.c_void_p()
        st = self._L.baemm_create(_ptr(V), V.shape[0], _ptr(T), T.shape[0], _ptr(M), M.shape[0], int(device),
                                  ctypes.byref(h))
        if st != _lib.OK:
            raw = self._L.baemm_last_error(None)
            raise BaemmError(st, "baemm_create failed: " + (raw.decode() if raw else ""))
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
        m = {"fp32": _lib.PRECISION_FP32, "mixed": _lib.PRECISION_MIXED,
             "fp32_mfma": _lib.PRECISION_FP32_MFMA}.get(mode, mode)
        self._check(self._L.baemm_set_precision(self._h, int(m)))

    def SetMixedKernel(self, variant):
        """MFMA kernel variant of the mixed precision: 0 single-wave, 1 wave-specialised,
        2 automatic (the default)."""
        self._check(self._L.baemm_set_mixed_kernel(self._h, int(variant)))

    def AutotuneSegments(self, repeats=3):
        """Times source-segment counts of the loaded operator and keeps the fastest;
        returns (segments, milliseconds per apply).  No reference counterpart."""
        seg, ms = ctypes.c_int(), ctypes.c_double()
        self._check(self._L.baemm_autotune_segments(self._h, int(repeats), ctypes.byref(seg), ctypes.byref(ms)))
        return seg.value, ms.value

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

    def VertexCoordinates(self):
        """Helmholtz_OpenCL.hpp:333-336: the vertex coordinates in FP32 (V x 3)."""
        v = np.zeros((self.VertexCount(), 3), np.float32)
        self._check(self._L.baemm_get_coordinates(self._h, _ptr(v), None))
        return v

    def TriangleCoordinates(self):
        """Helmholtz_OpenCL.hpp:338-341: the three vertices of each triangle as
        x, y, z, 0 (T x 12, FP32; Initialize.hpp:82-93)."""
        t = np.zeros((self.SimplexCount(), 12), np.float32)
        self._check(self._L.baemm_get_coordinates(self._h, None, _ptr(t)))
        return t

    def DeviceInfo(self):
        """Helmholtz_OpenCL.hpp:343-391: the handle's GPU as text."""
        buf = ctypes.create_string_buffer(8192)
        self._check(self._L.baemm_device_info(self._h, buf, len(buf)))
        return buf.value.decode()

    def ClassName(self):
        return "Helmholtz_HIP"

    _FLAG_NAMES = ("Re_mass_matrix", "Im_mass_matrix", "Re_single_layer", "Im_single_layer",
                   "Re_double_layer", "Im_double_layer", "Re_adjdbl_layer", "Im_adjdbl_layer")

    def PrintBooleans(self):
        """LoadParameters.hpp:85-95: prints the flags of the loaded coefficient
        columns (nonzero real / imaginary parts over all chunks) and returns them."""
        f = (ctypes.c_int * 8)()
        self._check(self._L.baemm_get_flags(self._h, f))
        flags = {n: bool(f[j]) for j, n in enumerate(self._FLAG_NAMES)}
        for n, v in flags.items():
            print(f"{n} = {'true' if v else 'false'}")
        return flags

    # ---- P0 buffers (InputOutput.hpp:5-89) ----
    def _p0_width(self, which):
        data = ctypes.POINTER(ctypes.c_float)()
        ld = ctypes.c_int64()
        self._check(self._L.baemm_p0_buffer(self._h, int(which), ctypes.byref(data), ctypes.byref(ld)))
        return data, ld.value

    def _p0_write(self, fn, which):
        _, W = self._p0_width(which)
        out = np.zeros((self.SimplexCount(), W), np.complex64)
        self._check(fn(self._h, _ptr(out), W))
        return out

    def _p0_read(self, fn, data):
        a = np.ascontiguousarray(data, dtype=np.complex64)
        if a.ndim != 2 or a.shape[0] != self.SimplexCount():
            raise ValueError("input must be simplex_count x wave_count")
        self._check(fn(self._h, _ptr(a), a.shape[1], a.shape[1]))

    def WriteB(self):
        """B = Av B_in of the last operator apply, simplex_count x wave_count."""
        return self._p0_write(self._L.baemm_write_b, 0)

    def WriteC(self):
        """C = (K + D) b of the last boundary-operator apply (before AvT)."""
        return self._p0_write(self._L.baemm_write_c, 1)

    def ReadB(self, data):
        self._p0_read(self._L.baemm_read_b, data)

    def ReadC(self, data):
        self._p0_read(self._L.baemm_read_c, data)

    def _p0_elem(self, which, i, k):
        data, W = self._p0_width(which)
        if not (0 <= int(i) < self.SimplexCount() and 0 <= int(k) < W):
            raise IndexError("P0 buffer index out of range")
        o = 2 * (W * int(i) + int(k))
        return complex(data[o], data[o + 1])

    def B(self, i, k):
        """B(i, k) of the P0 buffer B (a value; the C++ header returns a reference)."""
        return self._p0_elem(0, i, k)

    def C(self, i, k):
        return self._p0_elem(1, i, k)

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
        self._wave_count_set = None
        self._check(self._L.baemm_load_parameters(self._h, _ptr(k), _ptr(c), int(wave_count), int(wave_chunk_size),
                                                  float(factor)))

    def GetWaveCount(self):
        if getattr(self, "_wave_count_set", None) is not None:
            return self._wave_count_set
        W = ctypes.c_int64()
        self._check(self._L.baemm_get_parameters(self._h, None, None, ctypes.byref(W), None, None))
        return W.value

    def SetWaveCount(self, wave_count):
        """GetSetters.hpp:40-45: the wave count GetWaveCount reports until the
        next LoadParameters (which sets it, as the reference's does)."""
        self._wave_count_set = int(wave_count)

    def GetWaveChunkCount(self, wave_count):
        """GetSetters.hpp:59-62."""
        wcs = self.GetWaveChunkSize()
        return -(-int(wave_count) // wcs)

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


def _libc_free(ptr):
    ctypes.CDLL(None).free(ctypes.c_void_p(ptr))


class _SolverMixin:
    """Solver compositions of FarField.hpp on the device (see include/baemm.h)."""

    def SetGraphs(self, on: bool):
        """Replays each GMRES Arnoldi step from a captured hipGraph (no reference counterpart)."""
        self._check(self._L.baemm_set_graphs(self._h, 1 if on else 0))

    def GraphStats(self):
        c, l = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._L.baemm_get_graph_stats(self._h, ctypes.byref(c), ctypes.byref(l)))
        return {"captures": c.value, "launches": l.value}

    def SetGmresLimits(self, max_iter=30, max_restarts=3):
        """gmres_max_iter / gmres_max_restarts (MemberVariables.hpp:64-65)."""
        self._check(self._L.baemm_set_gmres_limits(self._h, int(max_iter), int(max_restarts)))

    def SolverStats(self):
        it, cyc, cg, ap = (ctypes.c_int64() for _ in range(4))
        res = ctypes.c_double()
        self._check(self._L.baemm_get_solver_stats(self._h, ctypes.byref(it), ctypes.byref(cyc), ctypes.byref(res),
                                                   ctypes.byref(cg), ctypes.byref(ap)))
        return {"gmres_iterations": it.value, "gmres_cycles": cyc.value, "rel_residual": res.value,
                "cg_iterations": cg.value, "operator_applies": ap.value}

    def ResetSolverStats(self):
        self._check(self._L.baemm_reset_solver_stats(self._h))

    def _solve_status(self, st):
        if st not in (_lib.OK, _lib.ERR_NOCONV):
            self._check(st)
        return st == _lib.OK

    def SolveLoaded(self, B, cg_tol=1e-10, gmres_tol=1e-6):
        """X with (P A) X = P B for the loaded boundary operator A (the GMRES of the
        solvers, baemm_gmres_dev); returns (X, iterations, converged).  No
        reference counterpart."""
        L = self._L
        W = self.GetWaveCount()
        Bh = np.ascontiguousarray(_c128_in(B, "B", self.VertexCount()))
        if Bh.shape[1] != W:
            raise ValueError(f"B has {Bh.shape[1]} columns, the loaded wave count is {W}")
        dB, dX = _DevArray(L, Bh), _DevArray(L, np.zeros_like(Bh))
        try:
            its = ctypes.c_int64(0)
            st = L.baemm_gmres_dev(self._h, dB.ptr, dX.ptr, float(cg_tol), float(gmres_tol), ctypes.byref(its), None)
            if st not in (_lib.OK, _lib.ERR_NOCONV):
                self._check(st)
            X = np.empty_like(Bh)
            dX.download(X)
            return X, int(its.value), st == _lib.OK
        finally:
            dB.close()
            dX.close()

    def ApplyMassInverse(self, X, Y, cg_tol):
        """ApplyMassInverse.hpp:6-156: Y = M^{-1} X (real or complex columns)."""
        X = np.asarray(X)
        if np.iscomplexobj(X):
            Xc = np.require(X.reshape(X.shape[0], -1), dtype=np.complex128, requirements=["C"])
            Yc = Y.reshape(Y.shape[0], -1)
            self._check(self._L.baemm_apply_mass_inverse_complex(self._h, _ptr(Xc), Xc.shape[1], _ptr(Yc),
                                                                 _ld(Yc, "Y"), Xc.shape[1], float(cg_tol)))
        else:
            Xr = np.require(X.reshape(X.shape[0], -1), dtype=np.float64, requirements=["C"])
            Yr = Y.reshape(Y.shape[0], -1)
            self._check(self._L.baemm_apply_mass_inverse(self._h, _ptr(Xr), Xr.shape[1], _ptr(Yr), _ld(Yr, "Y"),
                                                         Xr.shape[1], float(cg_tol)))
        return Y

    def ApplyLumpedMassInverse(self, X, Y):
        Xr = np.require(np.asarray(X, dtype=np.float64).reshape(len(X), -1), requirements=["C"])
        Yr = Y.reshape(Y.shape[0], -1)
        self._check(self._L.baemm_apply_lumped_mass_inverse(self._h, _ptr(Xr), Xr.shape[1], _ptr(Yr), _ld(Yr, "Y"),
                                                            Xr.shape[1]))
        return Y

    @staticmethod
    def _kd(kappa, dirs, wcs):
        k = np.ascontiguousarray(kappa, dtype=np.float64).ravel()
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        if d.shape[0] < wcs:
            raise ValueError("need wave_chunk_size directions")
        return k, d

    def BoundaryPotential(self, kappa, wave, wave_chunk_size, cg_tol, gmres_tol, eta=None):
        """FarField.hpp:626-771; returns (phi, coeff rows, converged)."""
        k = np.ascontiguousarray(kappa, dtype=np.float64).ravel()
        wave = np.require(wave, dtype=np.complex128, requirements=["C"])
        phi = np.zeros_like(wave)
        coeff = np.zeros((k.size, 4), np.complex128)
        e = None if eta is None else np.ascontiguousarray(eta, dtype=np.float64)
        st = self._L.baemm_boundary_potential(self._h, _ptr(k), k.size, None if e is None else _ptr(e), _ptr(coeff),
                                              _ptr(wave), _ptr(phi), int(wave_chunk_size), float(cg_tol),
                                              float(gmres_tol))
        return phi, coeff, self._solve_status(st)

    def DirichletToNeumann(self, kappa, wave, wave_chunk_size, cg_tol, gmres_tol):
        """FarField.hpp:790-879; returns (neumann_trace, converged)."""
        k = np.ascontiguousarray(kappa, dtype=np.float64).ravel()
        wave = np.require(wave, dtype=np.complex128, requirements=["C"])
        out = np.zeros_like(wave)
        st = self._L.baemm_dirichlet_to_neumann(self._h, _ptr(k), k.size, _ptr(wave), _ptr(out),
                                                int(wave_chunk_size), float(cg_tol), float(gmres_tol))
        return out, self._solve_status(st)

    def FarField(self, kappa, wave_chunk_count, inc_directions, wave_chunk_size, Y_out=None,
                 type=WaveType.Plane, cg_tol=1e-5, gmres_tol=1e-4, eta=None):
        """FarField / FarField_parameters (FarField.hpp:28-144)."""
        k, d = self._kd(kappa, inc_directions, wave_chunk_size)
        W = int(wave_chunk_count) * int(wave_chunk_size)
        Y = np.zeros((self.GetMeasCount(), W), np.complex128) if Y_out is None else Y_out
        e = None if eta is None else np.ascontiguousarray(eta, dtype=np.float64)
        st = self._L.baemm_far_field_parameters(self._h, _ptr(k), int(wave_chunk_count), _ptr(d),
                                                int(wave_chunk_size), _ptr(Y), int(type),
                                                None if e is None else _ptr(e), float(cg_tol), float(gmres_tol))
        self._solve_status(st)
        return Y

    def Derivative_FF_Helper(self, kappa, wave_chunk_count, inc_directions, wave_chunk_size,
                             type=WaveType.Plane, cg_tol=1e-5, gmres_tol=1e-4):
        k, d = self._kd(kappa, inc_directions, wave_chunk_size)
        W = int(wave_chunk_count) * int(wave_chunk_size)
        du = np.zeros((self.VertexCount(), W), np.complex128)
        st = self._L.baemm_derivative_ff_helper(self._h, _ptr(k), int(wave_chunk_count), _ptr(d),
                                                int(wave_chunk_size), _ptr(du), int(type), float(cg_tol),
                                                float(gmres_tol))
        self._solve_status(st)
        return du

    def _du_arg(self, du_dn):
        if du_dn is None:
            return ctypes.c_void_p(None), None
        du = np.require(du_dn, dtype=np.complex128, requirements=["C"])
        return ctypes.c_void_p(du.ctypes.data), du

    def _du_result(self, pp, keep, W):
        if keep is not None:
            return keep
        nv = self.VertexCount()
        arr = np.ctypeslib.as_array(ctypes.cast(pp, ctypes.POINTER(ctypes.c_double)), shape=(nv * W * 2,))
        du = arr.view(np.complex128).reshape(nv, W).copy()
        _libc_free(pp.value)
        return du

    def Derivative_FF(self, kappa, wave_chunk_count, inc_directions, wave_chunk_size, X_in, Y_out=None, du_dn=None,
                      type=WaveType.Plane, cg_tol=1e-5, gmres_tol=1e-4):
        """FarField.hpp:227-321; returns (Y_out, du_dn)."""
        k, d = self._kd(kappa, inc_directions, wave_chunk_size)
        W = int(wave_chunk_count) * int(wave_chunk_size)
        X = np.require(X_in, dtype=np.float64, requirements=["C"]).reshape(-1, 3)
        Y = np.zeros((self.GetMeasCount(), W), np.complex128) if Y_out is None else Y_out
        pp, keep = self._du_arg(du_dn)
        st = self._L.baemm_derivative_ff(self._h, _ptr(k), int(wave_chunk_count), _ptr(d), int(wave_chunk_size),
                                         _ptr(X), _ptr(Y), ctypes.byref(pp), int(type), float(cg_tol),
                                         float(gmres_tol))
        self._solve_status(st)
        return Y, self._du_result(pp, keep, W)

    def AdjointDerivative_FF(self, kappa, wave_chunk_count, inc_directions, wave_chunk_size, g, Y_out=None,
                             du_dn=None, type=WaveType.Plane, cg_tol=1e-5, gmres_tol=1e-4):
        """FarField.hpp:342-428; returns (Y_out (V x 3), du_dn)."""
        k, d = self._kd(kappa, inc_directions, wave_chunk_size)
        W = int(wave_chunk_count) * int(wave_chunk_size)
        gg = np.require(g, dtype=np.complex128, requirements=["C"])
        Y = np.zeros((self.VertexCount(), 3)) if Y_out is None else Y_out
        pp, keep = self._du_arg(du_dn)
        st = self._L.baemm_adjoint_derivative_ff(self._h, _ptr(k), int(wave_chunk_count), _ptr(d),
                                                 int(wave_chunk_size), _ptr(gg), _ptr(Y), ctypes.byref(pp),
                                                 int(type), float(cg_tol), float(gmres_tol))
        self._solve_status(st)
        return Y, self._du_result(pp, keep, W)

    def GaussNewtonSolve(self, kappa, wave_chunk_count, inc_directions, wave_chunk_size, M, P, alpha, B, beta, X,
                         du_dn=None, type=WaveType.Plane, cg_tol=1e-5, gmres_tol_inner=5e-3, gmres_tol_outer=1e-2):
        """FarField.hpp:457-604.  M, P: callables (x, y) on (V x 3) float64 arrays,
        writing y in place.  Returns (succeeded, du_dn)."""
        k, d = self._kd(kappa, inc_directions, wave_chunk_size)
        W = int(wave_chunk_count) * int(wave_chunk_size)
        nv = self.VertexCount()

        def wrap(fn):
            def cb(user, x, y):
                xa = np.ctypeslib.as_array(x, shape=(nv, 3))
                ya = np.ctypeslib.as_array(y, shape=(nv, 3))
                fn(xa, ya)
            return _lib.REAL_OP(cb)

        Mc, Pc = wrap(M), wrap(P)
        Bv = B
        if not (isinstance(B, np.ndarray) and B.dtype == np.float64 and B.ndim == 2 and B.strides[1] == 8):
            Bv = np.require(B, dtype=np.float64, requirements=["C"])   # keeps the row stride (ldB) otherwise
        pp, keep = self._du_arg(du_dn)
        ok = ctypes.c_int(0)
        st = self._L.baemm_gauss_newton_solve(self._h, _ptr(k), int(wave_chunk_count), _ptr(d), int(wave_chunk_size),
                                              ctypes.cast(Mc, ctypes.c_void_p), None,
                                              ctypes.cast(Pc, ctypes.c_void_p), None, float(alpha), _ptr(Bv),
                                              _ld(Bv, "B"), float(beta), _ptr(X), _ld(X, "X"), ctypes.byref(pp),
                                              int(type), float(cg_tol), float(gmres_tol_inner),
                                              float(gmres_tol_outer), ctypes.byref(ok))
        self._check(st)
        return bool(ok.value), self._du_result(pp, keep, W)


    def GaussNewtonSolveDevice(self, kappa, wave_chunk_count, inc_directions, wave_chunk_size, c, d, alpha, B, beta,
                               X, du_dn=None, type=WaveType.Plane, cg_tol=1e-5, gmres_tol_inner=5e-3,
                               gmres_tol_outer=1e-2, tangent_point=None):
        """GaussNewtonSolve with every vector on the device
        (baemm_gauss_newton_solve_dev) and a device metric with its Jacobi
        preconditioner: M = c I + d (Mass (x) I_3) (baemm_mass_metric_apply /
        _jacobi), or, with tangent_point = (q, p, regpar), M = regpar times the
        tangent-point metric as in Example_GaussNewton/main.cpp:301-346
        (baemm_tp_metric_apply / _jacobi; c and d are then unused).
        B, X: V x 3 float64 views (ld >= 3 allowed); X is updated in place.
        du_dn: V x W complex or None (then computed).  Returns (succeeded, du_dn)."""
        k, dr = self._kd(kappa, inc_directions, wave_chunk_size)
        W = int(wave_chunk_count) * int(wave_chunk_size)
        nv = self.VertexCount()
        L = self._L
        ldB, ldX = _ld(B, "B"), _ld(X, "X")
        # device copies with the callers' leading dimensions (columns 3.. unused)
        Bfull = np.zeros((nv, ldB))
        Bfull[:, :3] = B[:, :3]
        Xfull = np.zeros((nv, ldX))
        Xfull[:, :3] = X[:, :3]
        du = np.zeros((nv, W), np.complex128) if du_dn is None else np.require(du_dn, np.complex128, ["C"])
        bufs = []
        try:
            for a in (Bfull, Xfull, du):
                p = ctypes.c_void_p()
                _lib.check(L.baemm_dev_alloc(ctypes.byref(p), a.nbytes))
                bufs.append(p)
                _lib.check(L.baemm_dev_upload(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))
            if tangent_point is None:
                metric = _lib.MassMetric(self._h.value, float(c), float(d))
                Mf = ctypes.cast(L.baemm_mass_metric_apply, ctypes.c_void_p)
                Pf = ctypes.cast(L.baemm_mass_metric_jacobi, ctypes.c_void_p)
            else:
                q, p, regpar = (float(v) for v in tangent_point)
                self._check(L.baemm_tp_metric_prepare(self._h, q, p))
                metric = _lib.TPMetric(self._h.value, q, p, regpar)
                Mf = ctypes.cast(L.baemm_tp_metric_apply, ctypes.c_void_p)
                Pf = ctypes.cast(L.baemm_tp_metric_jacobi, ctypes.c_void_p)
            mp = ctypes.cast(ctypes.pointer(metric), ctypes.c_void_p)
            ok = ctypes.c_int(0)
            st = L.baemm_gauss_newton_solve_dev(self._h, _ptr(k), int(wave_chunk_count), _ptr(dr),
                                                int(wave_chunk_size), Mf, mp, Pf, mp, float(alpha), bufs[0], ldB,
                                                float(beta), bufs[1], ldX, bufs[2], 0 if du_dn is None else 1,
                                                int(type), float(cg_tol), float(gmres_tol_inner),
                                                float(gmres_tol_outer), ctypes.byref(ok), None)
            self._check(st)
            _lib.check(L.baemm_dev_download(Xfull.ctypes.data_as(ctypes.c_void_p), bufs[1], Xfull.nbytes))
            _lib.check(L.baemm_dev_download(du.ctypes.data_as(ctypes.c_void_p), bufs[2], du.nbytes))
        finally:
            for p in bufs:
                L.baemm_dev_free(p)
        X[:, :3] = Xfull[:, :3]
        return bool(ok.value), du


    def TangentPointEnergy(self, q=6.0, p=12.0, gradient=True):
        """The tangent-point energy of the mesh (baemm_tp_energy): the
        caller-side energy of Example_GaussNewton/main.cpp:308-309.
        Returns (E, dE/dx as V x 3 float64, or None)."""
        E = ctypes.c_double(0.0)
        g = np.zeros((self.VertexCount(), 3)) if gradient else None
        self._check(self._L.baemm_tp_energy(self._h, float(q), float(p), ctypes.byref(E),
                                            None if g is None else g.ctypes.data_as(ctypes.c_void_p)))
        return E.value, g

    def TangentPointMetric(self, x, q=6.0, p=12.0, scale=1.0):
        """scale times the tangent-point metric applied to x (V x 3 float64):
        the operator of Example_GaussNewton/main.cpp:324-332
        (baemm_tp_metric_apply_host)."""
        xv = np.require(np.asarray(x, np.float64).reshape(-1, 3), np.float64, ["C"])
        if xv.shape[0] != self.VertexCount():
            raise ValueError("x must have vertex_count rows")
        y = np.zeros_like(xv)
        self._check(self._L.baemm_tp_metric_apply_host(self._h, float(q), float(p), float(scale),
                                                       xv.ctypes.data_as(ctypes.c_void_p),
                                                       y.ctypes.data_as(ctypes.c_void_p)))
        return y


    def TangentPointPreconditioner(self, x, q=6.0, p=12.0, scale=1.0):
        """x / (scale d_v) row by row, d the diagonal of the tangent-point metric
        (baemm_tp_metric_jacobi_host)."""
        xv = np.require(np.asarray(x, np.float64).reshape(-1, 3), np.float64, ["C"])
        if xv.shape[0] != self.VertexCount():
            raise ValueError("x must have vertex_count rows")
        y = np.zeros_like(xv)
        self._check(self._L.baemm_tp_metric_jacobi_host(self._h, float(q), float(p), float(scale),
                                                        xv.ctypes.data_as(ctypes.c_void_p),
                                                        y.ctypes.data_as(ctypes.c_void_p)))
        return y


class Helmholtz_HIP(_HelmholtzBase, _SolverMixin):
    """Drop-in for ``BAEMM::Helmholtz_OpenCL`` (Helmholtz_OpenCL.hpp:34-399)."""


class _DevArray:
    """Device buffer of a complex128 host array (row-major), freed on close."""

    def __init__(self, L, host):
        self.L = L
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        _lib.check(L.baemm_dev_alloc(ctypes.byref(self.ptr), max(16, self.nbytes)))
        _lib.check(L.baemm_dev_upload(self.ptr, _ptr(host), self.nbytes))

    def download(self, host):
        _lib.check(self.L.baemm_dev_download(_ptr(host), self.ptr, self.nbytes))

    def close(self):
        if self.ptr:
            self.L.baemm_dev_free(self.ptr)
            self.ptr = ctypes.c_void_p()


def _batch_arrays(instances, B_list, C_list):
    n = len(instances)
    if n == 0 or len(B_list) != n:
        raise ValueError("one B per instance")
    Bs, Cs = [], []
    for k, H in enumerate(instances):
        W = H.GetWaveCount()
        Bs.append(np.ascontiguousarray(_c128_in(B_list[k], "B_in", H.VertexCount())))
        if Bs[-1].shape[1] != W:
            raise ValueError(f"instance {k}: B has {Bs[-1].shape[1]} columns, the loaded wave count is {W}")
        if C_list is not None and C_list[k] is not None:
            Cs.append(np.ascontiguousarray(C_list[k], dtype=np.complex128))
        else:
            Cs.append(np.zeros((H.VertexCount(), W), np.complex128))
    return Bs, Cs


def ApplyBoundaryOperatorsBatch(instances, B_list, C_list=None, alpha=None, beta=None):
    """C_k = alpha_k A_k B_k + beta_k C_k for several Helmholtz_HIP instances in one
    persistent launch per precision (baemm_apply_boundary_batch).

    Each instance has its operators loaded (LoadBoundaryOperators_PL) and its
    own precision.  Returns the list of C arrays (new arrays; C_list is read
    only when beta is given).  No reference counterpart.
    """
    L = _lib.load()
    n = len(instances)
    Bs, Cs = _batch_arrays(instances, B_list, C_list if beta is not None else None)
    dB = [_DevArray(L, b) for b in Bs]
    dC = [_DevArray(L, c) for c in Cs]
    try:
        hs = (ctypes.c_void_p * n)(*[H._h.value for H in instances])
        pB = (ctypes.c_void_p * n)(*[d.ptr.value for d in dB])
        pC = (ctypes.c_void_p * n)(*[d.ptr.value for d in dC])
        ldB = (ctypes.c_int64 * n)(*[b.shape[1] for b in Bs])
        ldC = (ctypes.c_int64 * n)(*[c.shape[1] for c in Cs])

        def scal(v):
            if v is None:
                return None
            v = np.broadcast_to(np.asarray(v, dtype=np.complex128), (n,))
            return (ctypes.c_double * (2 * n))(*np.column_stack([v.real, v.imag]).ravel())

        st = L.baemm_apply_boundary_batch(n, hs, scal(alpha), pB, ldB, _lib.C128, scal(beta), pC, ldC,
                                          _lib.C128, None)
        if st != _lib.OK:
            raw = L.baemm_last_error(None)
            raise BaemmError(st, raw.decode() if raw else "")
        _lib.check(L.baemm_dev_synchronize())
        out = []
        for d, c in zip(dC, Cs):
            r = np.empty_like(c)
            d.download(r)
            out.append(r)
        return out
    finally:
        for d in dB + dC:
            d.close()


def SolveLoadedBatch(instances, B_list, cg_tol=1e-10, gmres_tol=1e-6):
    """X_k with (P A_k) X_k = P B_k for the loaded operator A_k of each instance,
    all solved together (baemm_gmres_batch): each Arnoldi step applies every
    operator still iterating in one batched launch.  Returns (X list,
    iterations list, converged list).  No reference counterpart."""
    L = _lib.load()
    n = len(instances)
    Bs, Xs = _batch_arrays(instances, B_list, None)
    dB = [_DevArray(L, b) for b in Bs]
    dX = [_DevArray(L, x) for x in Xs]
    try:
        hs = (ctypes.c_void_p * n)(*[H._h.value for H in instances])
        pB = (ctypes.c_void_p * n)(*[d.ptr.value for d in dB])
        pX = (ctypes.c_void_p * n)(*[d.ptr.value for d in dX])
        its = (ctypes.c_int64 * n)()
        conv = (ctypes.c_int * n)()
        st = L.baemm_gmres_batch(n, hs, pB, pX, float(cg_tol), float(gmres_tol), its, conv, None)
        if st not in (_lib.OK, _lib.ERR_NOCONV):
            raw = L.baemm_last_error(None)
            raise BaemmError(st, raw.decode() if raw else "")
        out = []
        for d, x in zip(dX, Xs):
            r = np.empty_like(x)
            d.download(r)
            out.append(r)
        return out, list(its), [bool(c) for c in conv]
    finally:
        for d in dB + dX:
            d.close()


def BatchAutotune(instances, B_list, repeats=3):
    """Times the persistent launch's tile shapes on this batch and keeps the
    fastest process-wide; returns (tuning dict, best milliseconds)."""
    L = _lib.load()
    n = len(instances)
    Bs, Cs = _batch_arrays(instances, B_list, None)
    dB = [_DevArray(L, b) for b in Bs]
    dC = [_DevArray(L, c) for c in Cs]
    try:
        hs = (ctypes.c_void_p * n)(*[H._h.value for H in instances])
        pB = (ctypes.c_void_p * n)(*[d.ptr.value for d in dB])
        pC = (ctypes.c_void_p * n)(*[d.ptr.value for d in dC])
        ldB = (ctypes.c_int64 * n)(*[b.shape[1] for b in Bs])
        ldC = (ctypes.c_int64 * n)(*[c.shape[1] for c in Cs])
        best = ctypes.c_double()
        st = L.baemm_batch_autotune(n, hs, pB, ldB, _lib.C128, pC, ldC, _lib.C128, int(repeats), ctypes.byref(best))
        if st != _lib.OK:
            raw = L.baemm_last_error(None)
            raise BaemmError(st, raw.decode() if raw else "")
        return BatchTuning(), best.value
    finally:
        for d in dB + dC:
            d.close()


def BatchTuning(mixed_steps=None, plain_blocks=None, mixed_grid_mult=None, plain_grid_mult=None, own_rounds=None):
    """Reads (and with arguments, sets) the process-wide tile shapes of the
    persistent batched launch, and own_rounds: instances whose own pair launch
    fills at least that many rounds of resident work-groups run it instead of
    joining the persistent launch (0: every instance)."""
    L = _lib.load()
    v = [ctypes.c_int() for _ in range(4)]
    _lib.check(L.baemm_batch_get_tuning(*[ctypes.byref(x) for x in v]))
    cur = dict(zip(("mixed_steps", "plain_blocks", "mixed_grid_mult", "plain_grid_mult"), [x.value for x in v]))
    own = ctypes.c_double()
    _lib.check(L.baemm_batch_get_own_launch(ctypes.byref(own)))
    new = dict(cur)
    for k, val in (("mixed_steps", mixed_steps), ("plain_blocks", plain_blocks),
                   ("mixed_grid_mult", mixed_grid_mult), ("plain_grid_mult", plain_grid_mult)):
        if val is not None:
            new[k] = int(val)
    if new != cur:
        _lib.check(L.baemm_batch_set_tuning(new["mixed_steps"], new["plain_blocks"], new["mixed_grid_mult"],
                                            new["plain_grid_mult"]))
    new["own_rounds"] = own.value
    if own_rounds is not None and float(own_rounds) != own.value:
        _lib.check(L.baemm_batch_set_own_launch(float(own_rounds)))
        new["own_rounds"] = float(own_rounds)
    return new


def BatchStats():
    """Work items and persistent grid size of the last batched call."""
    L = _lib.load()
    v = [ctypes.c_int64() for _ in range(4)]
    _lib.check(L.baemm_batch_get_stats(*[ctypes.byref(x) for x in v]))
    return dict(zip(("items_mixed", "items_fp32", "grid_mixed", "grid_fp32"), [x.value for x in v]))
