"""CPU oracle for the BAEMM brute-force boundary-element hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker (or, for the baseline, as the thing timed on the host cores).  The
product in ``baemm_amd/`` never imports it.

What it holds:

* FP64 brute-force operator mathematics (``*_f64`` in ``orc.c`` plus the
  compositions below).  Each function cites the reference file:line it
  restates.  Pinning (tests/test_oracle_*.py) uses checks that do not retype
  the formula: sphere spectra of S, K, K' (Bessel/Hankel closed forms),
  symmetry A_S^T = A_S and A_K^T = A_K', the Gauss law K 1 = -1/2 at kappa=0,
  polar quadrature of the self term, Funk-Hecke for the far field and the
  Herglotz wave, dense LU solves for CG/GMRES, the Mie series for the
  sound-soft sphere, and the adjoint dot-product test for DF / DF*.
* An FP32 restatement of the reference OpenCL kernel's arithmetic and loop
  order (``orc_bdr_cl_f32``/``orc_far_cl_f32``) and of the FP32 host chain;
  this is the stand-in for the reference's own apply, which cannot run here.

Parity status per function (what pins it) is listed in ``PINS`` below; none is
"parity unpinned" except the Krylov iteration counts, which the reference
delegates to the absent Tensors library (see ``gmres_dense``).
"""
from __future__ import annotations

import ctypes
import math
import os
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_HERE = Path(__file__).resolve().parent
_LIB = None

FOUR_PI = 4.0 * math.pi

PINS = {
    "geometry_f64": "closed forms + polar quadrature of the self term (test_oracle_geometry)",
    "mass / Av / AvT": "1^T M 1 = sum a, M 1 = AvT 1 (test_oracle_geometry)",
    "bdr_f64": "sphere spectra of S, K, K'; A_S^T = A_S; A_K^T = A_K'; Gauss law (test_oracle_operators)",
    "bdr_cl_f32": "agrees with bdr_f64 to FP32 accuracy (test_oracle_operators)",
    "far_cl_f32": "agrees with far_f64 to FP32 accuracy in every Re/Im branch (test_oracle_operators)",
    "far_f64": "Funk-Hecke closed form, far field of the SL of a plane-wave density (test_oracle_operators)",
    "herglotz_f64": "Funk-Hecke / j0 closed form (test_oracle_operators)",
    "near_f64": "limit of the near field equals the far field; SL of the sphere exterior (test_oracle_operators)",
    "incident": "plane-wave closed form (test_oracle_operators)",
    "solves": "dense LU; Mie series for the sound-soft sphere; adjoint dot-product test (test_oracle_solvers)",
    "gauss_newton": "its dense operator reproduces the pinned derivative_ff / adjoint_derivative_ff compositions "
                    "column by column; the solution satisfies the normal equations (test_oracle_solvers)",
    "tp_energy / tp_energy_gradient": "the gradient agrees with central finite differences of the energy; the "
                                      "energy scales as lambda^(4-p+q) under dilation and is invariant under "
                                      "rotations and translations (test_oracle_tangent_point)",
    "tp_metric": "symmetric positive semi-definite; translations are in its kernel; it scales as lambda^(-2 sigma) "
                 "under dilation; the face gradients reproduce the tangential projector (test_oracle_tangent_point)",
}


def build():
    """Compile orc.c into oracle/liboracle.so (gcc, OpenMP)."""
    import subprocess

    subprocess.run(["make", "-s", "-B", "-C", str(_HERE)], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        L = ctypes.CDLL(str(path))
        P = ctypes.c_void_p
        I = ctypes.c_int64
        L.orc_geometry_f64.argtypes = [I, P, P, P, P, P, P]
        L.orc_geometry_f32.argtypes = [I, P, P, P, P, P, P]
        L.orc_bdr_f64.argtypes = [I, P, P, I, I, P, P, P, I, P, I, P, I]
        L.orc_far_f64.argtypes = [I, P, P, I, P, I, I, P, P, P, I, P, I]
        L.orc_near_f64.argtypes = [I, P, P, I, P, I, I, P, P, P, I, P, I]
        L.orc_herglotz_f64.argtypes = [I, P, P, I, P, I, I, P, P, P, I, P, I]
        L.orc_bdr_cl_f32.argtypes = [I, P, P, I, I, P, P, ctypes.c_int, P, P, I, I]
        L.orc_far_cl_f32.argtypes = [I, P, P, I, P, I, I, P, P, ctypes.c_int, P, P]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


# --------------------------------------------------------------------------
# Coefficients (LoadParameters.hpp:21-81)
# --------------------------------------------------------------------------

def load_parameters(kappa, coeff, W, wcs, factor=1.0):
    """Return (kappa[K], c[K,4]) with c_j (j=1..3) scaled by factor/(4 pi).

    LoadParameters.hpp:65-79: c_q0 = coeff[4q]; c_qj = coeff[4q+j] / (4 pi) * factor.
    K = ceil(W / wcs) (GetSetters.hpp:59-62).
    """
    K = -(-int(W) // int(wcs))
    kap = np.asarray(kappa, dtype=np.float64)[:K].copy()
    c = np.asarray(coeff, dtype=np.complex128).reshape(-1)[: 4 * K].reshape(K, 4).copy()
    c[:, 1:] *= factor / FOUR_PI
    return kap, c


def flags_of(c):
    """The 8 Re/Im flags of LoadParameters.hpp:47-80 (OR over chunks)."""
    f = {}
    names = ["mass_matrix", "single_layer", "double_layer", "adjdbl_layer"]
    for j, n in enumerate(names):
        f["Re_" + n] = bool(np.any(c[:, j].real != 0))
        f["Im_" + n] = bool(np.any(c[:, j].imag != 0))
    return f


# --------------------------------------------------------------------------
# Mesh: geometry and sparse operators (Initialize.hpp)
# --------------------------------------------------------------------------

class Mesh:
    """Geometry and P1/P0 sparse operators of a triangle mesh in FP64.

    Initialize.hpp:57-353: centroids, unit normals, areas, single-layer self
    term; Av (P1 -> P0, weights a_t/3), AvT, Mass (P1 Galerkin, a/6 diagonal
    and a/12 off-diagonal per triangle, duplicates summed), lumped inverse.
    """

    def __init__(self, V, T, meas=None):
        self.V = np.ascontiguousarray(V, dtype=np.float64)
        self.T = np.ascontiguousarray(T, dtype=np.int64)
        self.nv = self.V.shape[0]
        self.nt = self.T.shape[0]
        self.meas = None if meas is None else np.ascontiguousarray(meas, dtype=np.float64)
        nt = self.nt
        self.mid = np.zeros((nt, 3))
        self.nrm = np.zeros((nt, 3))
        self.area = np.zeros(nt)
        self.sdiag = np.zeros(nt)
        lib().orc_geometry_f64(nt, _p(self.V), _p(self.T), _p(self.mid), _p(self.nrm),
                               _p(self.area), _p(self.sdiag))
        rows = np.repeat(np.arange(nt), 3)
        cols = self.T.ravel()
        vals = np.repeat(self.area / 3.0, 3)
        self.Av = sp.csr_matrix((vals, (rows, cols)), shape=(nt, self.nv))
        self.AvT = self.Av.T.tocsr()
        loc = np.array([[2.0, 1.0, 1.0], [1.0, 2.0, 1.0], [1.0, 1.0, 2.0]]) / 12.0
        I = np.repeat(self.T, 3, axis=1).ravel()
        J = np.tile(self.T, (1, 3)).ravel()
        Mv = (self.area[:, None, None] * loc[None]).ravel()
        self.M = sp.csr_matrix((Mv, (I, J)), shape=(self.nv, self.nv))
        self.M.sum_duplicates()
        self._Mlu = None

    def geometry_f32(self):
        """FP32 geometry in the reference's arithmetic order (Initialize.hpp)."""
        nt = self.nt
        mid4 = np.zeros((nt, 4), np.float32)
        nrm4 = np.zeros((nt, 4), np.float32)
        area = np.zeros(nt, np.float32)
        sdiag = np.zeros(nt, np.float32)
        lib().orc_geometry_f32(nt, _p(self.V), _p(self.T), _p(mid4), _p(nrm4), _p(area), _p(sdiag))
        return mid4, nrm4, area, sdiag

    def mass_solve(self, X):
        """Exact M^{-1} X (the quantity ApplyMassInverse's CG approximates)."""
        if self._Mlu is None:
            self._Mlu = spla.splu(self.M.tocsc())
        X = np.asarray(X)
        if np.iscomplexobj(X):
            return self._Mlu.solve(np.ascontiguousarray(X.real)) + 1j * self._Mlu.solve(np.ascontiguousarray(X.imag))
        return self._Mlu.solve(np.ascontiguousarray(X, dtype=np.float64))


# --------------------------------------------------------------------------
# Pair kernels (FP64) on P0 data
# --------------------------------------------------------------------------

def bdr_offdiag_f64(mesh: Mesh, kappa, c, b, wcs, rows=None):
    """C = K_offdiag(b) for selected target rows (formula A1, orc_bdr_f64)."""
    b = _c128(b)
    W = b.shape[1]
    rows = np.arange(mesh.nt, dtype=np.int64) if rows is None else np.ascontiguousarray(rows, dtype=np.int64)
    C = np.zeros((rows.shape[0], W), np.complex128)
    kap = np.ascontiguousarray(kappa, dtype=np.float64)
    cc = _c128(c)
    lib().orc_bdr_f64(mesh.nt, _p(mesh.mid), _p(mesh.nrm), W, int(wcs), _p(kap), _p(cc),
                      _p(b), W, _p(C), W, _p(rows), rows.shape[0])
    return C


def far_f64(mesh: Mesh, dirs, kappa, c, b, wcs):
    b = _c128(b)
    W = b.shape[1]
    dirs = np.ascontiguousarray(dirs, dtype=np.float64)
    F = np.zeros((dirs.shape[0], W), np.complex128)
    kap = np.ascontiguousarray(kappa, dtype=np.float64)
    cc = _c128(c)
    lib().orc_far_f64(mesh.nt, _p(mesh.mid), _p(mesh.nrm), dirs.shape[0], _p(dirs), W, int(wcs),
                      _p(kap), _p(cc), _p(b), W, _p(F), W)
    return F


def near_f64(mesh: Mesh, pts, kappa, c, b, wcs):
    b = _c128(b)
    W = b.shape[1]
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    U = np.zeros((pts.shape[0], W), np.complex128)
    kap = np.ascontiguousarray(kappa, dtype=np.float64)
    cc = _c128(c)
    lib().orc_near_f64(mesh.nt, _p(mesh.mid), _p(mesh.nrm), pts.shape[0], _p(pts), W, int(wcs),
                       _p(kap), _p(cc), _p(b), W, _p(U), W)
    return U


def herglotz_f64(mesh: Mesh, dirs, kappa, c, g, wcs):
    g = _c128(g)
    W = g.shape[1]
    dirs = np.ascontiguousarray(dirs, dtype=np.float64)
    H = np.zeros((mesh.nt, W), np.complex128)
    kap = np.ascontiguousarray(kappa, dtype=np.float64)
    cc = _c128(c)
    lib().orc_herglotz_f64(mesh.nt, _p(mesh.mid), _p(mesh.nrm), dirs.shape[0], _p(dirs), W, int(wcs),
                           _p(kap), _p(cc), _p(g), W, _p(H), W)
    return H


def _chunk_cols(W, wcs, K):
    return [slice(q * wcs, min(W, (q + 1) * wcs)) for q in range(K)]


# --------------------------------------------------------------------------
# Operator applies (host compositions)
# --------------------------------------------------------------------------

def apply_boundary(mesh: Mesh, alpha, B, beta, C, kappa, coeff, W, wcs, use_diagonal=True, rows=None):
    """FP64 restatement of ApplyBoundaryOperators_PL (ApplyBoundaryOperators.hpp:77-178).

    C_out = alpha * A B + beta * C_out with
    A = sum_q [c_q0 M + AvT (K_q + D_q) Av] on the columns of chunk q, and
    D_q = c_q1 (d_t + i kappa_q) on the diagonal (ApplySingleLayerDiagonal.hpp:
    36-50).  If no coefficient is non-zero C is returned unchanged (the
    reference touches C_out only inside its two branches).
    ``rows`` (vertex subset) is not supported here: the P1 output needs all rows.
    """
    kap, c = load_parameters(kappa, coeff, W, wcs)
    K = kap.shape[0]
    f = flags_of(c)
    B = _c128(B)
    C = np.array(C, dtype=np.complex128, copy=True)
    kernel_on = any(f[k] for k in f if not k.endswith("mass_matrix"))
    mass_on = f["Re_mass_matrix"] or f["Im_mass_matrix"]
    cols = _chunk_cols(W, wcs, K)
    if kernel_on:
        b = mesh.Av @ B
        Ck = bdr_offdiag_f64(mesh, kap, c, b, wcs)
        if use_diagonal and (f["Re_single_layer"] or f["Im_single_layer"]):
            for q, sl in enumerate(cols):
                Ck[:, sl] += (c[q, 1] * (mesh.sdiag + 1j * kap[q]))[:, None] * b[:, sl]
        C = alpha * (mesh.AvT @ Ck) + beta * C
        add_to = 1.0
    else:
        add_to = beta
    if mass_on:
        MB = mesh.M @ B
        for q, sl in enumerate(cols):
            C[:, sl] = alpha * c[q, 0] * MB[:, sl] + add_to * C[:, sl]
    return C


def apply_boundary_rows(mesh: Mesh, alpha, B, beta, C, kappa, coeff, W, wcs, vrows, use_diagonal=True):
    """apply_boundary restricted to the vertex rows ``vrows`` (C has len(vrows) rows).

    Only the P0 rows of triangles incident to those vertices are evaluated, so
    the O(N^2) oracle stays cheap on large meshes (sampled-row parity)."""
    kap, c = load_parameters(kappa, coeff, W, wcs)
    K = kap.shape[0]
    f = flags_of(c)
    B = _c128(B)
    vrows = np.asarray(vrows, dtype=np.int64)
    C = np.array(C, dtype=np.complex128, copy=True)
    kernel_on = any(f[k] for k in f if not k.endswith("mass_matrix"))
    mass_on = f["Re_mass_matrix"] or f["Im_mass_matrix"]
    cols = _chunk_cols(W, wcs, K)
    if kernel_on:
        AvT_rows = mesh.AvT[vrows]
        tris = np.unique(AvT_rows.indices)
        b = mesh.Av @ B
        Ck = bdr_offdiag_f64(mesh, kap, c, b, wcs, rows=tris)
        if use_diagonal and (f["Re_single_layer"] or f["Im_single_layer"]):
            for q, sl in enumerate(cols):
                Ck[:, sl] += (c[q, 1] * (mesh.sdiag[tris] + 1j * kap[q]))[:, None] * b[tris][:, sl]
        full = np.zeros((mesh.nt, W), np.complex128)
        full[tris] = Ck
        C = alpha * (AvT_rows @ full) + beta * C
        add_to = 1.0
    else:
        add_to = beta
    if mass_on:
        MB = mesh.M[vrows] @ B
        for q, sl in enumerate(cols):
            C[:, sl] = alpha * c[q, 0] * MB[:, sl] + add_to * C[:, sl]
    return C


def apply_far_field(mesh: Mesh, alpha, B, beta, F, kappa, coeff, W, wcs):
    """ApplyFarFieldOperators_PL (ApplyFarFieldOperators.hpp:49-104), FP64."""
    kap, c = load_parameters(kappa, coeff, W, wcs)
    f = flags_of(c)
    F = np.array(F, dtype=np.complex128, copy=True)
    if f["Re_single_layer"] or f["Im_single_layer"] or f["Re_double_layer"] or f["Im_double_layer"]:
        b = mesh.Av @ _c128(B)
        F = alpha * far_f64(mesh, mesh.meas, kap, c, b, wcs) + beta * F
    return F


def apply_near_field(mesh: Mesh, alpha, B, beta, U, kappa, coeff, W, wcs, pts):
    """ApplyNearFieldOperators_PL (ApplyNearFieldOperators.hpp:160-226), FP64."""
    kap, c = load_parameters(kappa, coeff, W, wcs)
    f = flags_of(c)
    U = np.array(U, dtype=np.complex128, copy=True)
    if f["Re_single_layer"] or f["Im_single_layer"] or f["Re_double_layer"] or f["Im_double_layer"]:
        b = mesh.Av @ _c128(B)
        U = alpha * near_f64(mesh, pts, kap, c, b, wcs) + beta * U
    return U


def create_herglotz_wave(mesh: Mesh, alpha, g, beta, out, kappa, coeff, W, wcs):
    """CreateHerglotzWave_PL (CreateHerglotzWave.hpp:152-222), FP64: factor 4 pi / m."""
    m = mesh.meas.shape[0]
    kap, c = load_parameters(kappa, coeff, W, wcs, factor=FOUR_PI / m)
    f = flags_of(c)
    out = np.array(out, dtype=np.complex128, copy=True)
    if f["Re_single_layer"] or f["Im_single_layer"] or f["Re_double_layer"] or f["Im_double_layer"]:
        H = herglotz_f64(mesh, mesh.meas, kap, c, g, wcs)
        out = alpha * (mesh.AvT @ H) + beta * out
    return out


def incident_wave_p0(mesh: Mesh, dirs, kappa, coeff, W, wcs, wave_type="plane"):
    """P0 incident-wave values before AvT (IncidentWaveKernel_Plane.hpp:190-233,
    IncidentWaveKernel_Radial.hpp:5-57), factor 4 pi so c_j = coeff_j.

    Plane:  U_t = (c1 + c2 i kappa (d.n_t)) exp(i kappa d.x_t)
    Radial: U_t = (c1 + c2 (-(n.delta)/R^2 + i kappa (n.delta)/R)) exp(i kappa R) / (4 pi R),
            delta = x_t - s, R = |delta|.
    The same wcs directions are used for every chunk (Q11).
    """
    kap, c = load_parameters(kappa, coeff, W, wcs, factor=FOUR_PI)
    K = kap.shape[0]
    d = np.asarray(dirs, dtype=np.float64).reshape(-1, 3)[:wcs]
    U = np.zeros((mesh.nt, W), np.complex128)
    for q, sl in enumerate(_chunk_cols(W, wcs, K)):
        n = sl.stop - sl.start
        dd = d[:n]
        if wave_type == "plane":
            ph = mesh.mid @ dd.T
            dn = mesh.nrm @ dd.T
            U[:, sl] = (c[q, 1] + c[q, 2] * 1j * kap[q] * dn) * np.exp(1j * kap[q] * ph)
        else:
            delta = mesh.mid[:, None, :] - dd[None, :, :]
            R = np.linalg.norm(delta, axis=2)
            nd = np.einsum("tk,tjk->tj", mesh.nrm, delta)
            A = 1.0 / (FOUR_PI * R)
            Bf = -nd / R ** 2 + 1j * nd * kap[q] / R
            U[:, sl] = (c[q, 1] * A + c[q, 2] * A * Bf) * np.exp(1j * kap[q] * R)
    return U


def create_incident_wave(mesh: Mesh, alpha, dirs, beta, out, kappa, coeff, W, wcs, wave_type="plane"):
    """CreateIncidentWave_PL (CreateIncidentWave.hpp:55-127), FP64."""
    kap, c = load_parameters(kappa, coeff, W, wcs, factor=FOUR_PI)
    f = flags_of(c)
    out = np.array(out, dtype=np.complex128, copy=True)
    if f["Re_single_layer"] or f["Im_single_layer"] or f["Re_double_layer"] or f["Im_double_layer"]:
        U = incident_wave_p0(mesh, dirs, kappa, coeff, W, wcs, wave_type)
        out = alpha * (mesh.AvT @ U) + beta * out
    return out


# --------------------------------------------------------------------------
# FP32 restatement of the reference's apply chain (the CPU baseline)
# --------------------------------------------------------------------------

class F32Chain:
    """FP32 restatement of ApplyBoundaryOperators_PL's chain in the reference's
    precision: FP32 geometry (Initialize.hpp), b = Av B in complex64,
    BoundaryOperatorKernel.cl loop order (orc_bdr_cl_f32), SL diagonal,
    AvT and mass in FP64 accumulation of FP32 data."""

    def __init__(self, mesh: Mesh):
        self.mesh = mesh
        self.mid4, self.nrm4, self.area, self.sdiag = mesh.geometry_f32()
        nt = mesh.nt
        rows = np.repeat(np.arange(nt), 3)
        vals = np.repeat(self.area.astype(np.float64) / 3.0, 3).astype(np.float32)
        self.Av = sp.csr_matrix((vals, (rows, mesh.T.ravel())), shape=(nt, mesh.nv))
        self.AvT = self.Av.T.tocsr()
        if mesh.meas is not None:
            self.dirs4 = np.zeros((mesh.meas.shape[0], 4), np.float32)
            self.dirs4[:, :3] = mesh.meas

    def kernel(self, kappa32, coef32, flags, b32, wcs, row_begin=0, row_end=None):
        nt = self.mesh.nt
        row_end = nt if row_end is None else row_end
        W = b32.shape[1]
        C = np.zeros((nt, W), np.complex64)
        lib().orc_bdr_cl_f32(nt, _p(self.mid4), _p(self.nrm4), W, int(wcs), _p(kappa32), _p(coef32),
                             int(flags), _p(b32), _p(C), int(row_begin), int(row_end))
        return C

    @staticmethod
    def params32(kappa, coeff, W, wcs, factor=1.0):
        K = -(-int(W) // int(wcs))
        kap32 = np.asarray(kappa, dtype=np.float64)[:K].astype(np.float32)
        c = np.asarray(coeff, dtype=np.complex128).reshape(-1)[: 4 * K].astype(np.complex64).reshape(K, 4)
        s = np.float32(1.0 / (4.0 * math.pi)) * np.float32(factor)
        c32 = c.copy()
        c32[:, 1:] = c[:, 1:] * s
        fl = flags_of(c32.astype(np.complex128))
        bits = (int(fl["Re_single_layer"]) | int(fl["Im_single_layer"]) << 1 | int(fl["Re_double_layer"]) << 2
                | int(fl["Im_double_layer"]) << 3 | int(fl["Re_adjdbl_layer"]) << 4 | int(fl["Im_adjdbl_layer"]) << 5)
        return kap32, np.ascontiguousarray(c32), bits, fl

    def apply_boundary(self, alpha, B, beta, C, kappa, coeff, W, wcs, use_diagonal=True):
        kap32, c32, bits, fl = self.params32(kappa, coeff, W, wcs)
        K = kap32.shape[0]
        C = np.array(C, dtype=np.complex128, copy=True)
        B = _c128(B)
        cols = _chunk_cols(W, wcs, K)
        if bits:
            b32 = np.ascontiguousarray((self.Av @ B).astype(np.complex64))
            Ck = self.kernel(kap32, c32, bits, b32, wcs)
            if use_diagonal and (bits & 3):
                for q, sl in enumerate(cols):
                    fac = c32[q, 1] * (self.sdiag + 1j * kap32[q]).astype(np.complex64)
                    Ck[:, sl] += fac[:, None] * b32[:, sl]
            C = alpha * (self.AvT @ Ck.astype(np.complex128)) + beta * C
            add_to = 1.0
        else:
            add_to = beta
        if fl["Re_mass_matrix"] or fl["Im_mass_matrix"]:
            MB = self.mesh.M @ B
            for q, sl in enumerate(cols):
                C[:, sl] = alpha * np.complex128(c32[q, 0]) * MB[:, sl] + add_to * C[:, sl]
        return C

    def far_field(self, alpha, B, beta, F, kappa, coeff, W, wcs):
        """FP32 restatement of ApplyFarFieldOperators_PL (ApplyFarFieldOperators.hpp:
        49-104): b = Av B in complex64, FarFieldOperatorKernel.cl arithmetic
        (orc_far_cl_f32, coefficient columns 1 and 2 only), then alpha/beta."""
        kap32, c32, bits, fl = self.params32(kappa, coeff, W, wcs)
        F = np.array(F, dtype=np.complex128, copy=True)
        bits &= 15
        if not bits:
            return F   # no single or double layer: F_out untouched
        b32 = np.ascontiguousarray((self.Av @ _c128(B)).astype(np.complex64))
        m = self.dirs4.shape[0]
        out = np.zeros((m, int(W)), np.complex64)
        lib().orc_far_cl_f32(self.mesh.nt, _p(self.mid4), _p(self.nrm4), m, _p(self.dirs4), int(W), int(wcs),
                             _p(kap32), _p(c32), int(bits), _p(b32), _p(out))
        return alpha * out.astype(np.complex128) + beta * F


# --------------------------------------------------------------------------
# Dense operator assembly and solver compositions (FarField.hpp), FP64
# --------------------------------------------------------------------------

def dense_p0_matrix(mesh: Mesh, kap, c, use_diagonal=True):
    """Dense P0 kernel matrix (nt x nt) for one chunk: A1 off-diagonal + D."""
    mid, nrm = mesh.mid, mesh.nrm
    Z = mid[None, :, :] - mid[:, None, :]          # z = y_j - x_i
    r = np.linalg.norm(Z, axis=2)
    np.fill_diagonal(r, 1.0)
    rinv = 1.0 / r
    np.fill_diagonal(rinv, 0.0)
    E = np.exp(1j * kap * r)
    znj = np.einsum("ijk,jk->ij", Z, nrm)
    zni = np.einsum("ijk,ik->ij", Z, nrm)
    F = (1j * kap * r - 1.0) * E * rinv ** 3
    A = c[1] * E * rinv + (c[2] * znj - c[3] * zni) * F
    if use_diagonal and c[1] != 0:
        A[np.diag_indices(mesh.nt)] = c[1] * (mesh.sdiag + 1j * kap)
    else:
        A[np.diag_indices(mesh.nt)] = 0.0
    return A


def dense_operator(mesh: Mesh, kap, c, use_diagonal=True):
    """Dense weak-form operator c0 M + AvT (K + D) Av (nv x nv) for one chunk."""
    A = dense_p0_matrix(mesh, kap, c, use_diagonal)
    Av = mesh.Av.toarray()
    return c[0] * mesh.M.toarray() + Av.T @ A @ Av


def solve_boundary(mesh: Mesh, kappa, coeff_rows, rhs, wcs):
    """Exact solution of (sum_q operator_q) X = rhs per chunk (GMRES target)."""
    W = rhs.shape[1]
    K = W // wcs
    X = np.zeros_like(rhs, dtype=np.complex128)
    for q in range(K):
        kap, c = load_parameters([kappa[q]], coeff_rows[q], wcs, wcs)
        A = dense_operator(mesh, kap[0], c[0])
        sl = slice(q * wcs, (q + 1) * wcs)
        X[:, sl] = np.linalg.solve(A, rhs[:, sl])
    return X


def far_field(mesh: Mesh, kappa, dirs, wcs, wave_type="plane", eta=None):
    """FarField / FarField_parameters (FarField.hpp:28-144) with exact solves."""
    K = len(kappa)
    W = K * wcs
    eta = kappa if eta is None else eta
    inc = np.tile(np.array([0, -1, 0, 0], np.complex128), K)
    wave = create_incident_wave(mesh, 1.0, dirs, 0.0, np.zeros((mesh.nv, W)), kappa, inc, W, wcs, wave_type)
    rows = [np.array([0.5, -1j * eta[q], 1.0, 0.0]) for q in range(K)]
    phi = solve_boundary(mesh, kappa, rows, wave, wcs)
    coeff = np.concatenate(rows)
    return apply_far_field(mesh, 1.0, phi, 0.0, np.zeros((mesh.meas.shape[0], W)), kappa, coeff, W, wcs), phi


def dtn(mesh: Mesh, kappa, wave, wcs):
    """DirichletToNeumann (FarField.hpp:790-879): (1/2 M - i kappa S + K') psi = wave."""
    K = len(kappa)
    rows = [np.array([0.5, -1j * kappa[q], 0.0, 1.0]) for q in range(K)]
    return solve_boundary(mesh, kappa, rows, wave, wcs)


def derivative_ff_helper(mesh: Mesh, kappa, dirs, wcs, wave_type="plane"):
    """Derivative_FF_Helper (FarField.hpp:148-207): du_dn = DtN(du_i/dn - i kappa u_i)."""
    K = len(kappa)
    W = K * wcs
    inc = np.concatenate([np.array([0, -1j * kappa[q], 1.0, 0]) for q in range(K)])
    inc_wave = create_incident_wave(mesh, 1.0, dirs, 0.0, np.zeros((mesh.nv, W)), kappa, inc, W, wcs, wave_type)
    return dtn(mesh, kappa, inc_wave, wcs)


def dot_with_normals(mesh: Mesh, X):
    """DotWithNormals_PL (FarField.hpp:956-1022) with an exact mass solve."""
    Xt = mesh.Av @ X                                   # nt x 3 (area-weighted)
    Y = np.einsum("tk,tk->t", mesh.nrm, Xt) / mesh.area
    return mesh.mass_solve(mesh.AvT @ Y)


def multiply_with_normals(mesh: Mesh, x, factor):
    """MultiplyWithNormals_PL (FarField.hpp:886-952) with an exact mass solve."""
    xt = mesh.Av @ x
    Y = (xt.real / mesh.area)[:, None] * mesh.nrm
    return mesh.mass_solve(factor * (mesh.AvT @ Y))


def derivative_ff(mesh: Mesh, kappa, dirs, wcs, X, du_dn=None, wave_type="plane"):
    """Derivative_FF (FarField.hpp:227-321)."""
    K = len(kappa)
    W = K * wcs
    if du_dn is None:
        du_dn = derivative_ff_helper(mesh, kappa, dirs, wcs, wave_type)
    xn = dot_with_normals(mesh, X)
    bdr = -xn[:, None] * du_dn
    bdr_weak = mesh.M @ bdr
    rows = [np.array([0.5, -1j * kappa[q], 1.0, 0.0]) for q in range(K)]
    phi = solve_boundary(mesh, kappa, rows, bdr_weak, wcs)
    coeff = np.concatenate(rows)
    Y = apply_far_field(mesh, 1.0, phi, 0.0, np.zeros((mesh.meas.shape[0], W)), kappa, coeff, W, wcs)
    return Y, du_dn


def adjoint_derivative_ff(mesh: Mesh, kappa, dirs, wcs, g, du_dn=None, wave_type="plane"):
    """AdjointDerivative_FF (FarField.hpp:342-428).

    The row sum of du_dn .* dv_dn uses the plain (unconjugated) product; the
    reference's dot_buffers convention is fixed by the adjoint dot-product test
    (test_oracle_solvers) rather than by reading Tools.
    """
    K = len(kappa)
    W = K * wcs
    if du_dn is None:
        du_dn = derivative_ff_helper(mesh, kappa, dirs, wcs, wave_type)
    inc = np.concatenate([np.array([0, -1j * kappa[q], 1.0, 0]) for q in range(K)])
    herg = create_herglotz_wave(mesh, 1.0, g, 0.0, np.zeros((mesh.nv, W)), kappa, inc, W, wcs)
    dv_dn = dtn(mesh, kappa, herg, wcs)
    prod = np.sum(du_dn * dv_dn, axis=1)
    return multiply_with_normals(mesh, prod, -1.0 / W), du_dn


def gauss_newton_operator(mesh: Mesh, kappa, dirs, wcs, metric, du_dn=None, wave_type="plane"):
    """Dense matrix of the outer GMRES operator of GaussNewtonSolve
    (FarField.hpp:551-572): A x = metric x + (Mass (x) I_3) DF* DF x on vec(X),
    X V x 3 row-major (entry 3 v + d).  ``metric`` is a dense (3V x 3V) matrix
    (the caller's M).  Returns (A, du_dn).

    DF and DF* are derivative_ff and adjoint_derivative_ff with each of their
    linear stages written as a dense matrix (the boundary operators factorised
    once per chunk, exact mass solves), so A is built for all 3V unit vectors
    at once:
      DF:  xn = N vec(X) (DotWithNormals), Y[:, c] = -FF_q A_q^{-1} M (xn * du[:, c])
      DF*: dv[:, c] = D_q^{-1} H_q conj(g[:, c]), p = sum_c du[:, c] dv[:, c],
           vec(Y) = R Re(p) (MultiplyWithNormals with factor -1/W)
    with A_q the potential operator (1/2, -i kappa, 1, 0), D_q the DtN operator
    (1/2, -i kappa, 0, 1), FF_q the far field and H_q the Herglotz wave
    (0, -i kappa, 1, 0) of chunk q.
    """
    K = len(kappa)
    W = K * wcs
    nv, nt = mesh.nv, mesh.nt
    m = mesh.meas.shape[0]
    if du_dn is None:
        du_dn = derivative_ff_helper(mesh, kappa, dirs, wcs, wave_type)
    Avd = mesh.Av.toarray()
    # N: vec(X) -> M^{-1} AvT ((n_t . (Av X)_t) / a_t)
    T1 = np.zeros((nt, nv, 3))
    for d in range(3):
        T1[:, :, d] = Avd * (mesh.nrm[:, d] / mesh.area)[:, None]
    N = mesh.mass_solve(Avd.T @ T1.reshape(nt, 3 * nv))
    # R: p (nv) -> vec(-1/W M^{-1} AvT ((Av p) / a_t n_t))
    R = np.zeros((nv, 3, nv))
    for d in range(3):
        R[:, d, :] = mesh.mass_solve(Avd.T @ (Avd * (mesh.nrm[:, d] / mesh.area)[:, None])) * (-1.0 / W)
    R = R.reshape(3 * nv, nv)
    Mfull = mesh.M.toarray()
    DFx = np.zeros((m, W, 3 * nv), np.complex128)
    P = np.zeros((nv, 3 * nv), np.complex128)
    for q in range(K):
        kq = kappa[q]
        pot = np.array([0.5, -1j * kq, 1.0, 0.0])
        inc = np.array([0.0, -1j * kq, 1.0, 0.0])
        kp, cp = load_parameters([kq], pot, 1, 1)
        Aq = dense_operator(mesh, kp[0], cp[0])
        kd, cd = load_parameters([kq], [0.5, -1j * kq, 0.0, 1.0], 1, 1)
        Dq = dense_operator(mesh, kd[0], cd[0])
        FFq = apply_far_field(mesh, 1.0, np.eye(nv), 0.0, np.zeros((m, nv)), [kq], pot, nv, nv)
        Hq = create_herglotz_wave(mesh, 1.0, np.eye(m), 0.0, np.zeros((nv, m)), [kq], inc, m, m)
        G = FFq @ np.linalg.solve(Aq, Mfull)            # m x nv
        L = np.linalg.solve(Dq, Hq)                      # nv x m
        for c in range(q * wcs, (q + 1) * wcs):
            DFx[:, c, :] = -G @ (du_dn[:, c][:, None] * N)
            P += du_dn[:, c][:, None] * (L @ np.conj(DFx[:, c, :]))
    Y = R @ P.real                                        # DF* DF on unit vectors, 3V x 3V
    MY = (Mfull @ Y.reshape(nv, 3, 3 * nv).reshape(nv, 9 * nv)).reshape(nv, 3, 3 * nv).reshape(3 * nv, 3 * nv)
    return np.asarray(metric, dtype=np.float64) + MY, du_dn


def gauss_newton(mesh: Mesh, kappa, dirs, wcs, metric, B, alpha=1.0, beta=0.0, X=None, du_dn=None,
                 wave_type="plane"):
    """GaussNewtonSolve (FarField.hpp:457-604) in FP64 with an exact solve:
    X_out = alpha A^{-1} vec(B) + beta X_out (the combine_matrices step of
    :581-599) with A from gauss_newton_operator.  B, X: V x 3 real.
    Returns (X_out, du_dn)."""
    A, du_dn = gauss_newton_operator(mesh, kappa, dirs, wcs, metric, du_dn, wave_type)
    x = np.linalg.solve(A, np.asarray(B, dtype=np.float64).reshape(-1)).reshape(mesh.nv, 3)
    out = alpha * x
    if X is not None and beta != 0.0:
        out = out + beta * np.asarray(X, dtype=np.float64)
    return out, du_dn


# --------------------------------------------------------------------------
# Tangent-point energy and metric of the Gauss-Newton example's caller
# (Example_GaussNewton/main.cpp:301-346 takes them from Repulsor's
# TangentPointEnergy0 / TangentPointMetric0, which is not part of BAEMM and not
# vendored).  Restated from the published discretisation (C. Yu, C.
# Brakensiek, H. Schumacher, K. Crane, "Repulsive Surfaces", ACM Trans. Graph.
# 40(6), 2021): centroid quadrature over pairs of distinct triangles, P1
# vertex functions.  Parity with Repulsor is unpinned; these functions are
# pinned by finite differences and by structure (PINS).
# --------------------------------------------------------------------------

def _tp_pairs(mesh):
    c, n, a = mesh.mid, mesh.nrm, mesh.area
    d = c[:, None, :] - c[None, :, :]          # d[S, T] = c_S - c_T
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, 1.0)
    return c, n, a, d, r2


def tp_energy(mesh: Mesh, q: float, p: float) -> float:
    """E = sum_{S != T} a_S a_T |n_S . (c_S - c_T)|^q / |c_S - c_T|^p  (FP64)."""
    _, n, a, d, r2 = _tp_pairs(mesh)
    s = np.einsum("sk,stk->st", n, d)
    K = np.abs(s) ** q * r2 ** (-p / 2)
    np.fill_diagonal(K, 0.0)
    return float(a @ K @ a)


def tp_energy_gradient(mesh: Mesh, q: float, p: float) -> np.ndarray:
    """dE/dx_v (V x 3) of tp_energy.  Per ordered pair (S, T), with d = c_S - c_T,
    s = n_S . d, R = |d|^-p and N = a n the area vector:
      dTerm/dN_S = a_T R ((1 - q) |s|^q n_S + q |s|^(q-1) sgn(s) d)
      dTerm/dN_T = a_S |s|^q R n_T
      dTerm/dd   = a_S a_T (q |s|^(q-1) sgn(s) R n_S - p |s|^q R d / |d|^2),
    c_S = (x0 + x1 + x2) / 3 and N = (x1 - x0) x (x2 - x0) / 2, so
    d(g . N)/dx_k = (x_(k+1) - x_(k+2)) x g / 2."""
    _, n, a, d, r2 = _tp_pairs(mesh)
    nt = mesh.nt
    s = np.einsum("sk,stk->st", n, d)
    As = np.abs(s)
    R = r2 ** (-p / 2)
    np.fill_diagonal(R, 0.0)
    P = As ** q
    Pm = q * As ** (q - 1) * np.sign(s)
    gN = ((a[None, :] * R * (1 - q) * P).sum(1))[:, None] * n + np.einsum("st,stk->sk", a[None, :] * R * Pm, d)
    gN += ((a[:, None] * P * R).sum(0))[:, None] * n
    aa = np.outer(a, a)
    full = (aa * Pm * R)[:, :, None] * n[:, None, :] - (aa * p * P * R / r2)[:, :, None] * d
    gc = full.sum(1) - full.sum(0)
    X = mesh.V[mesh.T]                          # nt x 3 x 3
    G = np.zeros_like(mesh.V)
    for k in range(3):
        e = X[:, (k + 1) % 3] - X[:, (k + 2) % 3]
        np.add.at(G, mesh.T[:, k], gc / 3.0 + 0.5 * np.cross(e, gN))
    assert G.shape == (mesh.nv, 3) and nt == gN.shape[0]
    return G


def face_gradients(mesh: Mesh) -> np.ndarray:
    """Gradients of the three barycentric (hat) functions on each triangle,
    nt x 3 (vertex) x 3 (space): grad lambda_k = n x (x_(k+2) - x_(k+1)) / (2 a)."""
    X = mesh.V[mesh.T]
    g = np.zeros((mesh.nt, 3, 3))
    for k in range(3):
        e = X[:, (k + 2) % 3] - X[:, (k + 1) % 3]
        g[:, k] = np.cross(mesh.nrm, e) / (2.0 * mesh.area)[:, None]
    return g


def tp_metric(mesh: Mesh, q: float, p: float) -> np.ndarray:
    """Dense (3V x 3V) matrix, on vec(X) with X V x 3 row-major (entry 3 v + c),
    of <<u, v>> = B_sigma(u, v) + B0_sigma(u, v), sigma = (p - 2) / q - 1:
      B_sigma = sum_{S != T} w_ST <du_S - du_T, dv_S - dv_T>_F,
      B0      = sum_{S != T} w_ST k_ST <ubar_S - ubar_T, vbar_S - vbar_T>,
    with w_ST = a_S a_T |c_S - c_T|^(-2 sigma - 2), k_ST = (n_S . (c_S -
    c_T))^2 / |c_S - c_T|^4, du_S the face gradient (3 x 3) and ubar_S the
    face average of a P1 vector function."""
    sigma = (p - 2.0) / q - 1.0
    _, n, a, d, r2 = _tp_pairs(mesh)
    w = np.outer(a, a) * r2 ** (-(sigma + 1.0))
    np.fill_diagonal(w, 0.0)
    kst = np.einsum("sk,stk->st", n, d) ** 2 / r2 ** 2
    wt = w * (kst + kst.T)
    L = np.diag(w.sum(1)) - w
    Lt = np.diag(wt.sum(1)) - wt
    nt, nv = mesh.nt, mesh.nv
    g = face_gradients(mesh)
    # D: vec(u) -> (nt, 3 components, 3 space); A: vec(u) -> (nt, 3 components)
    D = np.zeros((nt, 3, 3, nv, 3))
    A = np.zeros((nt, 3, nv, 3))
    for k in range(3):
        for cc in range(3):
            D[np.arange(nt), cc, :, mesh.T[:, k], cc] += g[:, k, :]
            A[np.arange(nt), cc, mesh.T[:, k], cc] += 1.0 / 3.0
    # Bs = 2 sum_st L_st D_s^T D_t, B0 = sum_st Lt_st A_s^T A_t, as two products each
    D = D.reshape(nt, 9 * 3 * nv)
    A = A.reshape(nt, 3 * 3 * nv)
    LD = (L @ D).reshape(nt * 9, 3 * nv)
    LA = (Lt @ A).reshape(nt * 3, 3 * nv)
    Bs = 2.0 * D.reshape(nt * 9, 3 * nv).T @ LD
    B0 = A.reshape(nt * 3, 3 * nv).T @ LA
    return Bs + B0


def gmres_dense(A, b, tol=1e-10):
    """Reference dense solve standing in for Tensors::GMRES (not vendored).

    The iteration count and stopping rule of the reference's GMRES are not
    recoverable from the snapshot (parity unpinned for iteration counts); the
    converged solution is pinned by this dense LU.
    """
    return np.linalg.solve(A, b)


# --------------------------------------------------------------------------
# Closed forms of the unit sphere (pure mathematics: spherical Bessel / Hankel series)
# --------------------------------------------------------------------------

def sphere_eigenvalues(l, kappa):
    """Eigenvalues of S, K, K' on the unit sphere for degree l (G = e^{ikr}/(4 pi r)).

    S Y_l = i k j_l(k) h_l(k) Y_l;  K Y_l = (i k^2 j_l(k) h_l'(k) + 1/2) Y_l;
    K' Y_l = (i k^2 j_l'(k) h_l(k) - 1/2) Y_l.  kappa -> 0 limits:
    S: 1/(2l+1), K: -1/(2(2l+1)), K' = K.
    """
    from scipy.special import spherical_jn, spherical_yn

    if kappa == 0:
        lam = 1.0 / (2 * l + 1)
        return lam, -0.5 * lam, -0.5 * lam
    j = spherical_jn(l, kappa)
    jp = spherical_jn(l, kappa, derivative=True)
    h = j + 1j * spherical_yn(l, kappa)
    hp = jp + 1j * spherical_yn(l, kappa, derivative=True)
    S = 1j * kappa * j * h
    K = 1j * kappa ** 2 * j * hp + 0.5
    Kp = 1j * kappa ** 2 * jp * h - 0.5
    return S, K, Kp


def mie_far_field_soft(kappa, cos_theta, lmax=None):
    """Far field of the sound-soft unit sphere for the incident wave e^{i k x.d}.

    u_inf(xhat) = (i/k) sum_l (2l+1) j_l(k)/h_l(k) P_l(xhat.d) with the
    1/(4 pi) normalisation of G = e^{i k r} / (4 pi r) (sign fixed by the
    exp(-i k xhat.y) far-field kernel of FarFieldOperatorKernel.cl:91-101).
    """
    from scipy.special import eval_legendre, spherical_jn, spherical_yn

    if lmax is None:
        lmax = int(kappa + 4 * kappa ** (1.0 / 3.0) + 12)
    out = np.zeros_like(np.asarray(cos_theta, dtype=np.float64), dtype=np.complex128)
    for l in range(lmax + 1):
        j = spherical_jn(l, kappa)
        h = j + 1j * spherical_yn(l, kappa)
        out += (2 * l + 1) * j / h * eval_legendre(l, cos_theta)
    return -out / (1j * kappa)


def mie_neumann_soft(kappa, cos_theta, lmax=None):
    """Total-field Neumann data on the sound-soft unit sphere (outward normal):
    du/dn = (-i/k) sum_l i^l (2l+1) P_l(x.d) / h_l(k)  (Wronskian j h' - j' h = i/k^2)."""
    from scipy.special import eval_legendre, spherical_jn, spherical_yn

    if lmax is None:
        lmax = int(kappa + 4 * kappa ** (1.0 / 3.0) + 12)
    out = np.zeros_like(np.asarray(cos_theta, dtype=np.float64), dtype=np.complex128)
    for l in range(lmax + 1):
        h = spherical_jn(l, kappa) + 1j * spherical_yn(l, kappa)
        out += (1j ** l) * (2 * l + 1) * eval_legendre(l, cos_theta) / h
    return (-1j / kappa) * out


def mie_dilation_derivative(kappa, cos_theta, lmax=None):
    """Far-field derivative of the sound-soft unit sphere for h(x) = x (h.n = 1):
    DF h (xhat) = (1/k^2) sum_l (2l+1) P_l(xhat.d) / h_l(k)^2."""
    from scipy.special import eval_legendre, spherical_jn, spherical_yn

    if lmax is None:
        lmax = int(kappa + 4 * kappa ** (1.0 / 3.0) + 12)
    out = np.zeros_like(np.asarray(cos_theta, dtype=np.float64), dtype=np.complex128)
    for l in range(lmax + 1):
        h = spherical_jn(l, kappa) + 1j * spherical_yn(l, kappa)
        out += (2 * l + 1) * eval_legendre(l, cos_theta) / h ** 2
    return out / kappa ** 2


def j0(x):
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(np.abs(x) < 1e-12, 1.0, np.sin(x) / np.where(x == 0, 1, x))
