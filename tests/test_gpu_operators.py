"""GPU parity of the far-field, potential, Herglotz, incident-wave and mass
operators against the FP64 oracle (1e-5 relative; CG to its tolerance)."""
import copy
import math

import numpy as np
import pytest

import oracle
from baemm_amd import Helmholtz_HIP, WaveType
from baemm_amd.meshes import icosphere
from conftest import oracle_mesh, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-5
# the errors these cases achieve are 6e-8 to 1.4e-6: each apply is held to NEAR,
# inside the 1e-5 contract, so that a tenfold loss of accuracy fails
NEAR = 3e-6
CO = [0.5 + 0.25j, 1 - 0.5j, -0.75 + 1j, 0.3 - 0.9j]


@pytest.fixture(scope="module")
def mesh_h():
    m = oracle_mesh("perturbed", 3, 96)
    H = Helmholtz_HIP(m.V, m.T, m.meas)
    yield m, H
    H.close()


def _rand(rng, n, W):
    return rng.standard_normal((n, W)) + 1j * rng.standard_normal((n, W))


@pytest.mark.parametrize("W,wcs,kappa", [(8, 4, [3.0, 5.0]), (16, 16, [4.0]), (3, 3, [2.0])])
def test_far_field_apply(mesh_h, W, wcs, kappa):
    m, H = mesh_h
    rng = np.random.default_rng(W)
    B = _rand(rng, m.nv, W)
    F0 = _rand(rng, m.meas.shape[0], W)
    coeff = np.array(CO * (W // wcs))
    ref = oracle.apply_far_field(m, 0.7 + 0.2j, B, 0.5j, F0, kappa, coeff, W, wcs)
    F = F0.copy()
    H.ApplyFarFieldOperators_PL(0.7 + 0.2j, B, 0.5j, F, kappa, coeff, W, wcs)
    assert relerr(F, ref) < NEAR


@pytest.mark.parametrize("mode", ["fp32", "mixed", "fp32_mfma"])
def test_near_field_apply(mesh_h, mode):
    """Potential at 300 exterior points (radii 1.6-2.6) in each precision; the
    MFMA modes run the wave-specialised boundary kernel with the points as
    targets.  A second call with points close to the surface (radius 1.02) and
    alpha, beta != 0 checks the mixed-precision scale bound and the epilogue."""
    m, H = mesh_h
    H.SetPrecision(mode)
    try:
        rng = np.random.default_rng(7)
        W, wcs = 8, 4
        B = _rand(rng, m.nv, W)
        P = rng.standard_normal((300, 3))
        P *= (1.6 + rng.random(300))[:, None] / np.linalg.norm(P, axis=1, keepdims=True)
        coeff = np.array(CO * 2)
        ref = oracle.apply_near_field(m, 1.0, B, 0.0, np.zeros((300, W)), [2.0, 4.5], coeff, W, wcs, P)
        U = np.zeros((300, W), np.complex128)
        H.ApplyNearFieldOperators_PL(1.0, B, 0.0, U, [2.0, 4.5], coeff, W, wcs, P)
        print(mode, "near field", relerr(U, ref))
        assert relerr(U, ref) < NEAR
        # points near the surface, single- and double-layer rows with c2 = 0 in one chunk
        P2 = P / np.linalg.norm(P, axis=1, keepdims=True) * 1.02
        coeff2 = np.array([0, 1.0 - 0.5j, 0.0, 0, 0, 0.3j, 1.0 + 0.2j, 0], np.complex128)
        U0 = _rand(rng, 300, W)
        ref2 = oracle.apply_near_field(m, 0.5 - 1j, B, 2.0, U0, [2.0, 4.5], coeff2, W, wcs, P2)
        U2 = U0.copy()
        H.ApplyNearFieldOperators_PL(0.5 - 1j, B, 2.0, U2, [2.0, 4.5], coeff2, W, wcs, P2)
        print(mode, "near field at r = 1.02", relerr(U2, ref2))
        assert relerr(U2, ref2) < NEAR
    finally:
        H.SetPrecision("fp32")


def test_herglotz_wave(mesh_h):
    m, H = mesh_h
    rng = np.random.default_rng(3)
    W, wcs = 8, 4
    g = _rand(rng, m.meas.shape[0], W)
    C0 = _rand(rng, m.nv, W)
    coeff = np.array(CO * 2)
    ref = oracle.create_herglotz_wave(m, 1.0 - 1j, g, 2.0, C0, [2.0, 3.5], coeff, W, wcs)
    C = C0.copy()
    H.CreateHerglotzWave_PL(1.0 - 1j, g, 2.0, C, [2.0, 3.5], coeff, W, wcs)
    assert relerr(C, ref) < NEAR


@pytest.mark.parametrize("wtype", ["plane", "radial"])
def test_incident_wave(mesh_h, wtype):
    m, H = mesh_h
    rng = np.random.default_rng(5)
    W, wcs = 6, 3
    d = rng.standard_normal((wcs, 3))
    if wtype == "plane":
        d /= np.linalg.norm(d, axis=1, keepdims=True)
    else:
        d *= 0.2
    coeff = np.array([0, 1.0 - 0.3j, 0.4j, 0, 0, -2.0j, 1.0, 0])
    ref = oracle.create_incident_wave(m, 1.0, d, 0.0, np.zeros((m.nv, W)), [3.0, 5.0], coeff, W, wcs, wtype)
    C = np.zeros((m.nv, W), np.complex128)
    H.CreateIncidentWave_PL(1.0, d, 0.0, C, [3.0, 5.0], coeff, W, wcs,
                            WaveType.Plane if wtype == "plane" else WaveType.Radial)
    assert relerr(C, ref) < NEAR


def test_mass_inverse(mesh_h):
    m, H = mesh_h
    rng = np.random.default_rng(11)
    X = rng.standard_normal((m.nv, 5))
    Y = np.zeros_like(X)
    H.ApplyMassInverse(X, Y, 1e-10)
    # the reference stops after 20 PCG iterations (ApplyMassInverse.hpp:48)
    assert relerr(Y, m.mass_solve(X)) < 1e-6
    Xc = _rand(rng, m.nv, 3)
    Yc = np.zeros_like(Xc)
    H.ApplyMassInverse(Xc, Yc, 1e-10)
    assert relerr(Yc, m.mass_solve(Xc)) < 1e-6
    Yl = np.zeros_like(X)
    H.ApplyLumpedMassInverse(X, Yl)
    # areas_lumped_inv = 1 / (AvT (1/3)) = 9 / sum_{t ni v} a_t  (Initialize.hpp:337-353)
    lumped = 1.0 / (m.AvT @ np.full(m.T.shape[0], 1.0 / 3.0))
    assert relerr(Yl, lumped[:, None] * X) < 1e-6


@pytest.mark.parametrize("grid", [(60, 20), (80, 24), (240, 80)])
def test_mass_inverse_larger_meshes(grid):
    """The PCG paths of larger meshes against the exact sparse solve with the
    device's own mass matrix, with 5, 32 and 80 real columns (one, two and
    four columns per work-group).  A 60 x 20 torus (1,200 vertices) runs the
    LDS-resident block PCG on the CSR matrix with eight rows per thread; an
    80 x 24 torus (1,920 vertices) the LDS-resident block PCG on the triangle
    form of the matrix; a 240 x 80 torus (19,200 vertices) the multi-block PCG
    (direction update fused into the SpMV, column-major block partials; 80
    columns cover two 64-column chunks)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from baemm_amd.meshes import torus

    V, T = torus(1.0, 0.4, *grid)
    nv = V.shape[0]
    rng = np.random.default_rng(12)
    with Helmholtz_HIP(V, T) as H:
        ptr, idx, val = H.MassMatrix()
        M = sp.csr_matrix((val, idx, ptr), shape=(nv, nv)).tocsc()
        for X in (rng.standard_normal((nv, 5)), _rand(rng, nv, 16), _rand(rng, nv, 40)):
            Y = np.zeros_like(X)
            H.ApplyMassInverse(X, Y, 1e-10)
            assert relerr(Y, spla.spsolve(M, X)) < 1e-6
            # a second call on the same workspaces gives the same bits
            Y2 = np.zeros_like(X)
            H.ApplyMassInverse(X, Y2, 1e-10)
            assert np.array_equal(Y, Y2)


def test_mass_matrix_and_geometry(mesh_h):
    m, H = mesh_h
    ptr, idx, val = H.MassMatrix()
    import scipy.sparse as sp

    M = sp.csr_matrix((val, idx, ptr), shape=(m.nv, m.nv))
    assert relerr(M.toarray(), m.M.toarray()) < 1e-6
    assert relerr(H.TriangleAreas(), m.area) < 1e-6
    assert relerr(H.SingleLayerDiagonal(), m.sdiag) < 1e-6
    assert relerr(H.TriangleMidpoints()[:, :3], m.mid) < 1e-6
    X = np.random.default_rng(0).standard_normal((m.nv, 3))
    Y = np.ones_like(X)
    H.MassDot(2.0, X, 1.0, Y)
    assert relerr(Y, 2.0 * (m.M @ X) + 1.0) < 1e-6


def test_hardware_sine_accuracy_over_the_accepted_phase_range():
    """The pair kernels take e^{i kappa r} from the hardware sine and cosine
    (argument in revolutions) without a range reduction of their own.  Over the
    phases the library accepts (|kappa r / 2 pi| <= 65,536 revolutions; beyond
    that baemm_load_parameters refuses kappa) they stay within 5e-7 of
    sin/cos(2 pi t) evaluated in FP64 at the same FP32 argument t."""
    import ctypes

    from baemm_amd import _lib

    rng = np.random.default_rng(12)
    t = np.concatenate([
        rng.uniform(-1.0, 1.0, 20000),
        rng.uniform(-300.0, 300.0, 20000),
        rng.uniform(-65536.0, 65536.0, 40000),
        np.arange(-65536, 65537, 97, dtype=np.float64) + 0.125,
        [65535.75, -65535.75, 0.25, 0.5, 0.75, 1e-8],
    ]).astype(np.float32)
    s = np.zeros_like(t)
    c = np.zeros_like(t)
    L = _lib.load()
    st = L.baemm_hw_sincos(0, t.ctypes.data_as(ctypes.c_void_p), t.size, s.ctypes.data_as(ctypes.c_void_p),
                           c.ctypes.data_as(ctypes.c_void_p))
    assert st == 0
    t64 = t.astype(np.float64)
    es = np.abs(s - np.sin(2 * np.pi * t64))
    ec = np.abs(c - np.cos(2 * np.pi * t64))
    for lo, hi in ((0, 1), (1, 256), (256, 65537)):
        sel = (np.abs(t64) >= lo) & (np.abs(t64) < hi)
        print(f"|t| in [{lo}, {hi}): max error sin {es[sel].max():.2e} cos {ec[sel].max():.2e}")
    assert es.max() < 5e-7 and ec.max() < 5e-7


def test_wave_numbers_beyond_the_phase_range_are_refused():
    """kappa * mesh extent / (2 pi) above 65,536 revolutions is an argument
    error (the FP32 phase would resolve less than 2^-8 revolution there)."""
    from baemm_amd import BaemmError

    V, T = icosphere(1)
    with Helmholtz_HIP(V, T) as H:
        H.LoadBoundaryOperators_PL([1.0e5], [0, 1, 0, 0], 1, 1)   # extent 2 sqrt(3): about 55,000 revolutions
        with pytest.raises(BaemmError):
            H.LoadBoundaryOperators_PL([1.0e6], [0, 1, 0, 0], 1, 1)
        with pytest.raises(BaemmError):
            H.LoadBoundaryOperators_PL([float("nan")], [0, 1, 0, 0], 1, 1)


def test_potential_on_a_volume_grid_at_scale():
    """The potential (ApplyNearFieldOperators_PL) of the 81,920-triangle sphere
    at 65,536 points: a 60 x 32 x 32 box grid over [-1.5, 1.5]^3 (interior and
    exterior points) plus 4,096 points at distances -0.05 to 0.05 from the
    surface along the triangle normals.  Two wave numbers x 16 columns, single
    and double layer.  Every precision against the FP64 oracle on 512 sampled
    points of each kind (grid and shell), within 1e-5 relative.  Grid points
    come within 0.0023 of a centroid (an eighth of the mesh size), where
    rounding the geometry and the points to FP32 alone moves 1/r by about
    3e-5; against the oracle evaluated on the FP32-rounded centroids, normals
    and points (the reference's FP32 geometry, Helmholtz_OpenCL.hpp:187) the
    kernels are held to 4e-6 (measured 1.2e-6 to 2.1e-6)."""
    from baemm_amd.meshes import shell_points, volume_grid

    V, T = icosphere(6)
    m = oracle.Mesh(V, T)
    grid = volume_grid((60, 32, 32), (-1.5, -1.5, -1.5), (1.5, 1.5, 1.5))
    shell = shell_points(V, T, [-0.05, -0.02, 0.02, 0.05], 1024, seed=5)
    P = np.concatenate([grid, shell])
    assert P.shape[0] == 65536
    rng = np.random.default_rng(13)
    W, wcs = 32, 16
    B = _rand(rng, m.nv, W)
    coeff = np.array([0, 1.0, -0.5 + 1j, 0, 0, 0.5 - 0.25j, 1.0, 0], np.complex128)
    sel = {"grid": rng.choice(grid.shape[0], 512, replace=False),
           "shell": grid.shape[0] + rng.choice(shell.shape[0], 512, replace=False)}
    refs = {k: oracle.apply_near_field(m, 1.0, B, 0.0, np.zeros((512, W)), [3.0, 6.0], coeff, W, wcs, P[idx])
            for k, idx in sel.items()}
    # the same FP64 arithmetic on the geometry and points the kernels see (FP32)
    m32 = copy.copy(m)
    m32.mid = m.mid.astype(np.float32).astype(np.float64)
    m32.nrm = m.nrm.astype(np.float32).astype(np.float64)
    P32 = P.astype(np.float32).astype(np.float64)
    refs32 = {k: oracle.apply_near_field(m32, 1.0, B, 0.0, np.zeros((512, W)), [3.0, 6.0], coeff, W, wcs, P32[idx])
              for k, idx in sel.items()}
    with Helmholtz_HIP(V, T) as H:
        for mode in ("fp32", "mixed", "fp32_mfma"):
            H.SetPrecision(mode)
            U = np.zeros((P.shape[0], W), np.complex128)
            H.ApplyNearFieldOperators_PL(1.0, B, 0.0, U, [3.0, 6.0], coeff, W, wcs, P)
            assert np.all(np.isfinite(U))
            for k, idx in sel.items():
                e, e32 = relerr(U[idx], refs[k]), relerr(U[idx], refs32[k])
                print(f"volume potential {mode} {k}: rel. error vs FP64 on 512 points {e:.2e}, "
                      f"vs FP64 on the FP32 geometry {e32:.2e}")
                assert e < TOL and e32 < 4e-6, (mode, k)
