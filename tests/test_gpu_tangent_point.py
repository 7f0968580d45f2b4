"""The tangent-point energy, gradient and metric on the device (the caller side
of the Gauss-Newton example, Example_GaussNewton/main.cpp:301-346) against
the oracle's FP64 restatement, and a device Gauss-Newton solve with the
metric as device callbacks against the oracle's dense solve."""
import numpy as np
import pytest

import oracle
from baemm_amd import Helmholtz_HIP
from baemm_amd.meshes import fibonacci_sphere, icosphere, perturbed_sphere, torus
from conftest import relerr

pytestmark = pytest.mark.gpu

MEAS = fibonacci_sphere(16)


def _meshes():
    # 20 triangles (one source segment), 320 and 1280 (several segments, a
    # partial last tile), and a torus (genus one, open-looking neighbourhoods)
    yield "icosahedron", icosphere(0)
    yield "perturbed-320", perturbed_sphere(2, seed=11)
    yield "perturbed-1280", perturbed_sphere(3, seed=12)
    yield "torus-288", torus(1.0, 0.4, 24, 6)


@pytest.mark.parametrize("q,p", [(6.0, 12.0), (3.5, 8.25)])
def test_energy_and_gradient_match_oracle(q, p):
    for name, (V, T) in _meshes():
        m = oracle.Mesh(V, T)
        E_ref = oracle.tp_energy(m, q, p)
        G_ref = oracle.tp_energy_gradient(m, q, p)
        with Helmholtz_HIP(V, T, MEAS) as H:
            E, G = H.TangentPointEnergy(q, p)
            E2, none = H.TangentPointEnergy(q, p, gradient=False)
        e_E, e_G = abs(E - E_ref) / abs(E_ref), relerr(G, G_ref)
        print(f"{name} q={q} p={p}: energy {e_E:.2e}, gradient {e_G:.2e}")
        assert none is None and E2 == E
        assert e_E < 1e-5 and e_G < 5e-5, name


def test_metric_apply_matches_oracle():
    q, p = 6.0, 12.0
    rng = np.random.default_rng(21)
    for name, (V, T) in _meshes():
        m = oracle.Mesh(V, T)
        M = oracle.tp_metric(m, q, p)
        x = rng.standard_normal((m.nv, 3))
        with Helmholtz_HIP(V, T, MEAS) as H:
            y = H.TangentPointMetric(x, q, p, scale=0.25)
            # translations lie in the metric's kernel
            t = np.zeros((m.nv, 3))
            t[:, 1] = 1.0
            yt = H.TangentPointMetric(t, q, p)
            # a second exponent pair on the same handle
            y2 = H.TangentPointMetric(x, 3.5, 8.25)
        ref = 0.25 * (M @ x.ravel()).reshape(-1, 3)
        ref2 = (oracle.tp_metric(m, 3.5, 8.25) @ x.ravel()).reshape(-1, 3)
        e, e2 = relerr(y, ref), relerr(y2, ref2)
        e_t = np.abs(yt).max() / (np.abs(M).max() * m.nv)
        print(f"{name}: metric apply {e:.2e}, second (q, p) {e2:.2e}, translation {e_t:.2e}")
        assert e < 5e-6 and e2 < 5e-6, name
        assert e_t < 1e-5, name


def test_metric_diagonal_preconditioner():
    """baemm_tp_metric_jacobi: P x = x / (scale d_v), d_v the diagonal of the
    metric's matrix (the oracle's dense matrix, entries (3v + c, 3v + c))."""
    import ctypes

    from baemm_amd import _lib

    q, p, scale = 6.0, 12.0, 0.5
    V, T = perturbed_sphere(2, seed=13)
    m = oracle.Mesh(V, T)
    Md = np.diag(oracle.tp_metric(m, q, p)).reshape(-1, 3)
    assert np.abs(Md - Md[:, :1]).max() < 1e-12 * Md.max()
    x = np.random.default_rng(22).standard_normal((m.nv, 3))
    with Helmholtz_HIP(V, T, MEAS) as H:
        L = H._L
        H._check(L.baemm_tp_metric_prepare(H._h, q, p))
        bufs = []
        try:
            for a in (x, np.zeros_like(x)):
                ptr = ctypes.c_void_p()
                _lib.check(L.baemm_dev_alloc(ctypes.byref(ptr), a.nbytes))
                _lib.check(L.baemm_dev_upload(ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))
                bufs.append(ptr)
            met = _lib.TPMetric(H._h.value, q, p, scale)
            L.baemm_tp_metric_jacobi(ctypes.byref(met), bufs[0], bufs[1], None)
            _lib.check(L.baemm_dev_synchronize())
            y = np.zeros_like(x)
            _lib.check(L.baemm_dev_download(y.ctypes.data_as(ctypes.c_void_p), bufs[1], y.nbytes))
        finally:
            for ptr in bufs:
                L.baemm_dev_free(ptr)
    ref = x / (scale * Md)
    print(f"Jacobi of the tangent-point metric: {relerr(y, ref):.2e}")
    assert relerr(y, ref) < 1e-6


def test_gauss_newton_with_tangent_point_metric_matches_oracle():
    """GaussNewtonSolveDevice with M = regpar times the tangent-point metric
    and its diagonal as preconditioner, both device callbacks, as the
    Gauss-Newton example composes them (main.cpp:324-346), against the
    oracle's dense FP64 solve with the dense metric."""
    V, T = perturbed_sphere(2, seed=14)
    m = oracle.Mesh(V, T, fibonacci_sphere(48))
    q, p = 6.0, 12.0
    kappa, wcs = [2.0, 3.0], 2
    dirs = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8]])
    TP = oracle.tp_metric(m, q, p)
    A0, du0 = oracle.gauss_newton_operator(m, kappa, dirs, wcs, np.zeros_like(TP))
    regpar = 0.3 * np.abs(A0).max() / np.abs(TP).max()
    rng = np.random.default_rng(23)
    B = rng.standard_normal((m.nv, 3))
    X0 = rng.standard_normal((m.nv, 3))
    alpha, beta = 1.0, 0.5
    ref, _ = oracle.gauss_newton(m, kappa, dirs, wcs, regpar * TP, B, alpha, beta, X0, du_dn=du0)
    with Helmholtz_HIP(m.V, m.T, m.meas) as H:
        H.SetGmresLimits(100, 8)
        for mode in ("fp32", "mixed"):
            H.SetPrecision(mode)
            X = X0.copy()
            ok, du = H.GaussNewtonSolveDevice(kappa, 2, dirs, wcs, 0.0, 0.0, alpha, B, beta, X,
                                              cg_tol=1e-12, gmres_tol_inner=1e-10, gmres_tol_outer=1e-9,
                                              tangent_point=(q, p, regpar))
            e_x, e_du = relerr(X, ref), relerr(du, du0)
            print(f"Gauss-Newton with the tangent-point metric, {mode}: X {e_x:.2e}, du_dn {e_du:.2e}",
                  H.SolverStats())
            assert ok, mode
            assert e_du < 2e-5 and e_x < 1e-5, mode


def test_properties_at_scale():
    """At 20,480 triangles (beyond the oracle's dense sizes), through
    size-independent properties: the metric is symmetric (u . M v = v . M u)
    and has the translations in its kernel, and the energy gradient matches a
    central finite difference of the device energy along a random smooth
    direction."""
    V, T = perturbed_sphere(5, seed=31)
    q, p = 6.0, 12.0
    rng = np.random.default_rng(32)
    u = rng.standard_normal(V.shape)
    v = rng.standard_normal(V.shape)
    with Helmholtz_HIP(V, T, MEAS) as H:
        Mu = H.TangentPointMetric(u, q, p)
        Mv = H.TangentPointMetric(v, q, p)
        t = np.zeros_like(V)
        t[:, 0] = 1.0
        Mt = H.TangentPointMetric(t, q, p)
        E, G = H.TangentPointEnergy(q, p)
    sym = abs((u * Mv).sum() - (v * Mu).sum()) / np.sqrt((u * Mu).sum() * (v * Mv).sum())
    trans = np.linalg.norm(Mt) / (np.linalg.norm(Mu) / np.linalg.norm(u) * np.linalg.norm(t))
    # a smooth direction: a rotation-free stretch of the surface
    d = np.stack([V[:, 0] * V[:, 1], V[:, 2] ** 2, np.sin(V[:, 0])], axis=1)
    h = 5e-4
    with Helmholtz_HIP(V + h * d, T, MEAS) as Hp, Helmholtz_HIP(V - h * d, T, MEAS) as Hm:
        Ep, _ = Hp.TangentPointEnergy(q, p, gradient=False)
        Em, _ = Hm.TangentPointEnergy(q, p, gradient=False)
    fd = (Ep - Em) / (2 * h)
    dg = (G * d).sum()
    print(f"20,480 triangles: E {E:.6e}, symmetry {sym:.2e}, translation {trans:.2e}, "
          f"FD {fd:.6e} vs gradient {dg:.6e} ({abs(fd - dg) / abs(dg):.2e})")
    assert sym < 1e-8 and trans < 1e-6
    assert abs(fd - dg) < 3e-4 * abs(dg)


def test_host_callbacks_with_the_tangent_point_metric():
    """GaussNewtonSolve with host callbacks that apply the tangent-point metric,
    as Example_GaussNewton/main.cpp:324-346 composes them: a call on the
    solving handle from inside its own callback is refused with a state error
    (it would wait for the solve that runs the callback), and a second handle
    on the same mesh serves the callbacks; the result matches the device
    callbacks (baemm_tp_metric_apply / _jacobi)."""
    from baemm_amd import BaemmError

    V, T = perturbed_sphere(2, seed=14)
    meas = fibonacci_sphere(48)
    q, p, regpar = 6.0, 12.0, 2e-4
    kappa, wcs = [2.0], 2
    dirs = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8]])
    B = np.random.default_rng(33).standard_normal((V.shape[0], 3))
    refused = []
    with Helmholtz_HIP(V, T, meas) as H, Helmholtz_HIP(V, T, meas) as Hm:
        H.SetGmresLimits(100, 8)

        def M(x, y):
            try:
                H.TangentPointMetric(x, q, p, regpar)
            except BaemmError as e:
                refused.append(e.status)
            y[:] = Hm.TangentPointMetric(x, q, p, regpar)

        def P(x, y):
            y[:] = Hm.TangentPointPreconditioner(x, q, p, regpar)

        X = np.zeros_like(B)
        ok, du = H.GaussNewtonSolve(kappa, 1, dirs, wcs, M, P, 1.0, B, 0.0, X, cg_tol=1e-12,
                                    gmres_tol_inner=1e-10, gmres_tol_outer=1e-9)
        Xd = np.zeros_like(B)
        ok_d, _ = H.GaussNewtonSolveDevice(kappa, 1, dirs, wcs, 0.0, 0.0, 1.0, B, 0.0, Xd, du_dn=du, cg_tol=1e-12,
                                           gmres_tol_inner=1e-10, gmres_tol_outer=1e-9, tangent_point=(q, p, regpar))
    print(f"host callbacks vs device callbacks: {relerr(X, Xd):.2e}; refused calls: {len(refused)}")
    assert ok and ok_d
    assert refused and set(refused) == {3}
    assert relerr(X, Xd) < 1e-6
