"""GPU solver compositions (device GMRES + PCG) against dense FP64 oracle
solves, the Mie series and the defining equations of GaussNewtonSolve."""
import numpy as np
import pytest

import oracle
from baemm_amd import Helmholtz_HIP
from baemm_amd.meshes import fibonacci_sphere, icosphere, write_mesh_file
from conftest import oracle_mesh, relerr

pytestmark = pytest.mark.gpu

DIRS = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8], [0.0, 0.6, -0.8]])


@pytest.fixture(scope="module")
def small():
    m = oracle_mesh("sphere", 2, 64)
    H = Helmholtz_HIP(m.V, m.T, m.meas)
    H.SetGmresLimits(60, 4)
    yield m, H
    H.close()


def test_dirichlet_to_neumann_matches_dense_solve(small):
    m, H = small
    rng = np.random.default_rng(0)
    W, wcs = 6, 3
    kappa = [2.0, 3.0]
    wave = rng.standard_normal((m.nv, W)) + 1j * rng.standard_normal((m.nv, W))
    ref = oracle.dtn(m, kappa, wave, wcs)
    for mode in ("fp32", "mixed", "fp32_mfma"):
        H.SetPrecision(mode)
        psi, ok = H.DirichletToNeumann(kappa, wave, wcs, 1e-12, 1e-10)
        assert ok
        assert relerr(psi, ref) < 2e-5, mode
    H.SetPrecision("fp32")


def test_boundary_potential_and_coefficients(small):
    m, H = small
    rng = np.random.default_rng(1)
    wave = rng.standard_normal((m.nv, 4)) + 1j * rng.standard_normal((m.nv, 4))
    phi, coeff, ok = H.BoundaryPotential([2.5], wave, 4, 1e-12, 1e-10, eta=[1.5])
    assert ok
    np.testing.assert_allclose(coeff[0], [0.5, -1.5j, 1.0, 0.0])
    ref = oracle.solve_boundary(m, [2.5], [np.array([0.5, -1.5j, 1.0, 0.0])], wave, 4)
    assert relerr(phi, ref) < 2e-5


def test_far_field_matches_oracle_and_mie(small):
    m, H = small
    F = H.FarField([2.0], 1, DIRS, 3, cg_tol=1e-12, gmres_tol=1e-10)
    ref, _ = oracle.far_field(m, [2.0], DIRS, 3)
    assert relerr(F, ref) < 2e-5


def test_derivatives_match_oracle(small):
    m, H = small
    kappa, wcs = [2.0], 3
    du = H.Derivative_FF_Helper(kappa, 1, DIRS, wcs, cg_tol=1e-12, gmres_tol=1e-10)
    du_ref = oracle.derivative_ff_helper(m, kappa, DIRS, wcs)
    assert relerr(du, du_ref) < 2e-5
    rng = np.random.default_rng(2)
    X = rng.standard_normal((m.nv, 3))
    Y, du2 = H.Derivative_FF(kappa, 1, DIRS, wcs, X, du_dn=None, cg_tol=1e-12, gmres_tol=1e-10)
    assert relerr(du2, du_ref) < 2e-5          # du_dn allocated and filled when not given
    Yref, _ = oracle.derivative_ff(m, kappa, DIRS, wcs, X, du_ref)
    assert relerr(Y, Yref) < 5e-5
    g = rng.standard_normal((m.meas.shape[0], 3)) + 1j * rng.standard_normal((m.meas.shape[0], 3))
    Z, _ = H.AdjointDerivative_FF(kappa, 1, DIRS, wcs, g, du_dn=du, cg_tol=1e-12, gmres_tol=1e-10)
    Zref, _ = oracle.adjoint_derivative_ff(m, kappa, DIRS, wcs, g, du_ref)
    assert relerr(Z, Zref) < 5e-5


def test_gauss_newton_solves_its_equation(small):
    """X = (c I + Mass DF* DF)^{-1} B; check the residual with the GPU's own DF, DF*."""
    m, H = small
    kappa, wcs = [2.0], 3
    c = 0.05
    du = H.Derivative_FF_Helper(kappa, 1, DIRS, wcs, cg_tol=1e-12, gmres_tol=1e-10)

    def Mop(x, y):
        y[:] = c * x

    def Pop(x, y):
        y[:] = x / c

    rng = np.random.default_rng(4)
    B = rng.standard_normal((m.nv, 3))
    X = np.zeros_like(B)
    ok, _ = H.GaussNewtonSolve(kappa, 1, DIRS, wcs, Mop, Pop, 1.0, B, 0.0, X, du_dn=du, cg_tol=1e-12,
                               gmres_tol_inner=1e-10, gmres_tol_outer=1e-7)
    assert ok
    DFX, _ = H.Derivative_FF(kappa, 1, DIRS, wcs, X, du_dn=du, cg_tol=1e-12, gmres_tol=1e-10)
    Z, _ = H.AdjointDerivative_FF(kappa, 1, DIRS, wcs, DFX, du_dn=du, cg_tol=1e-12, gmres_tol=1e-10)
    res = c * X + m.M @ Z - B
    print('GN stats', H.SolverStats(), np.linalg.norm(res) / np.linalg.norm(B))
    assert np.linalg.norm(res) < 1e-4 * np.linalg.norm(B)


def test_gauss_newton_matches_dense_fp64_oracle():
    """GaussNewtonSolve (FarField.hpp:457-604) against the oracle's dense FP64
    solve of (M_user + Mass DF* DF) X = B, with DF and DF* the oracle's own
    compositions: two wave numbers, a metric c I + d Mass (x) I_3, B and X in
    scattered storage (ldB = 5, ldX = 4, FarField.hpp:516-536,581-599), and
    alpha/beta in the combine step; columns 3.. of X are left untouched."""
    m = oracle_mesh("sphere", 2, 64)
    kappa, wcs = [2.0, 3.5], 2
    dirs = DIRS[:2]
    c, d = 0.05, 0.5
    metric = c * np.eye(3 * m.nv) + d * np.kron(m.M.toarray(), np.eye(3))
    rng = np.random.default_rng(8)
    B = rng.standard_normal((m.nv, 3))
    X0 = rng.standard_normal((m.nv, 3))
    alpha, beta = 0.5, 2.0
    ref, du_ref = oracle.gauss_newton(m, kappa, dirs, wcs, metric, B, alpha, beta, X0)
    lumped = np.asarray(m.M.sum(axis=1)).ravel()
    with Helmholtz_HIP(m.V, m.T, m.meas) as H:
        H.SetGmresLimits(60, 4)

        def Mop(x, y):
            y[:] = c * x + d * (m.M @ x)

        def Pop(x, y):
            y[:] = x / (c + d * lumped)[:, None]

        Bw = np.zeros((m.nv, 5))
        Bw[:, :3] = B
        Xw = np.full((m.nv, 4), 7.0)
        Xw[:, :3] = X0
        for mode in ("fp32", "mixed"):
            H.SetPrecision(mode)
            Xw[:, :3] = X0
            ok, du = H.GaussNewtonSolve(kappa, 2, dirs, wcs, Mop, Pop, alpha, Bw[:, :3], beta, Xw[:, :3],
                                        du_dn=None, cg_tol=1e-12, gmres_tol_inner=1e-10, gmres_tol_outer=1e-9)
            assert ok
            e_du, e_x = relerr(du, du_ref), relerr(Xw[:, :3], ref)
            print(f"Gauss-Newton {mode}: du_dn rel. err {e_du:.2e}, X rel. err {e_x:.2e}", H.SolverStats())
            assert e_du < 2e-5 and e_x < 1e-4, mode
            assert np.all(Xw[:, 3] == 7.0)


def test_gauss_newton_on_the_device_matches_host_callbacks_and_oracle():
    """baemm_gauss_newton_solve_dev: the outer GMRES with its Krylov basis in
    HBM and the device metric c I + d (Mass (x) I_3) with its Jacobi
    preconditioner (baemm_mass_metric_apply / _jacobi), no host callback.
    Against the oracle's dense FP64 Gauss-Newton solve and against the same
    solve with the metric as host callbacks (baemm_gauss_newton_solve), with
    B and X in scattered storage (ldB = 5, ldX = 4) and alpha/beta; a second
    call that passes the returned du_dn (not recomputed) gives the same X."""
    m = oracle_mesh("sphere", 2, 64)
    kappa, wcs = [2.0, 3.5], 2
    dirs = DIRS[:2]
    c, d = 0.05, 0.5
    metric = c * np.eye(3 * m.nv) + d * np.kron(m.M.toarray(), np.eye(3))
    rng = np.random.default_rng(18)
    B = rng.standard_normal((m.nv, 3))
    X0 = rng.standard_normal((m.nv, 3))
    alpha, beta = 0.5, 2.0
    ref, du_ref = oracle.gauss_newton(m, kappa, dirs, wcs, metric, B, alpha, beta, X0)
    lumped = np.asarray(m.M.sum(axis=1)).ravel()
    with Helmholtz_HIP(m.V, m.T, m.meas) as H:
        H.SetGmresLimits(60, 4)
        for mode in ("fp32", "mixed"):
            H.SetPrecision(mode)
            Bw = np.zeros((m.nv, 5))
            Bw[:, :3] = B
            Xd = np.full((m.nv, 4), 7.0)
            Xd[:, :3] = X0
            ok, du = H.GaussNewtonSolveDevice(kappa, 2, dirs, wcs, c, d, alpha, Bw[:, :3], beta, Xd[:, :3],
                                              cg_tol=1e-12, gmres_tol_inner=1e-10, gmres_tol_outer=1e-9)
            assert ok
            assert np.all(Xd[:, 3] == 7.0)
            Xh = X0.copy()
            ok_h, _ = H.GaussNewtonSolve(kappa, 2, dirs, wcs, lambda x, y: y.__setitem__(slice(None), c * x + d * (m.M @ x)),
                                         lambda x, y: y.__setitem__(slice(None), x / (c + d * lumped)[:, None]),
                                         alpha, B, beta, Xh, du_dn=du, cg_tol=1e-12, gmres_tol_inner=1e-10,
                                         gmres_tol_outer=1e-9)
            assert ok_h
            Xd2 = X0.copy()
            ok2, _ = H.GaussNewtonSolveDevice(kappa, 2, dirs, wcs, c, d, alpha, B, beta, Xd2, du_dn=du,
                                              cg_tol=1e-12, gmres_tol_inner=1e-10, gmres_tol_outer=1e-9)
            e_x, e_du = relerr(Xd[:, :3], ref), relerr(du, du_ref)
            e_host, e_again = relerr(Xd[:, :3], Xh), relerr(Xd2, Xd[:, :3])
            print(f"device Gauss-Newton {mode}: X vs FP64 {e_x:.2e}, du_dn {e_du:.2e}, vs host callbacks {e_host:.2e}, "
                  f"with du_dn passed {e_again:.2e}", H.SolverStats())
            assert ok2
            assert e_du < 2e-5 and e_x < 1e-4, mode
            assert e_host < 1e-6 and e_again < 1e-6, mode


def test_far_field_mie_refinement():
    """Sound-soft unit sphere: far field error vs the Mie series shrinks under refinement."""
    errs = []
    meas = fibonacci_sphere(128)
    for L in (3, 4):
        V, T = icosphere(L)
        with Helmholtz_HIP(V, T, meas) as H:
            H.SetPrecision("mixed")
            F = H.FarField([4.0], 1, DIRS, 3, cg_tol=1e-10, gmres_tol=1e-8)
        ref = np.stack([oracle.mie_far_field_soft(4.0, meas @ d) for d in DIRS], axis=1)
        errs.append(relerr(F, ref))
    print("Mie errors", errs)
    assert errs[1] < 0.6 * errs[0] and errs[1] < 0.02


def test_cpp_dropin_example_matches_python(tmp_path):
    """The C++ drop-in class (include/Helmholtz_HIP.hpp) through the example
    program gives the same far-field norm as the Python mirror."""
    import re
    import subprocess
    from pathlib import Path

    exe = Path(__file__).resolve().parents[1] / "baemm_amd" / "examples" / "example_far_field"
    if not exe.exists():
        pytest.fail(f"{exe} not built")
    V, T = icosphere(3)
    Vm, Tm = icosphere(4)
    write_mesh_file(tmp_path / "obstacle.txt", V, T)
    write_mesh_file(tmp_path / "meas.txt", Vm, Tm)
    out = subprocess.run([str(exe), str(tmp_path / "obstacle.txt"), str(tmp_path / "meas.txt"), "--out",
                          str(tmp_path)], capture_output=True, text=True, timeout=240)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    norm_cpp = float(re.search(r"L\^2-norm of C = (\S+)", out.stdout).group(1))
    inc = np.array([[0.22663516023574246, 0.4654289185373517, 0.8555772472045235],
                    [-0.3474510994575331, -0.7864451523456505, -0.5106679506663584],
                    [-0.9541836471144971, -0.27451273343735705, -0.11906438073591856],
                    [0.43775029208027166, -0.8876694056416434, 0.1428905457734996],
                    [-0.44907802104219113, -0.8139693295211834, 0.36848726113078095],
                    [0.47666884531973414, 0.3192753380458682, -0.8190543757390276],
                    [-0.02194389227182042, 0.8818743114784301, -0.47097363444932655],
                    [0.786615210967289, -0.1684579992437276, 0.5940188653281036],
                    [0.42244725439759406, -0.5580931680624308, -0.7141920841160128],
                    [0.6792359217126083, 0.7269051927489542, 0.10122945919953094],
                    [-0.3904321861768485, 0.03454064212408624, -0.919983506394991],
                    [-0.3009524704632311, 0.8792870632871999, 0.36916374531886265],
                    [-0.006080463353741372, -0.38604914872328944, 0.9224581739761577],
                    [0.9742753958166932, -0.0853848807263007, -0.20855904499585115],
                    [-0.8068074072061289, 0.5272347230598002, -0.2666183686046023],
                    [-0.7266826985737691, 0.12599246204091846, 0.675320779409617]])
    with Helmholtz_HIP(V, T, Vm) as H:
        H.UseDiagonal(True)
        F = H.FarField([5 * np.pi], 1, inc, 16, cg_tol=1e-5, gmres_tol=1e-4)
        DFh, du = H.Derivative_FF([5 * np.pi], 1, inc, 16, V.copy(), cg_tol=1e-5, gmres_tol=1e-4)
        w, _ = H.AdjointDerivative_FF([5 * np.pi], 1, inc, 16, DFh, du_dn=du, cg_tol=1e-5, gmres_tol=1e-4)
    norm_py = np.sqrt(2 * np.pi / (Vm.shape[0] * 16)) * np.linalg.norm(F)
    assert abs(norm_cpp - norm_py) < 1e-6 * norm_py
    # the result files of Example_OpenCL/main.cpp:191,220,233 hold the same arrays
    def cplx(name):
        a = np.loadtxt(tmp_path / name)
        return a[:, 0::2] + 1j * a[:, 1::2]

    assert relerr(cplx("F_New.txt"), F) < 1e-12
    assert relerr(cplx("DF_New.txt"), DFh) < 1e-12
    assert relerr(np.loadtxt(tmp_path / "DFT_New.txt"), w) < 1e-12


def test_graph_captured_arnoldi_steps_match_eager(small):
    """GMRES with each Arnoldi step replayed from a captured hipGraph gives the
    eager result exactly, on every replay (eager work and synchronisation in
    between), in every precision.  The small sphere (162 vertices) runs the
    one-launch LDS-resident PCG; the torus with 3,000 vertices (its mass matrix
    needs 162 KB in the triangle form, more than the 150 KB the LDS PCG takes)
    runs the multi-block PCG: device-side convergence flag, ticket and
    active-column counters reset every iteration, double-buffered directions,
    launches that return at once after convergence."""
    from baemm_amd.meshes import torus

    m, H = small
    _graph_vs_eager(H, m.nv)
    V, T = torus(1.0, 0.4, 100, 30)
    assert V.shape[0] == 3000
    with Helmholtz_HIP(V, T) as Ht:
        _graph_vs_eager(Ht, V.shape[0])


def _graph_vs_eager(H, nv):
    rng = np.random.default_rng(9)
    W, wcs = 6, 3
    kappa = [2.0, 3.0]
    wave = rng.standard_normal((nv, W)) + 1j * rng.standard_normal((nv, W))
    for mode in ("fp32", "mixed", "fp32_mfma"):
        H.SetPrecision(mode)
        H.SetGraphs(False)
        eager, ok = H.DirichletToNeumann(kappa, wave, wcs, 1e-12, 1e-10)
        assert ok
        st0 = H.GraphStats()
        H.SetGraphs(True)
        for rep in range(3):
            got, ok = H.DirichletToNeumann(kappa, wave, wcs, 1e-12, 1e-10)
            assert ok
            assert np.array_equal(got, eager), (mode, rep)
            # eager work between the replays
            H.SetGraphs(False)
            again, _ = H.DirichletToNeumann(kappa, wave, wcs, 1e-12, 1e-10)
            assert np.array_equal(again, eager), (mode, rep)
            H.SetGraphs(True)
        st = H.GraphStats()
        print(mode, "graph stats", st)
        assert st["launches"] - st0["launches"] > 0
        assert st["captures"] - st0["captures"] > 0
        # replays reuse the captured graphs: far fewer captures than launches
        assert st["captures"] - st0["captures"] < st["launches"] - st0["launches"]
    H.SetGraphs(False)
    H.SetPrecision("fp32")


def test_c3_torus_small_matches_dense_fp64_solve():
    """Sound-soft torus (R=1, r=0.4), far field from GMRES(30) at 1e-6 against
    the oracle's dense FP64 LU solve (torus small enough for LU)."""
    m = oracle_mesh("torus", 10, 256)
    dirs = np.array([[0.0, 0.0, 1.0], [0.6, 0.8, 0.0], [0.0, 0.6, -0.8], [-0.48, 0.6, 0.64]])
    ref, _ = oracle.far_field(m, [4.0], dirs, 4)
    with Helmholtz_HIP(m.V, m.T, m.meas) as H:
        for mode in ("fp32", "mixed", "fp32_mfma"):
            H.SetPrecision(mode)
            F = H.FarField([4.0], 1, dirs, 4, cg_tol=1e-10, gmres_tol=1e-6)
            err = relerr(F, ref)
            print("torus", mode, err, H.SolverStats())
            assert err < 2e-5, mode


def test_c3_torus_full_size_mixed_vs_fp32():
    """The 38,400-triangle torus at kappa = 10 with 32 seeded plane waves and 256
    Fibonacci far-field directions: the mixed-precision solve (MFMA operator,
    FP32-operator refinement at restarts) agrees with the plain FP32 solve;
    both converge to 1e-6."""
    from baemm_amd.meshes import random_directions, torus

    V, T = torus(1.0, 0.4, 240, 80)
    assert T.shape[0] == 38400
    meas = fibonacci_sphere(256)
    dirs = random_directions(32, seed=2)
    out = {}
    with Helmholtz_HIP(V, T, meas) as H:
        for mode in ("fp32", "mixed", "fp32_mfma"):
            H.SetPrecision(mode)
            H.ResetSolverStats()
            out[mode] = H.FarField([10.0], 1, dirs, 32, cg_tol=1e-10, gmres_tol=1e-6)
            st = H.SolverStats()
            print("torus 38400", mode, st)
            assert st["rel_residual"] <= 1e-6
    for mode in ("mixed", "fp32_mfma"):
        d = relerr(out[mode], out["fp32"])
        print(mode, "vs fp32 far field", d)
        assert d < 1e-4


def test_c3b_icosphere_81920_against_mie():
    """Sound-soft unit sphere with 81,920 triangles at kappa = 10: far field of
    the mixed-precision solve against the Mie series."""
    V, T = icosphere(6)
    meas = fibonacci_sphere(256)
    with Helmholtz_HIP(V, T, meas) as H:
        H.SetPrecision("mixed")
        F = H.FarField([10.0], 1, DIRS, 3, cg_tol=1e-10, gmres_tol=1e-6)
        print("stats", H.SolverStats())
    ref = np.stack([oracle.mie_far_field_soft(10.0, meas @ d) for d in DIRS], axis=1)
    err = relerr(F, ref)
    print("Mie error 81920", err)
    assert err < 5e-3


def test_cpp_dropin_surface_checks():
    """Every method of the C++ drop-in class on a small sphere: overloads that
    must agree, M (M^-1 X) = X, the solvers' equations, the Gauss-Newton
    residual (baemm_amd/examples/dropin_check.cpp)."""
    import subprocess
    from pathlib import Path

    exe = Path(__file__).resolve().parents[1] / "baemm_amd" / "examples" / "dropin_check"
    if not exe.exists():
        pytest.fail(f"{exe} not built")
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "FAIL" not in out.stdout
    assert out.stdout.count("PASS") >= 20
