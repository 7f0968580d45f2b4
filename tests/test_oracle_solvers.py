"""Pins the oracle's solver compositions (FarField.hpp) against the sound-soft
sphere (Mie series), the total-field Neumann data, the radial-dilation
derivative and the adjoint dot-product identity; the conjugation convention of
the Hadamard product in AdjointDerivative_FF (HadamardProduct.hpp:21) is fixed
by the adjoint identity."""
import numpy as np
import pytest

import oracle
from conftest import oracle_mesh, relerr

DIRS = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8]])


@pytest.fixture(scope="module")
def mie_errors():
    out = {}
    for L in (2, 3):
        m = oracle_mesh("sphere", L, 64)
        kappa = [2.0]
        F, _ = oracle.far_field(m, kappa, DIRS, 2)
        ref = np.stack([oracle.mie_far_field_soft(2.0, m.meas @ d) for d in DIRS], axis=1)
        du = oracle.derivative_ff_helper(m, kappa, DIRS, 2)
        xhat = m.V / np.linalg.norm(m.V, axis=1, keepdims=True)
        refdu = np.stack([oracle.mie_neumann_soft(2.0, xhat @ d) for d in DIRS], axis=1)
        DFh, _ = oracle.derivative_ff(m, kappa, DIRS, 2, m.V.copy(), du)
        refD = np.stack([oracle.mie_dilation_derivative(2.0, m.meas @ d) for d in DIRS], axis=1)
        out[L] = (relerr(F, ref), relerr(du, refdu), relerr(DFh, refD))
    return out


def test_far_field_matches_mie(mie_errors):
    assert mie_errors[3][0] < 0.025 and mie_errors[3][0] < 0.6 * mie_errors[2][0]


def test_neumann_data_matches_closed_form(mie_errors):
    assert mie_errors[3][1] < 0.025 and mie_errors[3][1] < 0.6 * mie_errors[2][1]


def test_dilation_derivative_matches_closed_form(mie_errors):
    assert mie_errors[3][2] < 0.05 and mie_errors[3][2] < 0.6 * mie_errors[2][2]


def test_adjoint_identity_and_hadamard_convention():
    """(4 pi/(m W)) Re sum conj(g) DF h  =  h^T M DF* g  (-> 1 under refinement),
    with the unconjugated row sum sum_k du_dn * dv_dn in DF*."""
    ratios = {}
    for L, M in ((2, 128), (3, 256)):
        m = oracle_mesh("sphere", L, M)
        kappa, wcs = [2.0], 2
        rng = np.random.default_rng(0)
        dirs = rng.standard_normal((wcs, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        du = oracle.derivative_ff_helper(m, kappa, dirs, wcs)
        r = []
        for trial in range(4):
            h = np.stack([np.sin(1.3 * m.V[:, 1] + trial), np.cos(0.7 * m.V[:, 2]), m.V[:, 0] * (trial + 1)], axis=1)
            g = np.exp(1j * (trial + 1) * m.meas[:, [0]]) * np.array([[1.0, 0.5j]])
            a = oracle.derivative_ff(m, kappa, dirs, wcs, h, du)[0]
            Y = oracle.adjoint_derivative_ff(m, kappa, dirs, wcs, g, du)[0]
            lhs = (4 * np.pi / (M * wcs)) * np.sum(np.conj(g) * a).real
            r.append(lhs / np.sum(h * (m.M @ Y)))
        ratios[L] = np.array(r)
    assert np.max(np.abs(ratios[3] - 1)) < 0.06
    assert np.mean(np.abs(ratios[3] - 1)) < np.mean(np.abs(ratios[2] - 1))


def test_dense_solve_is_a_solution():
    m = oracle_mesh("sphere", 2)
    rng = np.random.default_rng(1)
    rhs = rng.standard_normal((m.nv, 3)) + 1j * rng.standard_normal((m.nv, 3))
    X = oracle.solve_boundary(m, [3.0], [np.array([0.5, -3j, 1.0, 0.0])], rhs, 3)
    back = oracle.apply_boundary(m, 1.0, X, 0.0, np.zeros_like(X), [3.0], [0.5, -3j, 1.0, 0.0], 3, 3)
    assert np.linalg.norm(back - rhs) < 1e-10 * np.linalg.norm(rhs)


def _gn_metric(m, c=0.05, d=0.5):
    """A stand-in SPD metric on vec(X) (V x 3 row-major): c I + d (Mass (x) I_3)."""
    Mf = m.M.toarray()
    return c * np.eye(3 * m.nv) + d * np.kron(Mf, np.eye(3))


def test_gauss_newton_operator_reproduces_the_pinned_compositions():
    """Every column family of the dense Gauss-Newton operator equals
    metric x + Mass DF*(DF x) evaluated with the pinned derivative_ff and
    adjoint_derivative_ff (FarField.hpp:551-572) on random X; the solution of
    gauss_newton then satisfies the normal equations through those same
    functions, with alpha/beta applied as in FarField.hpp:581-599."""
    m = oracle_mesh("sphere", 2, 48)
    kappa, wcs = [2.0, 3.5], 2
    dirs = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8]])
    metric = _gn_metric(m)
    A, du = oracle.gauss_newton_operator(m, kappa, dirs, wcs, metric)
    rng = np.random.default_rng(3)
    for trial in range(3):
        X = rng.standard_normal((m.nv, 3))
        DFX, _ = oracle.derivative_ff(m, kappa, dirs, wcs, X, du)
        Z, _ = oracle.adjoint_derivative_ff(m, kappa, dirs, wcs, DFX, du)
        ref = (metric @ X.reshape(-1)).reshape(m.nv, 3) + m.M @ Z
        assert relerr((A @ X.reshape(-1)).reshape(m.nv, 3), ref) < 1e-10
    B = rng.standard_normal((m.nv, 3))
    X0 = rng.standard_normal((m.nv, 3))
    X, _ = oracle.gauss_newton(m, kappa, dirs, wcs, metric, B, alpha=0.5, beta=2.0, X=X0, du_dn=du)
    x = (X - 2.0 * X0) / 0.5
    DFx, _ = oracle.derivative_ff(m, kappa, dirs, wcs, x, du)
    Zx, _ = oracle.adjoint_derivative_ff(m, kappa, dirs, wcs, DFx, du)
    res = (metric @ x.reshape(-1)).reshape(m.nv, 3) + m.M @ Zx - B
    assert np.linalg.norm(res) < 1e-9 * np.linalg.norm(B)


def test_gauss_newton_data_term_is_symmetric_positive():
    """Mass DF* DF is (up to discretisation) the Gauss-Newton Hessian of the
    misfit: x^T Mass DF* DF x = (4 pi / (m W)) |DF x|^2 >= 0 in the limit, so
    the data term of the operator is nearly symmetric and its quadratic form
    positive; a wrong sign or a dropped conjugation in DF* breaks both."""
    m = oracle_mesh("sphere", 3, 128)
    kappa, wcs = [2.0], 2
    dirs = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8]])
    A, du = oracle.gauss_newton_operator(m, kappa, dirs, wcs, np.zeros((3 * m.nv, 3 * m.nv)))
    asym = np.linalg.norm(A - A.T) / np.linalg.norm(A)
    assert asym < 0.05
    rng = np.random.default_rng(4)
    for _ in range(4):
        x = rng.standard_normal(3 * m.nv)
        DFx, _ = oracle.derivative_ff(m, kappa, dirs, wcs, x.reshape(m.nv, 3), du)
        quad = x @ A @ x
        data = 4 * np.pi / (128 * wcs) * np.sum(np.abs(DFx) ** 2)
        assert quad > 0 and abs(quad / data - 1) < 0.05
