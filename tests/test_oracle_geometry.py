"""Pins the oracle's geometry and sparse operators (Initialize.hpp:57-353)
against closed forms and an independent quadrature -- not against a retyped
formula."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import oracle
from baemm_amd.meshes import icosphere, read_mesh_file, signed_volume, torus
from conftest import REFERENCE, oracle_mesh, relerr


def _single(tri):
    V = np.asarray(tri, dtype=np.float64)
    return oracle.Mesh(V, np.array([[0, 1, 2]]))


def test_right_triangle_closed_forms():
    m = _single([[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    assert m.area[0] == pytest.approx(0.5, rel=1e-15)
    np.testing.assert_allclose(m.nrm[0], [0, 0, 1], atol=1e-15)
    np.testing.assert_allclose(m.mid[0], [1 / 3, 1 / 3, 0], atol=1e-15)


def test_equilateral_self_term_closed_form():
    # int_T 1/|c-y| dy / a = 4 ln(2 + sqrt 3) / s for the equilateral triangle of side s.
    for s in (0.01, 1.0, 3.5):
        m = _single([[0, 0, 0], [s, 0, 0], [s / 2, s * math.sqrt(3) / 2, 0]])
        assert m.sdiag[0] == pytest.approx(4 * math.log(2 + math.sqrt(3)) / s, rel=1e-13)


def _polar_self_term(P):
    """(1/a) int_T 1/|c-y| dy via polar coordinates around the centroid:
    int_0^{2 pi} R(theta) d theta, R the distance to the boundary."""
    P = np.asarray(P, dtype=np.float64)
    c = P.mean(axis=0)
    e1 = P[1] - P[0]
    e1 /= np.linalg.norm(e1)
    nrm = np.cross(P[1] - P[0], P[2] - P[0])
    area = 0.5 * np.linalg.norm(nrm)
    nrm /= np.linalg.norm(nrm)
    e2 = np.cross(nrm, e1)
    Q = np.stack([(P - c) @ e1, (P - c) @ e2], axis=1)
    ang = np.arctan2(Q[:, 1], Q[:, 0])
    total = 0.0
    for k in range(3):
        a, b = Q[k], Q[(k + 1) % 3]

        def R(th, a=a, b=b):
            d = np.array([math.cos(th), math.sin(th)])
            M = np.array([d, a - b]).T
            t, _ = np.linalg.solve(M, a)
            return t

        t0, t1 = ang[k], ang[(k + 1) % 3]
        if t1 < t0:
            t1 += 2 * math.pi
        total += quad(R, t0, t1, epsabs=0, epsrel=1e-13, limit=200)[0]
    return total / area


def test_self_term_vs_polar_quadrature():
    rng = np.random.default_rng(11)
    for _ in range(8):
        P = rng.standard_normal((3, 3))
        if np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0])) < 0.3:
            continue
        m = _single(P)
        # Initialize.hpp orders the triangle counter-clockwise around the normal; either
        # orientation gives the same integral.
        assert m.sdiag[0] == pytest.approx(_polar_self_term(P), rel=1e-9)


def test_f32_geometry_matches_f64():
    m = oracle_mesh("perturbed", 3)
    mid4, nrm4, area, sdiag = m.geometry_f32()
    assert relerr(mid4[:, :3], m.mid) < 1e-6
    assert relerr(nrm4[:, :3], m.nrm) < 1e-6
    assert relerr(area, m.area) < 1e-6
    assert relerr(sdiag, m.sdiag) < 1e-5


def test_mass_and_averaging_invariants():
    m = oracle_mesh("perturbed", 3)
    one_v = np.ones(m.nv)
    assert one_v @ (m.M @ one_v) == pytest.approx(m.area.sum(), rel=1e-13)
    np.testing.assert_allclose(m.M @ one_v, m.AvT @ np.ones(m.nt), rtol=1e-13)
    np.testing.assert_allclose(m.Av @ one_v, m.area, rtol=1e-13)
    # M is symmetric positive definite.
    assert abs(m.M - m.M.T).max() == 0.0
    w = np.linalg.eigvalsh(m.M.toarray())
    assert w.min() > 0


def test_mass_integrates_quadratics_on_sphere():
    # int_{S^2} x^2 dS = 4 pi / 3; the P1 mass form converges as O(h^2).
    errs = []
    for L in (2, 3, 4):
        m = oracle_mesh("sphere", L)
        x = m.V[:, 0]
        errs.append(abs(x @ (m.M @ x) - 4 * math.pi / 3))
    assert errs[2] < 0.3 * errs[1] < 0.09 * errs[0]
    assert errs[2] < 1.5e-2


def test_generated_meshes():
    for L, nt, nv in ((3, 1280, 642), (4, 5120, 2562), (6, 81920, 40962)):
        V, T = icosphere(L)
        assert T.shape == (nt, 3) and V.shape == (nv, 3)
        np.testing.assert_allclose(np.linalg.norm(V, axis=1), 1.0, atol=1e-14)
        assert signed_volume(V, T) > 0
        # closed: every edge shared by exactly two triangles
        E = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(E[:, 0] * nv + E[:, 1], return_counts=True)
        assert np.all(cnt == 2)
    V, T = torus(1.0, 0.4, 240, 80)
    assert T.shape == (38400, 3) and V.shape == (19200, 3)
    assert signed_volume(V, T) == pytest.approx(2 * math.pi ** 2 * 1.0 * 0.4 ** 2, rel=2e-3)


@pytest.mark.skipif(not (REFERENCE / "Meshes").exists(), reason="reference meshes not mounted")
def test_reference_mesh_files_match_generated_sizes():
    for name, L in (("Sphere_00001280T.txt", 3), ("Sphere_00005120T.txt", 4)):
        Vr, Tr = read_mesh_file(REFERENCE / "Meshes" / name)
        V, T = icosphere(L)
        assert Vr.shape == V.shape and Tr.shape == T.shape
        np.testing.assert_allclose(np.linalg.norm(Vr, axis=1), 1.0, atol=1e-6)
        mr = oracle.Mesh(Vr, Tr)
        mg = oracle.Mesh(V, T)
        assert mr.area.sum() == pytest.approx(mg.area.sum(), rel=1e-4)
        assert signed_volume(Vr, Tr) > 0


@pytest.mark.parametrize("binary", [True, False])
def test_stl_round_trip(tmp_path, binary):
    """read_stl (ReadMeshFromSTL.hpp input path) recovers the mesh written as
    STL: same triangles up to float32 rounding, corners merged into vertices."""
    from baemm_amd.meshes import icosphere, read_stl, write_stl

    V, T = icosphere(2)
    f = tmp_path / ("m.stl" if binary else "m_ascii.stl")
    write_stl(f, V, T, binary=binary)
    V2, T2 = read_stl(f)
    assert V2.shape == V.shape and T2.shape == T.shape
    assert np.allclose(V2[T2], V[T], atol=1e-6)
    # first-appearance numbering: the first triangle's corners are 0, 1, 2
    assert list(T2[0]) == [0, 1, 2]
    with pytest.raises(ValueError):
        bad = tmp_path / "bad.stl"
        bad.write_text("not a mesh")
        read_stl(bad)
