"""Pins of the oracle's tangent-point energy, energy gradient and metric
(the Gauss-Newton example's caller side, Example_GaussNewton/main.cpp:301-346;
restated from the published discretisation, Repulsor is not vendored):
finite differences, invariances, scaling laws and structure."""
import numpy as np
import pytest

import oracle
from baemm_amd.meshes import perturbed_sphere


def _mesh(V, T):
    return oracle.Mesh(V, T)


@pytest.mark.parametrize("q,p", [(6.0, 12.0), (3.5, 8.25)])
def test_energy_gradient_matches_finite_differences(q, p):
    V, T = perturbed_sphere(1, seed=4)
    m = _mesh(V, T)
    G = oracle.tp_energy_gradient(m, q, p)
    rng = np.random.default_rng(1)
    h = 1e-6
    for v, c in zip(rng.choice(m.nv, 12, replace=False), rng.integers(0, 3, 12)):
        Vp, Vm = V.copy(), V.copy()
        Vp[v, c] += h
        Vm[v, c] -= h
        fd = (oracle.tp_energy(_mesh(Vp, T), q, p) - oracle.tp_energy(_mesh(Vm, T), q, p)) / (2 * h)
        assert abs(fd - G[v, c]) <= 1e-6 * np.abs(G).max() + 1e-6 * abs(fd), (v, c, fd, G[v, c])
    # a random direction as well
    dV = rng.standard_normal(V.shape)
    fd = (oracle.tp_energy(_mesh(V + 1e-6 * dV, T), q, p) - oracle.tp_energy(_mesh(V - 1e-6 * dV, T), q, p)) / 2e-6
    assert abs(fd - (G * dV).sum()) < 1e-6 * abs(fd)


def test_energy_invariances_and_scaling():
    V, T = perturbed_sphere(1, seed=5)
    q, p = 6.0, 12.0
    E = oracle.tp_energy(_mesh(V, T), q, p)
    th = 0.7
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    assert abs(oracle.tp_energy(_mesh(V @ Rz.T + [0.3, -1.0, 2.0], T), q, p) - E) < 1e-12 * E
    lam = 1.7
    assert abs(oracle.tp_energy(_mesh(lam * V, T), q, p) - lam ** (4 - p + q) * E) < 1e-12 * lam ** (4 - p + q) * E
    # translations do not change the energy: the gradient sums to zero
    G = oracle.tp_energy_gradient(_mesh(V, T), q, p)
    assert np.abs(G.sum(0)).max() < 1e-10 * np.abs(G).max()


def test_face_gradients_reproduce_the_tangential_projector():
    V, T = perturbed_sphere(1, seed=6)
    m = _mesh(V, T)
    g = oracle.face_gradients(m)
    assert np.abs(g.sum(1)).max() < 1e-12 * np.abs(g).max()
    X = V[T]
    Pm = np.einsum("tkc,tkj->tcj", X, g)
    proj = np.eye(3)[None] - np.einsum("tc,tj->tcj", m.nrm, m.nrm)
    assert np.abs(Pm - proj).max() < 1e-12


def test_metric_structure_and_scaling():
    V, T = perturbed_sphere(1, seed=7)
    q, p = 6.0, 12.0
    sigma = (p - 2) / q - 1
    M = oracle.tp_metric(_mesh(V, T), q, p)
    assert np.abs(M - M.T).max() < 1e-12 * np.abs(M).max()
    ev = np.linalg.eigvalsh(0.5 * (M + M.T))
    assert ev.min() > -1e-10 * ev.max()
    for c in range(3):
        t = np.zeros((V.shape[0], 3))
        t[:, c] = 1.0
        assert np.abs(M @ t.ravel()).max() < 1e-10 * np.abs(M).max()
    # a rotation field is not in the kernel
    rot = np.cross([0.0, 0.0, 1.0], V)
    assert rot.ravel() @ M @ rot.ravel() > 1e-6 * ev.max() * (rot ** 2).sum()
    lam = 1.3
    M2 = oracle.tp_metric(_mesh(lam * V, T), q, p)
    assert np.abs(M2 - lam ** (-2 * sigma) * M).max() < 1e-10 * np.abs(M2).max()
