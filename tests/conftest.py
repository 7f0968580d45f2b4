"""Shared pytest configuration.

Markers: ``gpu`` -- needs an MI355X and the built HIP library.
CPU tests (``-m "not gpu"``) cover the oracle, host logic and the C-ABI's
exported symbols.
"""
import functools
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

REFERENCE = Path("/root/reference")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs an MI355X GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


@functools.lru_cache(maxsize=None)
def icosphere_mesh(level):
    from baemm_amd.meshes import icosphere

    return icosphere(level)


@functools.lru_cache(maxsize=None)
def oracle_mesh(kind, level, meas=0):
    import oracle
    from baemm_amd.meshes import fibonacci_sphere, icosphere, perturbed_sphere, torus

    if kind == "sphere":
        V, T = icosphere(level)
    elif kind == "perturbed":
        V, T = perturbed_sphere(level, seed=3)
    elif kind == "torus":
        V, T = torus(1.0, 0.4, 6 * level, 2 * level)
    else:
        raise ValueError(kind)
    return oracle.Mesh(V, T, fibonacci_sphere(meas) if meas else None)


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


GOLDEN = Path(__file__).resolve().parent / "golden"


@functools.lru_cache(maxsize=None)
def reference_mesh(name):
    """A mesh of the reference's Meshes/ directory, from its committed fixture
    (tests/golden/meshes/<name>.npz, written by tests/golden/make_mesh_fixtures.py)."""
    with np.load(GOLDEN / "meshes" / f"{name}.npz") as z:
        return z["V"].astype(np.float64), z["T"].astype(np.int64)
