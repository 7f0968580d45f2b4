"""The committed mesh fixtures (tests/golden/meshes) are the reference's own
graded meshes: identical to its Meshes/*.txt files where the reference is
present, and closed, consistently oriented surfaces in any case."""
import numpy as np
import pytest

from baemm_amd.meshes import read_mesh_file
from conftest import REFERENCE, reference_mesh

NAMES = ["Spot_00005856T", "Bob_00010688T", "Blub_00014208T", "Bunny_00086632T"]
GENUS = {"Spot_00005856T": 0, "Bob_00010688T": 1, "Blub_00014208T": 0, "Bunny_00086632T": 0}


@pytest.mark.parametrize("name", NAMES)
def test_fixture_is_a_closed_oriented_surface(name):
    V, T = reference_mesh(name)
    assert T.shape[0] == int(name.split("_")[1][:-1])
    assert T.min() == 0 and T.max() == V.shape[0] - 1
    # every edge is shared by exactly two triangles, traversed in opposite directions
    e = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    key = e[:, 0] * V.shape[0] + e[:, 1]
    rev = e[:, 1] * V.shape[0] + e[:, 0]
    assert np.unique(key).size == key.size
    assert np.array_equal(np.sort(key), np.sort(rev))
    # Euler characteristic 2 - 2 genus (Bob has one handle) and positive enclosed volume
    assert V.shape[0] - key.size // 2 + T.shape[0] == 2 - 2 * GENUS[name]
    vol = np.einsum("ij,ij->i", V[T[:, 0]], np.cross(V[T[:, 1]], V[T[:, 2]])).sum() / 6
    assert vol > 0


@pytest.mark.skipif(not (REFERENCE / "Meshes").exists(), reason="reference checkout not present")
@pytest.mark.parametrize("name", NAMES)
def test_fixture_equals_reference_file(name):
    V, T = reference_mesh(name)
    Vr, Tr = read_mesh_file(REFERENCE / "Meshes" / f"{name}.txt")
    assert np.array_equal(V, Vr) and np.array_equal(T, Tr)
