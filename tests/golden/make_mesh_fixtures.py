"""Writes tests/golden/meshes/<name>.npz from the reference's mesh files.

The fixtures are data: the vertex coordinates (float64, exactly the decimal
values of the file as parsed) and the zero-based triangles (int32) of the
graded meshes shipped in the reference's ``Meshes/`` directory, read with the
text-format reader of this package (``ReadMeshFromFile.hpp:30-97``).

Run in a container that has the reference checked out:
    python tests/golden/make_mesh_fixtures.py /root/reference/Meshes
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from baemm_amd.meshes import read_mesh_file  # noqa: E402

# Spot/Bob/Blub: graded character meshes (area ratio 57-162); Bunny: the most
# strongly graded mesh of the reference (largest/smallest triangle area 24,049).
NAMES = ["Spot_00005856T", "Bob_00010688T", "Blub_00014208T", "Bunny_00086632T"]


def main(src):
    out = HERE / "meshes"
    out.mkdir(exist_ok=True)
    for name in NAMES:
        V, T = read_mesh_file(Path(src) / f"{name}.txt")
        np.savez_compressed(out / f"{name}.npz", V=V.astype(np.float64), T=T.astype(np.int32))
        print(name, V.shape, T.shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/Meshes")
