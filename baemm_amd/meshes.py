"""Mesh generators and the plain-text mesh format used by the BAEMM examples.

The text format is the one read by ``ReadMeshFromFile.hpp:30-97``::

    dom_dim 2
    amb_dim 3
    vertex_count V
    simplex_count T
    <V lines of 3 coordinates>
    <T lines of 3 zero-based vertex indices>

Generators (all closed, outward-oriented triangle meshes):

* :func:`icosphere` -- midpoint subdivision of the unit icosahedron,
  projected to the unit sphere; level L has 20*4**L triangles and
  10*4**L + 2 vertices (level 4: 5120/2562, level 6: 81920/40962, matching
  ``Meshes/Sphere_00005120T.txt`` and ``Meshes/Sphere_00081920T.txt`` in size).
* :func:`torus` -- (u, v) grid split into two triangles per quad.
* :func:`perturbed_sphere` -- icosphere with radius 1 + eps * sum of real
  spherical harmonics of degree <= lmax with seeded N(0,1)/(l+1)^2 weights.
* :func:`fibonacci_sphere` -- m quasi-uniform unit directions.

STL (ASCII or binary) is read by :func:`read_stl`, the input path of
``ReadMeshFromSTL.hpp:5-44``: coordinates in single precision, coincident
corners merged into one vertex.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

__all__ = [
    "icosphere",
    "torus",
    "perturbed_sphere",
    "fibonacci_sphere",
    "random_directions",
    "read_mesh_file",
    "write_mesh_file",
    "read_stl",
    "write_stl",
    "signed_volume",
    "real_sph_harm",
]


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = np.array(
        [
            [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
            [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
            [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1],
        ],
        dtype=np.float64,
    )
    F = np.array(
        [
            [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
            [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
            [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
            [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
        ],
        dtype=np.int64,
    )
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    return V, F


def icosphere(level: int):
    """Unit icosphere: ``level``-fold midpoint subdivision of the icosahedron.

    Returns ``(V, T)`` with V float64 (nv, 3) and T int64 (nt, 3), outward
    orientation.
    """
    V, F = _icosahedron()
    for _ in range(level):
        nv = V.shape[0]
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
        e_sorted = np.sort(e, axis=1)
        key = e_sorted[:, 0] * (nv + 1) + e_sorted[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // (nv + 1)
        b = uniq % (nv + 1)
        mids = 0.5 * (V[a] + V[b])
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        nf = F.shape[0]
        m01 = nv + inv[:nf]
        m12 = nv + inv[nf:2 * nf]
        m20 = nv + inv[2 * nf:]
        F = np.concatenate(
            [
                np.stack([F[:, 0], m01, m20], axis=1),
                np.stack([F[:, 1], m12, m01], axis=1),
                np.stack([F[:, 2], m20, m12], axis=1),
                np.stack([m01, m12, m20], axis=1),
            ],
            axis=0,
        )
        V = np.concatenate([V, mids], axis=0)
    return np.ascontiguousarray(V), np.ascontiguousarray(F.astype(np.int64))


def torus(R: float = 1.0, r: float = 0.4, nu: int = 240, nv: int = 80):
    """Torus with major radius R, minor radius r on an nu x nv (u, v) grid.

    x = (R + r cos v) cos u, y = (R + r cos v) sin u, z = r sin v.  Each quad
    is split into two triangles; 2*nu*nv triangles, nu*nv vertices, outward.
    """
    u = 2.0 * np.pi * np.arange(nu) / nu
    v = 2.0 * np.pi * np.arange(nv) / nv
    U, Vv = np.meshgrid(u, v, indexing="ij")
    X = np.stack(
        [(R + r * np.cos(Vv)) * np.cos(U), (R + r * np.cos(Vv)) * np.sin(U), r * np.sin(Vv)],
        axis=-1,
    ).reshape(-1, 3)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    p00 = (i * nv + j).ravel()
    p10 = (((i + 1) % nu) * nv + j).ravel()
    p01 = (i * nv + (j + 1) % nv).ravel()
    p11 = (((i + 1) % nu) * nv + (j + 1) % nv).ravel()
    T = np.concatenate([np.stack([p00, p10, p11], 1), np.stack([p00, p11, p01], 1)], axis=0)
    if signed_volume(X, T) < 0:
        T = T[:, [0, 2, 1]]
    return np.ascontiguousarray(X), np.ascontiguousarray(T.astype(np.int64))


def _assoc_legendre(lmax, x):
    """Normalised associated Legendre functions for real spherical harmonics."""
    P = {}
    s = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    P[(0, 0)] = np.full_like(x, 1.0)
    for m in range(1, lmax + 1):
        P[(m, m)] = -(2 * m - 1) * s * P[(m - 1, m - 1)]
    for m in range(0, lmax):
        P[(m + 1, m)] = (2 * m + 1) * x * P[(m, m)]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            P[(l, m)] = ((2 * l - 1) * x * P[(l - 1, m)] - (l + m - 1) * P[(l - 2, m)]) / (l - m)
    return P


def real_sph_harm(lmax: int, X: np.ndarray):
    """Orthonormal real spherical harmonics Y_lm, l <= lmax, at unit vectors X.

    Returns a list of (l, m, values) with m in -l..l.
    """
    Xn = X / np.linalg.norm(X, axis=1, keepdims=True)
    ct = Xn[:, 2]
    phi = np.arctan2(Xn[:, 1], Xn[:, 0])
    P = _assoc_legendre(lmax, ct)
    out = []
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
            base = norm * P[(l, am)] * (-1) ** am  # remove Condon-Shortley phase
            if m > 0:
                val = math.sqrt(2.0) * base * np.cos(am * phi)
            elif m < 0:
                val = math.sqrt(2.0) * base * np.sin(am * phi)
            else:
                val = base
            out.append((l, m, val))
    return out


def perturbed_sphere(level: int, seed: int, eps: float = 0.1, lmax: int = 4):
    """Icosphere of the given level with radius 1 + eps * sum a_lm Y_lm.

    a_lm ~ N(0, 1) / (l + 1)**2 drawn from ``numpy.random.default_rng(seed)``.
    """
    V, T = icosphere(level)
    rng = np.random.default_rng(seed)
    rad = np.ones(V.shape[0])
    for l, m, y in real_sph_harm(lmax, V):
        rad += eps * rng.standard_normal() / (l + 1) ** 2 * y
    return np.ascontiguousarray(V * rad[:, None]), T


def fibonacci_sphere(m: int):
    """m Fibonacci directions: z_k = 1 - (2k+1)/m, phi_k = k * pi * (3 - sqrt 5)."""
    k = np.arange(m, dtype=np.float64)
    z = 1.0 - (2.0 * k + 1.0) / m
    phi = k * np.pi * (3.0 - math.sqrt(5.0))
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.ascontiguousarray(np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1))


def random_directions(n: int, seed: int):
    """n unit vectors, normalised N(0, I_3) draws from default_rng(seed)."""
    d = np.random.default_rng(seed).standard_normal((n, 3))
    return np.ascontiguousarray(d / np.linalg.norm(d, axis=1, keepdims=True))


def volume_grid(shape, lo, hi):
    """Points of a regular box grid: shape = (nx, ny, nz) samples from lo to hi
    (inclusive) per axis, x slowest; (nx ny nz) x 3.  Evaluation points for the
    potential (ApplyNearFieldOperators_PL) on a volume, for visualisation."""
    axes = [np.linspace(l, h, n) for n, l, h in zip(shape, lo, hi)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def shell_points(V: np.ndarray, T: np.ndarray, offsets, per_offset: int, seed: int):
    """Near-surface points: for each offset d, per_offset triangles drawn at
    random (seeded) and the point c_t + d n_t at signed distance d from the
    centroid c_t along the unit outward normal n_t (d < 0: inside)."""
    rng = np.random.default_rng(seed)
    A = V[T[:, 0]]
    c = (A + V[T[:, 1]] + V[T[:, 2]]) / 3.0
    n = np.cross(V[T[:, 1]] - A, V[T[:, 2]] - A)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    out = []
    for d in offsets:
        t = rng.choice(T.shape[0], per_offset, replace=False)
        out.append(c[t] + d * n[t])
    return np.concatenate(out, axis=0)


def signed_volume(V: np.ndarray, T: np.ndarray) -> float:
    """Signed enclosed volume (positive for outward orientation)."""
    a, b, c = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def read_mesh_file(path):
    """Read the text mesh format of ``ReadMeshFromFile.hpp:30-97``.

    Raises ValueError on a malformed header (the reference prints
    "Invalid file. Aborting." and returns false).
    """
    tok = Path(path).read_text().split()
    hdr = {}
    pos = 0
    for key in ("dom_dim", "amb_dim", "vertex_count", "simplex_count"):
        if pos >= len(tok) or tok[pos] != key:
            raise ValueError(f"{path}: invalid mesh file (expected '{key}')")
        hdr[key] = int(tok[pos + 1])
        pos += 2
    nv, nt, ad, dd = hdr["vertex_count"], hdr["simplex_count"], hdr["amb_dim"], hdr["dom_dim"]
    vals = np.array(tok[pos:pos + nv * ad], dtype=np.float64).reshape(nv, ad)
    pos += nv * ad
    simp = np.array(tok[pos:pos + nt * (dd + 1)], dtype=np.int64).reshape(nt, dd + 1)
    return vals, simp


def write_mesh_file(path, V: np.ndarray, T: np.ndarray):
    """Write the text mesh format read by :func:`read_mesh_file`."""
    lines = [
        "dom_dim 2",
        f"amb_dim {V.shape[1]}",
        f"vertex_count {V.shape[0]}",
        f"simplex_count {T.shape[0]}",
    ]
    lines += [" ".join(repr(float(x)) for x in row) for row in V]
    lines += [" ".join(str(int(x)) for x in row) for row in T]
    Path(path).write_text("\n".join(lines) + "\n")


def read_stl(path):
    """Read an ASCII or binary STL file (``ReadMeshFromSTL.hpp:5-44``).

    Corners with identical single-precision coordinates become one vertex
    (numbered in order of first appearance); returns ``(V, T)`` with V float64
    (values of the file's float32 coordinates) and T int64.  Raises ValueError
    for a file that is neither form.
    """
    data = Path(path).read_bytes()
    corners = None
    if len(data) >= 84:
        n = int.from_bytes(data[80:84], "little")
        if len(data) == 84 + 50 * n:
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                                count=n, offset=84)
            corners = np.ascontiguousarray(rec["v"].reshape(-1, 3))
    if corners is None:
        text = data.decode("ascii", errors="replace")
        if not text.lstrip().lower().startswith("solid"):
            raise ValueError(f"{path}: not an STL file")
        vals = []
        for line in text.splitlines():
            tok = line.split()
            if tok and tok[0] == "vertex":
                if len(tok) != 4:
                    raise ValueError(f"{path}: malformed vertex line {line!r}")
                vals.append([float(t) for t in tok[1:4]])
        if len(vals) % 3:
            raise ValueError(f"{path}: vertex count is not a multiple of 3")
        corners = np.array(vals, dtype=np.float32).reshape(-1, 3)
    uniq, first, inv = np.unique(corners, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)                 # number vertices by first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    V = uniq[order].astype(np.float64)
    T = rank[inv.reshape(-1)].reshape(-1, 3).astype(np.int64)
    return np.ascontiguousarray(V), np.ascontiguousarray(T)


def write_stl(path, V: np.ndarray, T: np.ndarray, binary: bool = True):
    """Write a triangle mesh as STL (facet normals from the corner order)."""
    P = np.asarray(V, np.float64)[np.asarray(T)]
    nrm = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-300)
    if binary:
        rec = np.zeros(P.shape[0], dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
        rec["n"] = nrm
        rec["v"] = P
        Path(path).write_bytes(b"baemm".ljust(80, b" ") + np.uint32(P.shape[0]).tobytes() + rec.tobytes())
    else:
        out = ["solid baemm"]
        for n, tri in zip(nrm, P):
            out.append(f"  facet normal {n[0]:.9e} {n[1]:.9e} {n[2]:.9e}")
            out.append("    outer loop")
            out += [f"      vertex {x:.9e} {y:.9e} {z:.9e}" for x, y, z in tri.astype(np.float32)]
            out.append("    endloop")
            out.append("  endfacet")
        out.append("endsolid baemm")
        Path(path).write_text("\n".join(out) + "\n")
