"""GPU parity of the boundary-operator apply on the reference's own graded
meshes (Meshes/Spot_00005856T, Bob_00010688T, Blub_00014208T and
Bunny_00086632T, committed as fixtures): largest/smallest triangle area from
57 (Bob) to 24,049 (Bunny).  Three wave numbers kappa = 1, 8, 16 (one chunk of
8 columns each), in every precision, against the FP64 oracle on 256 sampled
vertex rows:
  * single layer only (c2 = 0: the general form of the MFMA kernels),
  * the potential operator (1/2, -i kappa, 1, 0) (the form divided by c2),
  * the Dirichlet-to-Neumann operator (1/2, -i kappa, 0, 1) (the form divided
    by c3 on the target normals),
  * all four operators with complex coefficients.
The contract is 1e-5 relative (Frobenius); each precision is also held near
the error it achieves, so that a tenfold loss of accuracy fails."""
import functools

import numpy as np
import pytest

import oracle
from baemm_amd import Helmholtz_HIP
from conftest import reference_mesh, relerr

pytestmark = pytest.mark.gpu

CONTRACT = 1e-5
KAPPA = [1.0, 8.0, 16.0]
W, WCS = 24, 8
NAMES = ["Spot_00005856T", "Bob_00010688T", "Blub_00014208T", "Bunny_00086632T"]
COEFFS = {
    "SL": np.array([0, 1, 0, 0] * 3, np.complex128),
    "SL+DL": np.array(sum(([0.5, -1j * k, 1, 0] for k in KAPPA), []), np.complex128),
    "DtN": np.array(sum(([0.5, -1j * k, 0, 1] for k in KAPPA), []), np.complex128),
    "all": np.array([0.5 + 0.25j, 1 - 0.5j, -0.75 + 1j, 0.3 - 0.9j] * 3, np.complex128),
}
# achieved-error guards per precision (about 5x the largest error measured on these meshes)
BOUND = {"fp32": 5e-6, "fp32_mfma": 5e-6, "mixed": 5e-6}


@functools.lru_cache(maxsize=None)
def _case(name):
    V, T = reference_mesh(name)
    m = oracle.Mesh(V, T)
    rng = np.random.default_rng(len(name))
    B = rng.standard_normal((m.nv, W)) + 1j * rng.standard_normal((m.nv, W))
    rows = np.sort(rng.choice(m.nv, 256, replace=False))
    refs = {k: oracle.apply_boundary_rows(m, 1.0, B, 0.0, np.zeros((256, W)), KAPPA, c, W, WCS, rows)
            for k, c in COEFFS.items()}
    return m, B, rows, refs


@pytest.mark.parametrize("name", NAMES)
def test_graded_reference_mesh_sampled_rows(name):
    m, B, rows, refs = _case(name)
    worst = {}
    with Helmholtz_HIP(m.V, m.T) as H:
        for mode in ("fp32", "mixed", "fp32_mfma"):
            H.SetPrecision(mode)
            for key, coeff in COEFFS.items():
                C = np.zeros_like(B)
                H.ApplyBoundaryOperators_PL(1.0, B, 0.0, C, KAPPA, coeff, W, WCS)
                e = relerr(C[rows], refs[key])
                # per wave number (chunk) as well: low kappa is where the mixed
                # path's A scale spans the widest range
                eq = [relerr(C[rows, q * WCS:(q + 1) * WCS], refs[key][:, q * WCS:(q + 1) * WCS]) for q in range(3)]
                print(f"{name} {mode:9s} {key:6s} rel. err {e:.2e}  per kappa "
                      + " ".join(f"{x:.2e}" for x in eq), flush=True)
                worst[(mode, key)] = max([e] + eq)
    for (mode, key), e in worst.items():
        assert e < CONTRACT, (mode, key, e)
        assert e < BOUND[mode], (mode, key, e)
