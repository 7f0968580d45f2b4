"""The rest of the reference's class surface in the Python mirror: the P0
buffers (InputOutput.hpp:5-89), PrintBooleans (LoadParameters.hpp:85-95),
VertexCoordinates / TriangleCoordinates / DeviceInfo
(Helmholtz_OpenCL.hpp:333-391) and the wave-count getters and setters
(GetSetters.hpp:40-62), checked against the oracle's Av and the input mesh."""
import numpy as np
import pytest

from baemm_amd import Helmholtz_HIP
from conftest import oracle_mesh, relerr

pytestmark = pytest.mark.gpu


def test_mirror_buffers_flags_and_coordinates():
    m = oracle_mesh("sphere", 2)
    rng = np.random.default_rng(41)
    W, wcs = 8, 4
    kappa = [2.0, 3.0]
    coeff = np.array([0.5, -2j, 1.0, 0.0, 0.5, 1.5, 1.0, 0.0], np.complex128)
    B = rng.standard_normal((m.nv, W)) + 1j * rng.standard_normal((m.nv, W))
    with Helmholtz_HIP(m.V, m.T) as H:
        assert H.ClassName() == "Helmholtz_HIP"
        assert "DEVICE_NAME" in H.DeviceInfo()
        assert np.array_equal(H.VertexCoordinates(), m.V.astype(np.float32))
        tc = H.TriangleCoordinates().reshape(-1, 3, 4)
        assert np.array_equal(tc[:, :, :3], m.V[m.T].astype(np.float32)) and not tc[:, :, 3].any()
        with pytest.raises(Exception):
            H.WriteB()                       # no apply yet: the buffer has no value
        C = np.zeros_like(B)
        H.ApplyBoundaryOperators_PL(1.0, B, 0.0, C, kappa, coeff, W, wcs)
        flags = H.PrintBooleans()
        assert flags == {"Re_mass_matrix": True, "Im_mass_matrix": False, "Re_single_layer": True,
                         "Im_single_layer": True, "Re_double_layer": True, "Im_double_layer": False,
                         "Re_adjdbl_layer": False, "Im_adjdbl_layer": False}
        # B = Av B_in of the last apply
        Bp = H.WriteB()
        ref = m.Av.toarray() @ B
        assert Bp.shape == (m.nt, W) and relerr(Bp, ref) < 1e-6
        assert H.B(5, 3) == Bp[5, 3] and H.B(m.nt - 1, W - 1) == Bp[-1, -1]
        assert H.WriteC().shape == (m.nt, W)
        with pytest.raises(IndexError):
            H.B(m.nt, 0)
        x = (rng.standard_normal((m.nt, W)) + 1j * rng.standard_normal((m.nt, W))).astype(np.complex64)
        H.ReadB(x)
        assert np.array_equal(H.WriteB(), x) and H.B(7, 2) == x[7, 2]
        # wave counts
        assert H.GetWaveCount() == W and H.GetWaveChunkCount(10) == -(-10 // H.GetWaveChunkSize())
        H.SetWaveCount(5)
        assert H.GetWaveCount() == 5
        H.LoadBoundaryOperators_PL(kappa, coeff, W, wcs)
        assert H.GetWaveCount() == W
