"""baemm_apply_boundary_dev_noalloc (no reference counterpart: the
device-pointer boundary apply for callers that cannot allocate) against the
existing entry point baemm_apply_boundary_dev, through the C ABI: two calls in
a row in every precision, the second on a smaller mesh with fewer columns, and
an apply that does not fit the fixed scratch array, which is refused before
anything runs."""
import ctypes

import numpy as np
import pytest

import oracle
from baemm_amd import Helmholtz_HIP, _lib
from baemm_amd.meshes import icosphere
from conftest import relerr

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p


def _dev(L, host):
    ptr = P()
    assert L.baemm_dev_alloc(ctypes.byref(ptr), max(16, host.nbytes)) == 0
    assert L.baemm_dev_upload(ptr, host.ctypes.data_as(P), host.nbytes) == 0
    return ptr


def _apply_both(H, B, kappa, coeff, W, wcs, noalloc_first):
    """C = A B through baemm_apply_boundary_dev and through the no-allocation
    entry point, from the same device input; returns (existing, no-alloc).
    With noalloc_first the no-allocation call is the handle's first apply (it
    then also finds the minimum centroid distance in its scratch array)."""
    L = _lib.load()
    H.LoadBoundaryOperators_PL(kappa, coeff, W, wcs)
    dB = _dev(L, B)
    outs = []
    try:
        fns = [L.baemm_apply_boundary_dev, L.baemm_apply_boundary_dev_noalloc]
        for fn in (fns[::-1] if noalloc_first else fns):
            C = np.full_like(B, 3.0 + 1.0j)
            dC = _dev(L, C)
            try:
                one = (ctypes.c_double * 2)(1.0, 0.0)
                zero = (ctypes.c_double * 2)(0.0, 0.0)
                st = fn(H._h, one, dB, W, _lib.C128, zero, dC, W, _lib.C128, None)
                assert st == 0, (st, L.baemm_last_error(H._h))
                assert L.baemm_dev_synchronize() == 0
                assert L.baemm_dev_download(C.ctypes.data_as(P), dC, C.nbytes) == 0
            finally:
                L.baemm_dev_free(dC)
            outs.append(C)
    finally:
        L.baemm_dev_free(dB)
    return outs[::-1] if noalloc_first else outs


def _rows(form, kappa):
    """Coefficient rows per chunk: all four operators, single + double layer
    (c1 with a zero real or imaginary part), or the Dirichlet-to-Neumann rows
    (adjoint double layer without double layer)."""
    if form == "full":
        return np.array([0.5 + 0.25j, 1 - 0.5j, -0.75 + 1j, 0.3 - 0.9j] * len(kappa), np.complex128)
    if form == "sl_dl":
        return np.array([c for k in kappa for c in (0.5, -1j * k, 1.0, 0.0)], np.complex128)
    return np.array([c for k in kappa for c in (0.5, -1j * k, 0.0, 1.0)], np.complex128)


@pytest.mark.parametrize("form", ["full", "sl_dl", "dtn"])
@pytest.mark.parametrize("mode", ["fp32", "mixed", "fp32_mfma"])
def test_noalloc_apply_matches_allocating_apply(mode, form):
    rng = np.random.default_rng(31)
    cases = [(icosphere(4), [4.0, 9.0], 32, 16), (icosphere(3), [2.5], 8, 8)]   # second call: smaller input
    for first, ((V, T), kappa, W, wcs) in enumerate(cases):
        coeff = _rows(form, kappa)
        B = rng.standard_normal((V.shape[0], W)) + 1j * rng.standard_normal((V.shape[0], W))
        with Helmholtz_HIP(V, T) as H:
            H.SetPrecision(mode)
            ref_dev, got = _apply_both(H, B, kappa, coeff, W, wcs, noalloc_first=first == 0)
        ref = oracle.apply_boundary(oracle.Mesh(V, T), 1.0, B, 0.0, np.zeros_like(B), kappa, coeff, W, wcs)
        print(f"{mode} {form} {T.shape[0]} triangles W={W}: no-alloc vs existing {relerr(got, ref_dev):.2e}, "
              f"vs FP64 oracle {relerr(got, ref):.2e}")
        # the same plan, kernels and segment split: identical results
        assert np.array_equal(got, ref_dev)
        assert relerr(got, ref) < 1e-5


def test_noalloc_apply_refuses_what_does_not_fit():
    L = _lib.load()
    cap = L.baemm_noalloc_scratch_bytes()
    assert cap == 256 << 20
    V, T = icosphere(6)   # 81,920 triangles at 64 columns: about 300 MB of workspace
    W = 64
    B = np.ones((V.shape[0], W), np.complex128)
    with Helmholtz_HIP(V, T) as H:
        H.SetPrecision("mixed")
        H.LoadBoundaryOperators_PL([8.0, 16.0], [0, 1, 1, 0] * 2, W, 32)
        dB = _dev(L, B)
        dC = _dev(L, B)
        try:
            one = (ctypes.c_double * 2)(1.0, 0.0)
            st = L.baemm_apply_boundary_dev_noalloc(H._h, one, dB, W, _lib.C128, None, dC, W, _lib.C128, None)
            msg = L.baemm_last_error(H._h).decode()
            print(st, msg)
            assert st == _lib.ERR_ARG and "workspace" in msg
            assert L.baemm_dev_synchronize() == 0
            # nothing ran: C still holds its input
            out = np.zeros_like(B)
            assert L.baemm_dev_download(out.ctypes.data_as(P), dC, out.nbytes) == 0
            assert np.array_equal(out, B)
        finally:
            L.baemm_dev_free(dB)
            L.baemm_dev_free(dC)
