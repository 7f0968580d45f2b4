"""Batched persistent apply (baemm_apply_boundary_batch): uneven meshes,
wave numbers and RHS widths in one launch per precision, against the FP64
oracle and the single-instance apply; an instance's result is the same alone
and in the mix."""
import numpy as np
import pytest

import oracle
from baemm_amd import (ApplyBoundaryOperatorsBatch, BaemmError, BatchAutotune, BatchStats, BatchTuning,
                       Helmholtz_HIP)
from baemm_amd.meshes import perturbed_sphere
from conftest import relerr

pytestmark = pytest.mark.gpu

# (level, W, wcs, adjoint-double-layer on, precision)
SPECS = [(2, 1, 1, False, "mixed"), (3, 2, 2, True, "fp32"), (2, 4, 4, False, "fp32"), (3, 8, 8, True, "mixed"),
         (2, 16, 16, True, "mixed"), (3, 32, 32, False, "mixed"), (2, 64, 32, True, "mixed"),
         (3, 3, 3, False, "fp32"), (4, 64, 64, False, "mixed"), (3, 16, 8, True, "fp32")]


def _instance(k, level, W, wcs, adj, prec, rng):
    V, T = perturbed_sphere(level, seed=10 + k)
    K = W // wcs
    kappa = list(rng.uniform(1.0, 16.0, K))
    coeff = []
    for _ in range(K):
        row = rng.standard_normal(4) + 1j * rng.standard_normal(4)
        if not adj:
            row[3] = 0
        coeff.extend(row)
    B = rng.standard_normal((V.shape[0], W)) + 1j * rng.standard_normal((V.shape[0], W))
    return V, T, kappa, np.array(coeff), B


@pytest.fixture(scope="module")
def batch():
    rng = np.random.default_rng(3)
    items = []
    for k, (level, W, wcs, adj, prec) in enumerate(SPECS):
        V, T, kappa, coeff, B = _instance(k, level, W, wcs, adj, prec, rng)
        H = Helmholtz_HIP(V, T)
        H.SetPrecision(prec)
        H.LoadBoundaryOperators_PL(kappa, coeff, W, wcs)
        items.append(dict(H=H, V=V, T=T, kappa=kappa, coeff=coeff, B=B, W=W, wcs=wcs, prec=prec))
    yield items
    for it in items:
        it["H"].close()


def test_batch_matches_oracle_and_single(batch):
    Hs = [it["H"] for it in batch]
    out = ApplyBoundaryOperatorsBatch(Hs, [it["B"] for it in batch])
    st = BatchStats()
    print("batch stats", st)
    assert st["items_mixed"] > 0 and st["items_fp32"] > 0
    for k, (it, C) in enumerate(zip(batch, out)):
        m = oracle.Mesh(it["V"], it["T"])
        ref = oracle.apply_boundary(m, 1.0, it["B"], 0.0, np.zeros_like(it["B"]), it["kappa"], it["coeff"], it["W"],
                                    it["wcs"])
        err = relerr(C, ref)
        single = np.zeros_like(it["B"])
        it["H"].ApplyBoundaryOperators_PL(1.0, it["B"], 0.0, single)
        print(k, it["prec"], it["W"], "oracle", err, "single", relerr(C, single))
        assert err < 1e-5, k
        assert relerr(C, single) < 2e-6, k


def test_batch_result_independent_of_mix(batch):
    Hs = [it["H"] for it in batch]
    Bs = [it["B"] for it in batch]
    alpha = [0.5 + 0.25j] * len(batch)
    beta = [0.0] * len(batch)
    full = ApplyBoundaryOperatorsBatch(Hs, Bs, alpha=alpha)
    for k in (0, 3, 8, 9):
        alone = ApplyBoundaryOperatorsBatch([Hs[k]], [Bs[k]], alpha=[alpha[k]])
        assert np.array_equal(alone[0], full[k]), k
    # alpha / beta per instance
    C0 = [np.ones_like(b) for b in Bs]
    ab = ApplyBoundaryOperatorsBatch(Hs, Bs, C_list=C0, alpha=[1.0] * len(batch), beta=[2.0] * len(batch))
    for k in range(len(batch)):
        assert relerr(ab[k], full[k] / alpha[k] + 2.0) < 1e-12


def test_batch_rejects_duplicate_handles(batch):
    H = batch[0]["H"]
    with pytest.raises(BaemmError):
        ApplyBoundaryOperatorsBatch([H, H], [batch[0]["B"], batch[0]["B"]])


def test_batch_autotune_keeps_results(batch):
    Hs = [it["H"] for it in batch]
    Bs = [it["B"] for it in batch]
    before = ApplyBoundaryOperatorsBatch(Hs, Bs)
    saved = BatchTuning()
    tuning, ms = BatchAutotune(Hs, Bs, repeats=2)
    print("autotuned", tuning, ms)
    assert ms > 0 and tuning["mixed_steps"] >= 8 and tuning["plain_blocks"] >= 1
    after = ApplyBoundaryOperatorsBatch(Hs, Bs)
    for k in range(len(batch)):
        assert relerr(after[k], before[k]) < 2e-6
    BatchTuning(**saved)


def test_batch_rejects_fp32_mfma_instances(batch):
    """The batched launch has FP32 and mixed-precision bodies only: an instance
    set to BAEMM_PRECISION_FP32_MFMA is refused with an argument error."""
    H = batch[0]["H"]
    H.SetPrecision("fp32_mfma")
    try:
        with pytest.raises(BaemmError):
            ApplyBoundaryOperatorsBatch([H], [batch[0]["B"]])
    finally:
        H.SetPrecision(batch[0]["prec"])


def test_batch_mass_only_and_empty_operators():
    """Instances whose operator has no kernel part (coefficients (c0, 0, 0, 0):
    the mass term only, no staging) or nothing at all (all zero: C is left as
    it was, as the reference leaves C_out untouched) in one batch with ordinary
    FP32 and mixed instances; each result equals its single-instance apply."""
    rng = np.random.default_rng(31)
    specs = [("mixed", [0.5, 1.0 - 0.5j, 1.0, 0.0]), ("fp32", [0.0, 1.0, 0.0, 0.0]),
             ("mixed", [0.75 - 0.25j, 0.0, 0.0, 0.0]), ("fp32", [0.0, 0.0, 0.0, 0.0]),
             ("mixed", [0.0, 0.0, 0.0, 0.0])]
    Hs, Bs, Cs, singles = [], [], [], []
    try:
        for k, (prec, coeff) in enumerate(specs):
            V, T = perturbed_sphere(2 + k % 2, seed=300 + k)
            W = 4
            H = Helmholtz_HIP(V, T)
            H.SetPrecision(prec)
            H.LoadBoundaryOperators_PL([3.0], np.array(coeff, np.complex128), W, W)
            B = rng.standard_normal((V.shape[0], W)) + 1j * rng.standard_normal((V.shape[0], W))
            C0 = np.full_like(B, 1.0 + 2.0j)
            single = C0.copy()
            H.ApplyBoundaryOperators_PL(1.0, B, 0.0, single)
            Hs.append(H)
            Bs.append(B)
            Cs.append(C0)
            singles.append(single)
        # beta = 0 given, so that C_list is uploaded (the empty operators must leave it)
        out = ApplyBoundaryOperatorsBatch(Hs, Bs, C_list=[c.copy() for c in Cs], beta=[0.0] * len(Hs))
        for k, (C, single) in enumerate(zip(out, singles)):
            print(k, specs[k], "batch vs single", relerr(C, single))
            if not any(specs[k][1]):
                assert np.array_equal(C, Cs[k]), k          # untouched
            elif not any(specs[k][1][1:]):
                assert relerr(C, single) < 1e-14, k         # mass term: same FP64 epilogue
            else:
                assert relerr(C, single) < 2e-6, k
    finally:
        for H in Hs:
            H.close()


def _c4_mix(seed=3):
    """The batched-instances workload: 24 perturbed icospheres (radius
    1 + 0.1 sum of real spherical harmonics of degree <= 4 with seeded
    N(0,1)/(l+1)^2 weights) of 1,280, 5,120 or 20,480 triangles, kappa uniform
    in [1, 16], W = 2^p columns with p uniform in 0..6, SL+DL with complex
    coefficients; the same draws as the batch leg of bench.py."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(24):
        level = int(rng.integers(3, 6))
        V, T = perturbed_sphere(level, seed=1000 + k)
        W = 2 ** int(rng.integers(0, 7))
        kappa = [float(rng.uniform(1.0, 16.0))]
        coeff = np.array([0.0, 1.0 + 0.2j, 1.0 - 0.1j, 0.0])
        B = rng.standard_normal((V.shape[0], W)) + 1j * rng.standard_normal((V.shape[0], W))
        out.append(dict(V=V, T=T, W=W, kappa=kappa, coeff=coeff, B=B))
    return out


def test_c4_mix_in_one_launch_matches_oracle():
    """The full C4 mix (24 instances, including the 20,480-triangle meshes) in
    one persistent mixed-precision launch with autotuned tile shapes: every
    instance within 1e-5 of the FP64 oracle on 48 sampled vertex rows, and
    bit-identical to the same instance batched alone."""
    mix = _c4_mix()
    assert {it["T"].shape[0] for it in mix} == {1280, 5120, 20480}
    Hs = []
    try:
        for it in mix:
            H = Helmholtz_HIP(it["V"], it["T"])
            H.SetPrecision("mixed")
            H.LoadBoundaryOperators_PL(it["kappa"], it["coeff"], it["W"], it["W"])
            Hs.append(H)
        Bs = [it["B"] for it in mix]
        saved = BatchTuning()
        BatchAutotune(Hs, Bs, repeats=1)
        out = ApplyBoundaryOperatorsBatch(Hs, Bs)
        rng = np.random.default_rng(4)
        worst = 0.0
        for k, (it, C) in enumerate(zip(mix, out)):
            m = oracle.Mesh(it["V"], it["T"])
            rows = np.sort(rng.choice(m.nv, 48, replace=False))
            ref = oracle.apply_boundary_rows(m, 1.0, it["B"], 0.0, np.zeros((48, it["W"])), it["kappa"],
                                             it["coeff"], it["W"], it["W"], rows)
            e = relerr(C[rows], ref)
            worst = max(worst, e)
            assert e < 1e-5, (k, it["T"].shape[0], it["W"], e)
        print("C4 mix: worst sampled-row error vs FP64", worst)
        for k in (0, 11, 23):
            alone = ApplyBoundaryOperatorsBatch([Hs[k]], [Bs[k]])
            assert np.array_equal(alone[0], out[k]), k
        BatchTuning(**saved)
    finally:
        for H in Hs:
            H.close()


def test_batched_gmres_matches_single_solves_and_oracle():
    """baemm_gmres_batch on 8 instances (162 to 2,562 vertices, 1 to 16 columns,
    mixed and FP32): the indirect potential operator (1/2, -i kappa, 1, 0) of
    each instance solved together, every Arnoldi step one batched apply.  Each
    solution agrees with the same instance solved alone (baemm_gmres_dev) and
    with the oracle's dense FP64 solve."""
    from baemm_amd import SolveLoadedBatch

    rng = np.random.default_rng(21)
    specs = [(2, 1, "mixed"), (3, 4, "fp32"), (2, 16, "mixed"), (4, 8, "mixed"), (3, 2, "mixed"),
             (2, 3, "fp32"), (4, 1, "mixed"), (3, 16, "mixed")]
    Hs, Bs, refs, singles = [], [], [], []
    try:
        for k, (level, W, prec) in enumerate(specs):
            V, T = perturbed_sphere(level, seed=600 + k)
            kappa = float(rng.uniform(1.0, 8.0))
            coeff = np.array([0.5, -1j * kappa, 1.0, 0.0])
            H = Helmholtz_HIP(V, T)
            H.SetPrecision(prec)
            H.LoadBoundaryOperators_PL([kappa], coeff, W, W)
            B = rng.standard_normal((V.shape[0], W)) + 1j * rng.standard_normal((V.shape[0], W))
            Hs.append(H)
            Bs.append(B)
            m = oracle.Mesh(V, T)
            # GMRES solves the preconditioned system; its solution is A^{-1} B
            refs.append(oracle.solve_boundary(m, [kappa], [coeff], B, W))
        X, its, conv = SolveLoadedBatch(Hs, Bs, cg_tol=1e-12, gmres_tol=1e-9)
        for k, H in enumerate(Hs):
            Xs, its1, ok1 = H.SolveLoaded(Bs[k], cg_tol=1e-12, gmres_tol=1e-9)
            e_single, e_oracle = relerr(X[k], Xs), relerr(X[k], refs[k])
            print(f"instance {k} {specs[k]}: batched {its[k]} iterations, alone {its1}; vs alone {e_single:.2e}, "
                  f"vs FP64 dense solve {e_oracle:.2e}")
            assert conv[k] and ok1, k
            assert e_single < 1e-6 and e_oracle < 1e-5, k
    finally:
        for H in Hs:
            H.close()


def test_batched_gmres_on_the_c4_mix():
    """The 24-instance mix of 1,280 to 20,480 triangles in mixed precision,
    each instance's potential operator (1/2, -i kappa, 1, 0) solved to 1e-6
    with batched Arnoldi steps: every system converges and agrees with its
    solve alone."""
    from baemm_amd import SolveLoadedBatch

    mix = _c4_mix()
    Hs = []
    try:
        for it in mix:
            H = Helmholtz_HIP(it["V"], it["T"])
            H.SetPrecision("mixed")
            H.LoadBoundaryOperators_PL(it["kappa"], [0.5, -1j * it["kappa"][0], 1.0, 0.0], it["W"], it["W"])
            Hs.append(H)
        X, its, conv = SolveLoadedBatch(Hs, [it["B"] for it in mix], cg_tol=1e-10, gmres_tol=1e-6)
        assert all(conv)
        for k in (0, 5, 17, 23):
            Xs, its1, _ = Hs[k].SolveLoaded(mix[k]["B"], cg_tol=1e-10, gmres_tol=1e-6)
            e = relerr(X[k], Xs)
            print(f"C4 instance {k}: {its[k]} / {its1} iterations, batched vs alone {e:.2e}")
            assert e < 1e-5, k
    finally:
        for H in Hs:
            H.close()
