#!/usr/bin/env python3
"""Benchmark of the BAEMM hot path on one MI355X.

Workload (the operator-apply configuration of the metric): icosphere with
81,920 triangles / 40,962 vertices (6-fold midpoint subdivision of the unit
icosahedron), wave numbers kappa in {8, 16} with 32 right-hand sides each
(W = 64 columns, chunk size 32), vertex values i.i.d. complex N(0,1) from seed 1,
single + double layer with coefficient rows (0, 1, 1, 0) and (1/2, -i kappa, 1, 0).

One step = one device-resident SL+DL boundary-operator apply
C = A B  (Av gather, pair kernel, SL diagonal + AvT + mass epilogue), inputs and
outputs resident in HBM.  Reported:

  value              mixed-precision (FP16 hi+lo MFMA) pair*RHS/s = N_t^2 W / t_step
  fp32_value         plain FP32 apply, same workload (the >= 3x target's denominator)
  fp32_mfma          FP32 kernel values contracted with FP32 MFMA, same workload
  roofline           dominant kernel (MFMA pair kernel): algorithmic FLOP/s vs dense
                     FP16 MFMA peak; roofline_valu: kernel-generation FLOP/s vs FP32 peak
  cpu_baseline       FP32 restatement of the reference OpenCL kernel (oracle) on the
                     host cores, bounded sample of target rows, same workload
  mesh_sizes         (--mesh-sizes 3,4,...) the same operator on icospheres of other
                     sizes, mixed and FP32 step and kernel times
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "pair·RHS/s for SL+DL apply (mixed vs FP32); solves/s + p50/p99 at 1e-6"
PEAK_MFMA_F16 = 2500.0   # TFLOP/s, dense FP16 MFMA, MI355X
PEAK_FP32 = 157.3        # TFLOP/s, FP32 vector, MI355X
G_SLDL = 43              # algorithmic flops per pair per wave number, SL+DL (BoundaryOperatorKernel.cl:104-170)


TRAFFIC_PROFILE = ROOT / "profiles" / "bdr_mixed_kernel_icosphere81920_w64_traffic.txt"


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def profiled_traffic():
    """HBM bytes per launch of the MFMA pair kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent)."""
    try:
        for line in TRAFFIC_PROFILE.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch bdr_mixed_kernel:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def workload():
    from baemm_amd.meshes import icosphere

    V, T = icosphere(6)
    W, wcs = 64, 32
    kappa = [8.0, 16.0]
    coeff = np.array([0, 1, 1, 0, 0.5, -1j * 16.0, 1, 0], np.complex128)
    rng = np.random.default_rng(1)
    B = rng.standard_normal((V.shape[0], W)) + 1j * rng.standard_normal((V.shape[0], W))
    return V, T, W, wcs, kappa, coeff, B


def run_gpu(args):
    import ctypes

    from baemm_amd import Helmholtz_HIP, _lib

    V, T, W, wcs, kappa, coeff, B = workload()
    nt, nv = T.shape[0], V.shape[0]
    L = _lib.load()
    H = Helmholtz_HIP(V, T)
    h = H._h
    nbytes = nv * W * 16
    dB = ctypes.c_void_p()
    dC = ctypes.c_void_p()
    _lib.check(L.baemm_dev_alloc(ctypes.byref(dB), nbytes))
    _lib.check(L.baemm_dev_alloc(ctypes.byref(dC), nbytes))
    Bc = np.ascontiguousarray(B)
    _lib.check(L.baemm_dev_upload(dB, Bc.ctypes.data_as(ctypes.c_void_p), nbytes))
    one = (ctypes.c_double * 2)(1.0, 0.0)
    zero = (ctypes.c_double * 2)(0.0, 0.0)
    H.LoadBoundaryOperators_PL(kappa, coeff, W, wcs)

    def step():
        _lib.check(L.baemm_apply_boundary_dev(h, one, dB, W, _lib.C128, zero, dC, W, _lib.C128, None), h)

    results = {}
    # mixed: the library's automatic MFMA kernel choice (the single-wave kernel
    # for this operator); mixed_wave_specialised: the other variant, forced
    for mode in ("mixed", "mixed_wave_specialised", "fp32", "fp32_mfma"):
        log(f"{mode} apply: {args.warmup} warm-up + {args.steps} timed steps")
        H.SetPrecision({"fp32": "fp32", "fp32_mfma": "fp32_mfma"}.get(mode, "mixed"))
        H.SetMixedKernel(1 if mode == "mixed_wave_specialised" else 2)
        for _ in range(args.warmup):
            step()
        _lib.check(L.baemm_dev_synchronize())
        H.SetTiming(True)
        _lib.check(L.baemm_dev_synchronize())
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        _lib.check(L.baemm_dev_synchronize())
        t1 = time.perf_counter()
        kms, launches = H.GetTiming()
        H.SetTiming(False)
        out = np.zeros((nv, W), np.complex128)
        _lib.check(L.baemm_dev_download(out.ctypes.data_as(ctypes.c_void_p), dC, nbytes))
        results[mode] = {
            "ms_per_step": (t1 - t0) * 1e3 / args.steps,
            "kernel_ms": kms / max(1, args.steps),
            "kernel_launch_groups": launches,
            "out": out,
        }
    # host-pointer path (PCIe transfers of B and C included), mixed precision
    H.SetPrecision("mixed")
    H.SetMixedKernel(2)
    Ch = np.zeros((nv, W), np.complex128)
    H.ApplyBoundaryOperators_PL(1.0, Bc, 0.0, Ch)
    t0 = time.perf_counter()
    for _ in range(3):
        H.ApplyBoundaryOperators_PL(1.0, Bc, 0.0, Ch)
    results["host_path_ms"] = (time.perf_counter() - t0) * 1e3 / 3
    # one wave number with 16 columns (the reference example's shape): the first
    # 16 columns of the same vertex values, kappa 8, SL+DL row (1/2, -8i, 1, 0)
    H.LoadBoundaryOperators_PL([8.0], np.array([0.5, -8j, 1.0, 0.0]), 16, 16)
    small = {}
    for mode in ("fp32", "mixed", "fp32_mfma"):
        H.SetPrecision(mode)
        step16 = lambda: _lib.check(L.baemm_apply_boundary_dev(h, one, dB, W, _lib.C128, zero, dC, W, _lib.C128,
                                                               None), h)
        step16()
        _lib.check(L.baemm_dev_synchronize())
        H.SetTiming(True)
        t0 = time.perf_counter()
        for _ in range(5):
            step16()
        _lib.check(L.baemm_dev_synchronize())
        dt = (time.perf_counter() - t0) / 5
        kms, _ = H.GetTiming()
        H.SetTiming(False)
        small[mode] = {"ms_per_step": dt * 1e3, "kernel_ms": kms / 5,
                       "pair_rhs_per_s": float(nt) * nt * 16 / dt}
    results["w16"] = small
    # accuracy of the mixed apply relative to the FP32 apply (same inputs)
    dm = results["mixed"]["out"] - results["fp32"]["out"]
    rel_mixed_vs_fp32 = float(np.linalg.norm(dm) / np.linalg.norm(results["fp32"]["out"]))
    L.baemm_dev_free(dB)
    L.baemm_dev_free(dC)
    H.close()
    return nt, nv, W, results, rel_mixed_vs_fp32


def serving_workload(n, max_level, seed=7):
    """Gauss-Newton-step requests: perturbed icospheres with triangle counts
    log-uniform over 1280 * 4^(0..max_level-3) (levels 3..max_level), a few
    distinct meshes per level (mesh ids, so geometry can stay resident),
    kappa uniform in [1, min(32, 2 pi / (10 h))] (>= 10 points per
    wavelength, h the mean edge length), W = 2^p RHS with p uniform in 0..6,
    256 Fibonacci measurement directions, seeded synthetic far-field data."""
    from baemm_amd.meshes import fibonacci_sphere, perturbed_sphere, random_directions

    rng = np.random.default_rng(seed)
    meas = fibonacci_sphere(256)
    meshes = {}
    reqs = []
    for _ in range(n):
        level = int(rng.integers(3, max_level + 1))
        variant = int(rng.integers(0, 2))
        key = (level, variant)
        if key not in meshes:
            meshes[key] = perturbed_sphere(level, seed=100 + 10 * level + variant)
        V, T = meshes[key]
        e = V[T[:, 1]] - V[T[:, 0]]
        h = float(np.mean(np.linalg.norm(e, axis=1)))
        kmax = min(32.0, 2.0 * math.pi / (10.0 * h))
        kappa = float(rng.uniform(1.0, max(1.0, kmax)))
        W = 2 ** int(rng.integers(0, 7))
        dirs = random_directions(W, seed=int(rng.integers(1 << 30)))
        fobs = 0.05 * (rng.standard_normal((meas.shape[0], W)) + 1j * rng.standard_normal((meas.shape[0], W)))
        reqs.append(dict(V=V, T=T, meas=meas, kappa=[kappa], dirs=dirs, fobs=fobs, mesh_id=100 * level + variant,
                         nt=T.shape[0], W=W))
    return reqs


def run_serving(reqs, rate, slots, precision, seed=11, policy="fifo", fusion=False):
    """Submits the requests at Poisson arrival times (rate per second) from a
    driver thread and waits for all; returns the server statistics.  policy:
    the admission order ("fifo", or "sjf": least estimated work first, with
    aging); fusion: batches of up to `slots` requests with batched GMRES."""
    import threading

    from baemm_amd.serving import GNServer

    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    ids = [None] * len(reqs)
    with GNServer(slots=slots, hbm_budget_bytes=200 << 30, policy=policy, fusion=fusion) as srv:
        t0 = time.perf_counter()

        def driver():
            for k, r in enumerate(reqs):
                dt = arrivals[k] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                ids[k] = srv.submit(r["V"], r["T"], r["meas"], r["kappa"], r["dirs"], far_field_data=r["fobs"],
                                    mesh_id=r["mesh_id"], precision=precision, cg_tol=1e-10, gmres_tol=1e-6)

        th = threading.Thread(target=driver)
        th.start()
        th.join()
        statuses = []
        for k, i in enumerate(ids):
            statuses.append(srv.wait(i)[0])
            log(f"serving ({precision}, {slots} slots, {policy}{', fused' if fusion else ''}): "
                f"request {k + 1}/{len(ids)} done")
        wall = time.perf_counter() - t0
        st = srv.stats()
    st["wall_s"] = wall
    st["not_converged"] = int(sum(1 for x in statuses if x != 0))
    return st


def run_solve_leg(repeats=2, graphs=True):
    """Full sound-soft solve: torus R=1, r=0.4 on a 240 x 80 grid (38,400
    triangles), kappa = 10, 32 seeded plane waves, FarField (incident wave,
    GMRES(30) to 1e-6 with the PCG mass-inverse preconditioner, far field at 256
    Fibonacci directions).  Seconds per solve in FP32, mixed, and mixed with
    hipGraph-replayed Arnoldi steps."""
    from baemm_amd import Helmholtz_HIP
    from baemm_amd.meshes import fibonacci_sphere, random_directions, torus

    V, T = torus(1.0, 0.4, 240, 80)
    meas = fibonacci_sphere(256)
    dirs = random_directions(32, seed=2)
    out = {}
    ref = None
    with Helmholtz_HIP(V, T, meas) as H:
        legs = [("fp32", "fp32", False), ("mixed", "mixed", False)]
        if graphs:
            legs.append(("mixed_graphs", "mixed", True))
        for leg, prec, use_graphs in legs:
            log(f"solve leg {leg}")
            H.SetPrecision(prec)
            H.SetGraphs(use_graphs)
            H.FarField([10.0], 1, dirs, 32, cg_tol=1e-10, gmres_tol=1e-6)    # warm-up (workspaces, graphs)
            H.ResetSolverStats()
            t0 = time.perf_counter()
            for _ in range(repeats):
                F = H.FarField([10.0], 1, dirs, 32, cg_tol=1e-10, gmres_tol=1e-6)
            dt = (time.perf_counter() - t0) / repeats
            st = H.SolverStats()
            if ref is None:
                ref = F
            out[leg] = {"s_per_solve": dt, "gmres_iterations_per_solve": st["gmres_iterations"] / repeats,
                        "rel_residual": st["rel_residual"],
                        "far_field_rel_diff_vs_fp32": float(np.linalg.norm(F - ref) / np.linalg.norm(ref))}
        H.SetGraphs(False)
    return out


def run_potential_leg(repeats=3):
    """Potential (near field) of the same SL+DL operator at 65,536 points on the
    sphere of radius 1.5 (Fibonacci lattice) around the 81,920-triangle
    icosphere, the same 64 vertex-value columns and kappa {8, 16}.  Kernel time
    (HIP events around the pair kernels) and host-path wall time per apply in
    each precision (host arrays: the wall time includes the PCIe transfers of
    B, the points and the 65,536 x 64 result); mixed precision runs the
    automatic MFMA kernel choice (the single-wave kernel for this operator) and
    FP32 MFMA the wave-specialised kernel, both with the points as targets."""
    from baemm_amd import Helmholtz_HIP
    from baemm_amd.meshes import fibonacci_sphere

    V, T, W, wcs, kappa, coeff, B = workload()
    P = 65536
    pts = 1.5 * fibonacci_sphere(P)
    out, ref = {}, None
    with Helmholtz_HIP(V, T) as H:
        H.LoadParameters(kappa, coeff, W, wcs)
        for mode in ("fp32", "mixed", "fp32_mfma"):
            log(f"potential leg {mode}")
            H.SetPrecision(mode)
            U = np.zeros((P, W), np.complex128)
            H.ApplyNearFieldOperators_PL(1.0, B, 0.0, U, evaluation_points=pts)   # warm-up
            H.SetTiming(True)
            t0 = time.perf_counter()
            for _ in range(repeats):
                H.ApplyNearFieldOperators_PL(1.0, B, 0.0, U, evaluation_points=pts)
            wall = (time.perf_counter() - t0) / repeats
            kms, _ = H.GetTiming()
            H.SetTiming(False)
            if ref is None:
                ref = U.copy()
            out[mode] = {"kernel_ms": kms / repeats, "wall_ms": wall * 1e3,
                         "pair_rhs_per_s_kernel": float(P) * T.shape[0] * W / (kms / repeats * 1e-3),
                         "rel_diff_vs_fp32": float(np.linalg.norm(U - ref) / np.linalg.norm(ref))}
    out["workload"] = ("potential at 65536 points (radius 1.5) of the SL+DL operator on the 81920-triangle "
                       "icosphere, kappa {8,16} x 32 RHS")
    return out


def run_mesh_size_leg(levels):
    """The operator of the headline workload (SL+DL, kappa {8, 16} x 32 RHS =
    64 columns, complex N(0,1) vertex values) on icospheres of the given
    subdivision levels (3: 1,280 ... 7: 327,680 triangles), device-resident
    applies in mixed precision (automatic kernel choice) and FP32: wall time per
    step over enough steps for about 0.2 s, and the pair kernels' HIP-event time
    per step."""
    import ctypes

    from baemm_amd import Helmholtz_HIP, _lib
    from baemm_amd.meshes import icosphere

    L = _lib.load()
    W, wcs = 64, 32
    kappa = [8.0, 16.0]
    coeff = np.array([0, 1, 1, 0, 0.5, -1j * 16.0, 1, 0], np.complex128)
    one = (ctypes.c_double * 2)(1.0, 0.0)
    zero = (ctypes.c_double * 2)(0.0, 0.0)
    out = {}
    for level in levels:
        V, T = icosphere(level)
        nt, nv = T.shape[0], V.shape[0]
        log(f"mesh-size leg: {nt} triangles")
        rng = np.random.default_rng(1)
        B = np.ascontiguousarray(rng.standard_normal((nv, W)) + 1j * rng.standard_normal((nv, W)))
        nbytes = B.nbytes
        dB, dC = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(L.baemm_dev_alloc(ctypes.byref(dB), nbytes))
        _lib.check(L.baemm_dev_alloc(ctypes.byref(dC), nbytes))
        _lib.check(L.baemm_dev_upload(dB, B.ctypes.data_as(ctypes.c_void_p), nbytes))
        H = Helmholtz_HIP(V, T)
        h = H._h
        H.LoadBoundaryOperators_PL(kappa, coeff, W, wcs)
        res = {"triangles": nt}
        for mode in ("mixed", "fp32"):
            H.SetPrecision(mode)

            def step():
                _lib.check(L.baemm_apply_boundary_dev(h, one, dB, W, _lib.C128, zero, dC, W, _lib.C128, None), h)

            step()
            _lib.check(L.baemm_dev_synchronize())
            t0 = time.perf_counter()
            step()
            _lib.check(L.baemm_dev_synchronize())
            n = int(min(2000, max(3, 0.2 / max(time.perf_counter() - t0, 1e-6))))
            t0 = time.perf_counter()
            for _ in range(n):
                step()
            _lib.check(L.baemm_dev_synchronize())
            wall = (time.perf_counter() - t0) / n
            H.SetTiming(True)
            for _ in range(n):
                step()
            _lib.check(L.baemm_dev_synchronize())
            kms, _ = H.GetTiming()
            H.SetTiming(False)
            res[mode] = {"steps": n, "ms_per_step": wall * 1e3, "kernel_ms": kms / n,
                         "pair_rhs_per_s": float(nt) * nt * W / wall}
        H.close()
        L.baemm_dev_free(dB)
        L.baemm_dev_free(dC)
        out[str(nt)] = res
    return out


def run_batch_leg(repeats=5, seed=3, n_inst=24, levels=(3, 6), gmres=True):
    """Batched uneven instances: n_inst perturbed icospheres (radius 1 + 0.1 sum
    of real spherical harmonics of degree <= 4, seeded N(0,1)/(l+1)^2 weights)
    of icosphere levels uniform in [levels[0], levels[1]) (the default: 1,280,
    5,120 or 20,480 triangles), kappa uniform in [1, 16], W = 2^p RHS with p
    uniform in 0..6 (one wave-number chunk each), SL+DL with complex
    coefficients, mixed precision.  Times one persistent launch per apply
    (autotuned tile shapes, or every instance on its own launch where that is
    faster) against the single-instance applies."""
    import ctypes

    from baemm_amd import ApplyBoundaryOperatorsBatch, BatchAutotune, BatchStats, BatchTuning, Helmholtz_HIP, _lib
    from baemm_amd.meshes import perturbed_sphere

    rng = np.random.default_rng(seed)
    inst, Bs, kappas = [], [], []
    pairs_rhs = 0.0
    saved_tuning = BatchTuning()
    for k in range(n_inst):
        level = int(rng.integers(*levels))
        V, T = perturbed_sphere(level, seed=1000 + k)
        W = 2 ** int(rng.integers(0, 7))
        kappa = [float(rng.uniform(1.0, 16.0))]
        kappas.append(kappa)
        coeff = np.array([0.0, 1.0 + 0.2j, 1.0 - 0.1j, 0.0])
        H = Helmholtz_HIP(V, T)
        H.SetPrecision("mixed")
        H.LoadBoundaryOperators_PL(kappa, coeff, W, W)
        inst.append(H)
        Bs.append(rng.standard_normal((V.shape[0], W)) + 1j * rng.standard_normal((V.shape[0], W)))
        pairs_rhs += float(T.shape[0]) ** 2 * W
    L = _lib.load()
    # device-resident inputs and outputs for the timed loops
    dB, dC = [], []
    for b in Bs:
        for lst in (dB, dC):
            p = ctypes.c_void_p()
            _lib.check(L.baemm_dev_alloc(ctypes.byref(p), b.nbytes))
            lst.append(p)
        _lib.check(L.baemm_dev_upload(dB[-1], b.ctypes.data_as(ctypes.c_void_p), b.nbytes))
    n = len(inst)
    hs = (ctypes.c_void_p * n)(*[H._h.value for H in inst])
    pB = (ctypes.c_void_p * n)(*[d.value for d in dB])
    pC = (ctypes.c_void_p * n)(*[d.value for d in dC])
    ld = (ctypes.c_int64 * n)(*[b.shape[1] for b in Bs])
    log("batch leg: autotune")
    tuning, best = BatchAutotune(inst, Bs, repeats=2)

    def batched():
        _lib.check(L.baemm_apply_boundary_batch(n, hs, None, pB, ld, _lib.C128, None, pC, ld, _lib.C128, None))

    one = (ctypes.c_double * 2)(1.0, 0.0)
    zero = (ctypes.c_double * 2)(0.0, 0.0)

    def single():
        for k, H in enumerate(inst):
            _lib.check(L.baemm_apply_boundary_dev(H._h, one, dB[k], ld[k], _lib.C128, zero, dC[k], ld[k], _lib.C128,
                                                  None), H._h)

    # Each apply is complete before the next one starts (as in a solver,
    # whose next step needs the result): the batched call joins its
    # instances' streams, the 24 single applies are joined by a device
    # synchronisation.  "single_pipelined" lets consecutive rounds of single
    # applies overlap (no join), a lower bound no dependent use reaches.
    times = {}
    for name, fn, join in (("batched", batched, True), ("single", single, True),
                           ("single_pipelined", single, False)):
        fn()
        _lib.check(L.baemm_dev_synchronize())
        t0 = time.perf_counter()
        for _ in range(repeats):
            fn()
            if join:
                _lib.check(L.baemm_dev_synchronize())
        _lib.check(L.baemm_dev_synchronize())
        times[name] = (time.perf_counter() - t0) / repeats
    batched()
    _lib.check(L.baemm_dev_synchronize())
    stats = BatchStats()
    res = {
        "pair_rhs_per_apply": pairs_rhs,
        "batched_ms": times["batched"] * 1e3,
        "single_ms": times["single"] * 1e3,
        "single_pipelined_ms": times["single_pipelined"] * 1e3,
        "batched_pair_rhs_per_s": pairs_rhs / times["batched"],
        "single_pair_rhs_per_s": pairs_rhs / times["single"],
        "speedup": times["single"] / times["batched"],
        "speedup_vs_pipelined": times["single_pipelined"] / times["batched"],
        "tuning": tuning,
        "work_items": stats["items_mixed"],
        "persistent_grid": stats["grid_mixed"],
    }
    if not gmres:
        for d in dB + dC:
            L.baemm_dev_free(d)
        for H in inst:
            H.close()
        BatchTuning(**saved_tuning)
        return res
    # GMRES on the same instances: each one's indirect potential operator
    # (1/2, -i kappa, 1, 0) solved to 1e-6, all together with batched Arnoldi
    # steps against one after the other
    log("batch leg: batched GMRES")
    for H, kap, b in zip(inst, kappas, Bs):
        H.LoadBoundaryOperators_PL(kap, [0.5, -1j * kap[0], 1.0, 0.0], b.shape[1], b.shape[1])
    its = (ctypes.c_int64 * n)()
    conv = (ctypes.c_int * n)()

    def gm_batched():
        st = L.baemm_gmres_batch(n, hs, pB, pC, 1e-10, 1e-6, its, conv, None)
        assert st == _lib.OK, st

    def gm_sequential():
        one_its = ctypes.c_int64(0)
        for k, H in enumerate(inst):
            _lib.check(L.baemm_gmres_dev(H._h, dB[k], dC[k], 1e-10, 1e-6, ctypes.byref(one_its), None), H._h)

    gm_times = {}
    for name, fn in (("batched", gm_batched), ("sequential", gm_sequential)):
        fn()
        t0 = time.perf_counter()
        for _ in range(2):
            fn()
        gm_times[name] = (time.perf_counter() - t0) / 2
    for d in dB + dC:
        L.baemm_dev_free(d)
    for H in inst:
        H.close()
    BatchTuning(**saved_tuning)
    return {
        "workload": "24 perturbed icospheres of 1280/5120/20480 triangles, kappa in [1,16], W = 2^p (p in 0..6), "
                    "SL+DL, mixed precision",
        **res,
        "gmres": {
            "operator": "(1/2, -i kappa, 1, 0) per instance, GMRES(30) to 1e-6, PCG mass-inverse preconditioner",
            "batched_s": gm_times["batched"],
            "sequential_s": gm_times["sequential"],
            "batched_solves_per_s": n / gm_times["batched"],
            "sequential_solves_per_s": n / gm_times["sequential"],
            "speedup": gm_times["sequential"] / gm_times["batched"],
            "iterations": [int(x) for x in its],
        },
    }


def cpu_baseline(rows_target_seconds=12.0):
    """Times the FP32 restatement of BoundaryOperatorKernel.cl (oracle) on the
    host cores for a bounded sample of target rows of the same workload."""
    import oracle

    V, T, W, wcs, kappa, coeff, B = workload()
    m = oracle.Mesh(V, T)
    ch = oracle.F32Chain(m)
    kap32, c32, bits, _ = ch.params32(kappa, coeff, W, wcs)
    b32 = np.ascontiguousarray((ch.Av @ B).astype(np.complex64))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or 1))
    oracle.set_threads(threads)
    rows = 64
    t = 0.0
    while True:
        t0 = time.perf_counter()
        ch.kernel(kap32, c32, bits, b32, wcs, 0, rows)
        t = time.perf_counter() - t0
        if t > 1.0 or rows >= m.nt:
            break
        rows = min(m.nt, rows * 4)
    rows = int(min(m.nt, max(rows, rows * rows_target_seconds / max(t, 1e-3))))
    rows = max(64, (rows // 64) * 64)
    t0 = time.perf_counter()
    ch.kernel(kap32, c32, bits, b32, wcs, 0, rows)
    t = time.perf_counter() - t0
    value = rows * m.nt * W / t
    return {
        "value": value,
        "unit": "pair·RHS/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP32 restatement of BoundaryOperatorKernel.cl (64-source blocks, per-chunk loop), "
                  f"{rows} of {m.nt} target rows x all {m.nt} sources x {W} RHS, {t:.1f} s",
        "seconds": t,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="skip the torus solve leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-instances leg")
    ap.add_argument("--no-graphs", action="store_true", help="skip the hipGraph-replay solve of the solve leg")
    ap.add_argument("--no-potential", action="store_true", help="skip the potential (near-field) leg")
    ap.add_argument("--serve-requests", type=int, default=24, help="requests of the serving leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=40.0,
                    help="Poisson arrival rate of the serving leg (1/s); the default saturates the server")
    ap.add_argument("--serve-max-level", type=int, default=6,
                    help="largest icosphere level of the serving leg (6: 81920, 7: 327680 triangles)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--mesh-sizes", default="",
                    help="comma-separated icosphere levels (3: 1,280 ... 7: 327,680 triangles) for an extra leg "
                         "with the headline operator on each mesh (default: none)")
    ap.add_argument("--only-mesh-sizes", action="store_true",
                    help="run only the --mesh-sizes leg and print its JSON object")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "the apply would shard by target rows (one all-reduce of vertex sums) or by request.", file=sys.stderr)
        sys.exit(2)
    levels = [int(x) for x in args.mesh_sizes.split(",") if x.strip()]
    if args.only_mesh_sizes:
        print(json.dumps({"mesh_sizes": run_mesh_size_leg(levels)}))
        return
    nt, nv, W, res, rel = run_gpu(args)
    pairs_rhs = float(nt) * nt * W
    mixed = res["mixed"]
    fp32 = res["fp32"]
    flops = float(nt) * nt * (8 * W + G_SLDL * 2)
    kt = mixed["kernel_ms"] * 1e-3
    achieved = flops / kt / 1e12
    gen_tf = float(nt) * nt * G_SLDL * 2 / kt / 1e12
    line = {
        "metric": METRIC,
        "value": pairs_rhs / (mixed["ms_per_step"] * 1e-3),
        "unit": "pair·RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "arithmetic": "kernel values generated in f32; complex contraction as f16 hi+lo products on MFMA with f32 "
                      "accumulation (split FP32); segment partials summed in f64",
        "data": "synthetic",
        "config": {
            "workload": "SL+DL boundary-operator apply, icosphere 81920 triangles / 40962 vertices, "
                        "kappa {8,16} x 32 RHS = 64 columns, complex N(0,1) vertex values (seed 1)",
            "n_triangles": nt,
            "n_vertices": nv,
            "rhs_columns": W,
            "wave_chunk_size": 32,
        },
        "mixed_kernel": "bdr_mixed_kernel (single-wave, automatic choice)",
        "mixed_wave_specialised_ms_per_step": res["mixed_wave_specialised"]["ms_per_step"],
        "mixed_wave_specialised_kernel_ms": res["mixed_wave_specialised"]["kernel_ms"],
        "rel_diff_mixed_variants": float(np.linalg.norm(res["mixed_wave_specialised"]["out"] - mixed["out"]) /
                                         np.linalg.norm(mixed["out"])),
        "fp32_value": pairs_rhs / (fp32["ms_per_step"] * 1e-3),
        "fp32_ms_per_step": fp32["ms_per_step"],
        "speedup_mixed_vs_fp32": fp32["ms_per_step"] / mixed["ms_per_step"],
        "rel_diff_mixed_vs_fp32": rel,
        "w16": dict(workload="same mesh, one wave number (kappa 8) x 16 RHS, SL+DL", **res["w16"]),
        "host_path": {"ms_per_apply": res["host_path_ms"], "value": pairs_rhs / (res["host_path_ms"] * 1e-3),
                      "note": "baemm_apply_boundary with host arrays: 42 MB in, 42 MB out over PCIe per apply"},
        "fp32_mfma": {
            "value": pairs_rhs / (res["fp32_mfma"]["ms_per_step"] * 1e-3),
            "ms_per_step": res["fp32_mfma"]["ms_per_step"],
            "kernel_ms": res["fp32_mfma"]["kernel_ms"],
            "speedup_vs_fp32": fp32["ms_per_step"] / res["fp32_mfma"]["ms_per_step"],
            "rel_diff_vs_fp32": float(np.linalg.norm(res["fp32_mfma"]["out"] - fp32["out"]) /
                                      np.linalg.norm(fp32["out"])),
            "arithmetic": "kernel values in f32, contraction on v_mfma_f32_32x32x2_f32 with f32 accumulation",
            # the contraction (8 flops per pair and RHS column) against the dense FP32 matrix peak
            "roofline": {"bound": "mfma", "unit": "TFLOP/s", "peak": 157.3,
                         "achieved": float(nt) * nt * 8 * W / (res["fp32_mfma"]["kernel_ms"] * 1e-3) / 1e12,
                         "frac": float(nt) * nt * 8 * W / (res["fp32_mfma"]["kernel_ms"] * 1e-3) / 1e12 / 157.3,
                         "kernel": "bdr_mixed_ws_kernel<32, 2, false, true>"},
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_MFMA_F16,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_MFMA_F16,
            "traffic": profiled_traffic(),
            "kernel": "bdr_mixed_kernel",
            "kernel_ms": mixed["kernel_ms"],
            "flops_per_launch": flops,
        },
        "roofline_valu": {
            "bound": "fp32-valu",
            "achieved": gen_tf,
            "peak": PEAK_FP32,
            "unit": "TFLOP/s",
            "frac": gen_tf / PEAK_FP32,
            "note": "kernel-value generation flops (43 per pair per wave number) of the MFMA kernel",
        },
        "roofline_fp32_kernel": {
            "bound": "fp32-valu",
            "achieved": flops / (fp32["kernel_ms"] * 1e-3) / 1e12,
            "peak": PEAK_FP32,
            "unit": "TFLOP/s",
            "frac": flops / (fp32["kernel_ms"] * 1e-3) / 1e12 / PEAK_FP32,
            "kernel": "pair_plain_kernel",
            "kernel_ms": fp32["kernel_ms"],
        },
        "cpu_baseline": None,
    }
    if not args.no_batch:
        line["batch"] = run_batch_leg()
        log("batch leg: many small instances")
        small = run_batch_leg(seed=5, n_inst=96, levels=(2, 4), gmres=False)
        line["batch"]["small_instances"] = {
            "workload": "96 perturbed icospheres of 320 or 1280 triangles, otherwise as above", **small}
    if not args.no_potential:
        line["potential"] = run_potential_leg()
    if levels:
        line["mesh_sizes"] = run_mesh_size_leg(levels)
    if not args.no_solve:
        line["solve"] = dict(workload="sound-soft torus 38400 triangles, kappa 10, 32 plane waves, FarField, "
                                      "GMRES(30) 1e-6, 256 far-field directions", **run_solve_leg(graphs=not args.no_graphs))
    if args.serve_requests > 0:
        log("serving leg: building requests")
        reqs = serving_workload(args.serve_requests, args.serve_max_level)
        mixed_st = run_serving(reqs, args.serve_rate, args.serve_slots, "mixed")
        sjf_st = run_serving(reqs, args.serve_rate, args.serve_slots, "mixed", policy="sjf")
        fused_st = run_serving(reqs, args.serve_rate, args.serve_slots, "mixed", fusion=True)
        seq_st = run_serving(reqs, args.serve_rate, 1, "fp32")
        keys = ("completed", "failed", "not_converged", "solves_per_s", "p50_ms", "p99_ms", "mean_ms", "max_in_flight",
                "peak_bytes", "gmres_iterations", "wall_s")
        line["serving"] = {
            "workload": f"{args.serve_requests} Gauss-Newton-step requests (forward solve + far field, Neumann data, "
                        f"adjoint derivative of the misfit), GMRES 1e-6, Poisson arrivals {args.serve_rate}/s, "
                        f"perturbed icospheres 1280..{1280 * 4 ** (args.serve_max_level - 3)} triangles "
                        f"(log-uniform), W = 1..64, 256 far-field directions",
            "offered_rate_per_s": args.serve_rate,
            "continuous_mixed": dict({k: mixed_st[k] for k in keys}, slots=args.serve_slots),
            "continuous_mixed_shortest_first": dict({k: sjf_st[k] for k in keys}, slots=args.serve_slots),
            "fused_mixed": dict({k: fused_st[k] for k in keys}, slots=args.serve_slots,
                                note="batches of up to `slots` queued requests, each solve one batched GMRES"),
            "sequential_fp32": dict({k: seq_st[k] for k in keys}, slots=1),
            "speedup_solves_per_s": mixed_st["solves_per_s"] / max(seq_st["solves_per_s"], 1e-12),
        }
        # a stream of small requests only (1,280 and 5,120 triangles), where the
        # launches of each request, not the pair kernels, dominate
        log("serving leg: small requests")
        small = serving_workload(2 * args.serve_requests, 4, seed=8)
        sm_cont = run_serving(small, 4 * args.serve_rate, args.serve_slots, "mixed")
        sm_fused = run_serving(small, 4 * args.serve_rate, args.serve_slots, "mixed", fusion=True)
        line["serving"]["small_requests"] = {
            "workload": f"{2 * args.serve_requests} requests of 1280 or 5120 triangles, Poisson arrivals "
                        f"{4 * args.serve_rate}/s, otherwise as above",
            "continuous_mixed": dict({k: sm_cont[k] for k in keys}, slots=args.serve_slots),
            "fused_mixed": dict({k: sm_fused[k] for k in keys}, slots=args.serve_slots),
        }
    if not args.no_cpu_baseline:
        log("cpu baseline (bounded sample)")
        try:
            line["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # the baseline must not hide the GPU result
            line["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
