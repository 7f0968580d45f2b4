"""Request loop (baemm_server_*): Gauss-Newton-step requests run concurrently
give the same results as the same requests run alone, and the same as the
Python mirror's FarField / Derivative_FF_Helper / AdjointDerivative_FF calls;
admission respects the device-memory budget."""
import numpy as np
import pytest

from baemm_amd import BaemmError, Helmholtz_HIP
from baemm_amd.meshes import fibonacci_sphere, perturbed_sphere, random_directions
from baemm_amd.serving import GNServer
from conftest import relerr

pytestmark = pytest.mark.gpu

MEAS = fibonacci_sphere(64)


def _requests():
    rng = np.random.default_rng(5)
    out = []
    specs = [(2, 0, [2.0], 4, "fp32"), (3, 1, [3.0, 4.0], 8, "mixed"), (2, 2, [1.5], 16, "mixed"),
             (3, 1, [5.0], 2, "fp32"), (2, 0, [2.5], 1, "mixed"), (3, 3, [4.0], 32, "mixed")]
    for level, mesh_seed, kappa, wcs, prec in specs:
        V, T = perturbed_sphere(level, seed=20 + mesh_seed)
        dirs = random_directions(wcs, seed=int(rng.integers(1000)))
        W = len(kappa) * wcs
        fobs = 0.1 * (rng.standard_normal((MEAS.shape[0], W)) + 1j * rng.standard_normal((MEAS.shape[0], W)))
        out.append(dict(V=V, T=T, kappa=kappa, dirs=dirs, fobs=fobs, mesh_id=100 * level + mesh_seed, prec=prec))
    return out


def _run(server, reqs):
    ids = [server.submit(r["V"], r["T"], MEAS, r["kappa"], r["dirs"], far_field_data=r["fobs"],
                         mesh_id=r["mesh_id"], precision=r["prec"], cg_tol=1e-10, gmres_tol=1e-6) for r in reqs]
    return [server.wait(i) for i in ids]


@pytest.fixture(scope="module")
def reqs():
    return _requests()


def test_results_same_alone_and_in_the_mix(reqs):
    alone = []
    with GNServer(slots=1) as srv:
        for r in reqs:
            alone.extend(_run(srv, [r]))
    with GNServer(slots=4) as srv:
        mixed = _run(srv, reqs)
        st = srv.stats()
    print("mix stats", st)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["max_in_flight"] >= 2
    assert st["p50_ms"] <= st["p99_ms"] and st["solves_per_s"] > 0
    for k, (a, b) in enumerate(zip(alone, mixed)):
        assert a[0] == 0 and b[0] == 0, (k, a[0], b[0])
        # identical kernels on identical inputs: bit-identical results
        assert np.array_equal(a[1], b[1]), k
        assert np.array_equal(a[2], b[2]), k


def _completion_order(srv, reqs):
    """Submits all requests at once, waits, returns (results, ids in completion order)."""
    res = _run(srv, reqs)
    lat = [r[3] for r in res]
    return res, sorted(range(len(reqs)), key=lambda k: lat[k])


def test_shortest_first_policy(reqs):
    """Shortest-estimated-work-first admission (with aging) changes only the
    order: every request completes with the bits of the FIFO run.  With one
    slot, the requests queued behind the first admitted one run in order of
    estimated work N_t^2 * W; with max_skips = 0 every request ages at once
    and the order is the arrival order again."""
    work = [r["T"].shape[0] ** 2 * len(r["kappa"]) * r["dirs"].shape[0] for r in reqs]
    assert len(set(work)) == len(work)
    with GNServer(slots=1) as srv:
        fifo, order_fifo = _completion_order(srv, reqs)
    with GNServer(slots=1, policy="sjf", max_skips=100) as srv:
        sjf, order_sjf = _completion_order(srv, reqs)
        st = srv.stats()
    with GNServer(slots=1, policy="sjf", max_skips=0) as srv:
        aged, order_aged = _completion_order(srv, reqs)
    print("work", work, "fifo", order_fifo, "sjf", order_sjf, "aged", order_aged)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for k in range(len(reqs)):
        assert fifo[k][0] == 0 and sjf[k][0] == 0 and aged[k][0] == 0, k
        assert np.array_equal(fifo[k][1], sjf[k][1]) and np.array_equal(fifo[k][2], sjf[k][2]), k
        assert np.array_equal(fifo[k][1], aged[k][1]), k
    # the first admission may happen before the queue is full; the rest follow the policy
    assert order_fifo[1:] == sorted(order_fifo[1:])
    assert [work[k] for k in order_sjf[1:]] == sorted(work[k] for k in order_sjf[1:])
    assert order_aged[1:] == sorted(order_aged[1:])
    with pytest.raises(KeyError):
        GNServer(slots=1, policy="lifo")


def test_request_matches_python_mirror(reqs):
    r = reqs[1]
    with GNServer(slots=2) as srv:
        st, F, G, lat, its = _run(srv, [r])[0]
    K = len(r["kappa"])
    wcs = r["dirs"].shape[0]
    with Helmholtz_HIP(r["V"], r["T"], MEAS) as H:
        H.SetPrecision(r["prec"])
        Fm = H.FarField(r["kappa"], K, r["dirs"], wcs, cg_tol=1e-10, gmres_tol=1e-6)
        du = H.Derivative_FF_Helper(r["kappa"], K, r["dirs"], wcs, cg_tol=1e-10, gmres_tol=1e-6)
        Gm, _ = H.AdjointDerivative_FF(r["kappa"], K, r["dirs"], wcs, Fm - r["fobs"], du_dn=du, cg_tol=1e-10,
                                       gmres_tol=1e-6)
    assert relerr(F, Fm) < 1e-12
    assert relerr(G, Gm) < 1e-12
    assert its > 0


def test_admission_by_memory(reqs):
    # budget for one request at a time: never two in flight
    small = [reqs[0], reqs[2], reqs[4]]
    need = [GNServer.request_bytes(r["V"].shape[0], r["T"].shape[0], MEAS.shape[0], len(r["kappa"]),
                                   r["dirs"].shape[0], r["prec"]) for r in small]
    budget = int(1.5 * max(need))
    assert 2 * min(need) > budget
    with GNServer(slots=4, hbm_budget_bytes=budget) as srv:
        res = _run(srv, small)
        st = srv.stats()
    print("budget stats", st)
    assert all(x[0] == 0 for x in res)
    assert st["max_in_flight"] == 1
    # a request larger than the whole budget fails instead of waiting forever
    with GNServer(slots=2, hbm_budget_bytes=1 << 20) as srv:
        rid = srv.submit(reqs[0]["V"], reqs[0]["T"], MEAS, reqs[0]["kappa"], reqs[0]["dirs"])
        with pytest.raises(BaemmError):
            srv.wait(rid)


def test_c5_mix_with_large_meshes_alone_and_in_the_mix():
    """Serving at the sizes of the request stream: perturbed icospheres of 320
    to 327,680 triangles, kappa up to 32 (at least 10 points per wavelength),
    1 to 64 columns, GMRES to 1e-6, mixed and FP32 requests.  Every request
    completes; each result is bit-identical run alone and in the mix of four
    concurrent slots; the two small requests match the FP64 oracle's dense
    solves (FarField, Derivative_FF_Helper, AdjointDerivative_FF)."""
    import oracle

    rng = np.random.default_rng(55)
    specs = [  # level, kappa, wcs, precision
        (2, [2.0], 64, "mixed"),
        (3, [2.0, 4.0], 2, "fp32"),
        (4, [8.0], 1, "mixed"),
        (6, [16.0], 8, "mixed"),
        (7, [32.0], 2, "mixed"),
    ]
    reqs = []
    for k, (level, kappa, wcs, prec) in enumerate(specs):
        V, T = perturbed_sphere(level, seed=300 + k)
        dirs = random_directions(wcs, seed=400 + k)
        W = len(kappa) * wcs
        fobs = 0.1 * (rng.standard_normal((MEAS.shape[0], W)) + 1j * rng.standard_normal((MEAS.shape[0], W)))
        reqs.append(dict(V=V, T=T, kappa=kappa, dirs=dirs, fobs=fobs, mesh_id=5000 + k, prec=prec))
    assert [r["T"].shape[0] for r in reqs] == [320, 1280, 5120, 81920, 327680]
    alone = []
    with GNServer(slots=1) as srv:
        for r in reqs:
            alone.extend(_run(srv, [r]))
    with GNServer(slots=4) as srv:
        mixed = _run(srv, reqs[::-1])[::-1]   # the largest first, so the small ones run beside it
        st = srv.stats()
    print("C5 mix stats", st)
    assert st["completed"] == len(reqs) and st["failed"] == 0 and st["max_in_flight"] >= 2
    for k, (a, b) in enumerate(zip(alone, mixed)):
        assert a[0] == 0 and b[0] == 0, (k, a[0], b[0])
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), k
    for k in (0, 1):
        r = reqs[k]
        m = oracle.Mesh(r["V"], r["T"], MEAS)
        wcs = r["dirs"].shape[0]
        F, _ = oracle.far_field(m, r["kappa"], r["dirs"], wcs)
        du = oracle.derivative_ff_helper(m, r["kappa"], r["dirs"], wcs)
        G, _ = oracle.adjoint_derivative_ff(m, r["kappa"], r["dirs"], wcs, F - r["fobs"], du)
        eF, eG = relerr(alone[k][1], F), relerr(alone[k][2], G)
        print(f"C5 request {k} ({r['T'].shape[0]} triangles, {r['prec']}): far field vs FP64 {eF:.2e}, "
              f"gradient vs FP64 {eG:.2e}")
        assert eF < 1e-5 and eG < 2e-5, k   # measured 1.4e-7 to 1.2e-6


def test_fused_batches_match_unfused_and_do_not_depend_on_the_batch(reqs):
    """Fused mode (baemm_server_set_fusion): queued requests run as batches whose
    solves are batched GMRES.  Each request's result is bit-identical whether
    its batch holds it alone (one slot) or four requests, and agrees with the
    unfused loop within the GMRES tolerance."""
    fused_reqs = [r for r in reqs if r["prec"] in ("fp32", "mixed")]
    with GNServer(slots=4) as srv:
        plain = _run(srv, fused_reqs)
    with GNServer(slots=1, fusion=True) as srv:
        alone = _run(srv, fused_reqs)
    with GNServer(slots=4, fusion=True) as srv:
        together = _run(srv, fused_reqs)
        st = srv.stats()
    print("fused stats", st)
    assert st["completed"] == len(fused_reqs) and st["failed"] == 0 and st["max_in_flight"] >= 2
    for k in range(len(fused_reqs)):
        assert alone[k][0] == 0 and together[k][0] == 0, k
        assert np.array_equal(alone[k][1], together[k][1]) and np.array_equal(alone[k][2], together[k][2]), k
        eF, eG = relerr(together[k][1], plain[k][1]), relerr(together[k][2], plain[k][2])
        print(f"request {k}: fused vs unfused far field {eF:.2e}, gradient {eG:.2e}")
        assert eF < 1e-5 and eG < 1e-4, k
