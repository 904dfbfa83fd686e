"""Continuous-batching server: every request's result is the same alone and in
the mix (bit for bit), requests admitted while others run, and the HBM
footprint budget is respected."""
import pytest

from casmcode_monte_amd import serving

pytestmark = pytest.mark.gpu

CLASSES = ((64, 3, True), (256, 2, True), (512, 2, False))


def trace():
    reqs = []
    edges = [16, 20, 64, 100, 130, 200, 24, 300, 410, 48, 256, 90]
    for i, e in enumerate(edges):
        reqs.append(serving.Request(rid=i, edge=e, temperature=800.0 + 230.0 * i,
                                    exchange_potential=-0.5 + 0.09 * i, seed=9000 + i,
                                    arrival=0.01 * i, max_passes=2500 if e < 200 else 1000,
                                    precision=2e-3))
    return reqs


def test_mix_equals_alone():
    reqs = trace()
    print("case: serving mix of", len(reqs), "requests", flush=True)
    srv = serving.Server(classes=CLASSES, chunk=100, check_period=500)
    mixed, stats = srv.serve(reqs)
    srv.close()
    assert len(mixed) == len(reqs)
    assert stats["p50_latency_s"] <= stats["p99_latency_s"]
    # some requests had to wait for a slot (3 + 2 + 2 slots for 12 requests)
    assert any(r.admitted > r.arrival + 1e-3 for r in mixed.values())
    assert any(r.converged for r in mixed.values())
    for req in reqs:
        alone_srv = serving.Server(classes=CLASSES, chunk=100, check_period=500)
        req0 = serving.Request(**{**req.__dict__, "arrival": 0.0})
        alone, _ = alone_srv.serve([req0])
        alone_srv.close()
        a, m = alone[req.rid], mixed[req.rid]
        key = lambda r: (r.passes, r.n_samples, r.converged, r.n_equilibration, r.mean_x,
                         r.precision_x, r.lattice_crc)
        assert key(a) == key(m), (req, a, m)


def test_hbm_budget_admission():
    reqs = trace()[:6]
    budget = max(r.footprint() for r in reqs) * 2
    print("case: serving budget", budget, flush=True)
    srv = serving.Server(classes=CLASSES, chunk=100, check_period=500, hbm_budget=budget)
    res, stats = srv.serve(reqs)
    srv.close()
    assert len(res) == len(reqs)
    assert 0 < stats["max_hbm_in_use"] <= budget


FCC_CLASSES = ((6, 2, True), (10, 2, False))


def test_ising_fcc_mix_equals_alone():
    """Ising and FCC requests served together; each result equals its run alone."""
    reqs = []
    for i in range(10):
        fcc = i % 2 == 1
        edge = (4 + i // 2) if fcc else 16 + 12 * i
        reqs.append(serving.Request(rid=i, edge=edge, temperature=700.0 + 150.0 * i,
                                    exchange_potential=-0.3 + 0.07 * i, seed=7700 + i,
                                    arrival=0.005 * i, max_passes=1500, precision=3e-3,
                                    model="fcc" if fcc else "ising"))
    print("case: serving ising+fcc mix", flush=True)
    srv = serving.Server(classes=CLASSES, fcc_classes=FCC_CLASSES, chunk=100, check_period=500)
    mixed, stats = srv.serve(reqs)
    srv.close()
    assert len(mixed) == len(reqs)
    for req in reqs:
        alone_srv = serving.Server(classes=CLASSES, fcc_classes=FCC_CLASSES, chunk=100,
                                   check_period=500)
        alone, _ = alone_srv.serve([serving.Request(**{**req.__dict__, "arrival": 0.0})])
        alone_srv.close()
        a, m = alone[req.rid], mixed[req.rid]
        key = lambda r: (r.passes, r.n_samples, r.converged, r.n_equilibration, r.mean_x,
                         r.precision_x, r.lattice_crc)
        assert key(a) == key(m), (req, a, m)
