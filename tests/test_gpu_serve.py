"""The serving loop: requests arriving while others run are admitted
against the HBM budget, advanced together, and each result equals the
result of the same request run alone (within the stated tolerance of the
form: the J sum and the order of markers inside a cell are the only
differences)."""
import numpy as np
import pytest

import mandor3_amd as ma
from mandor3_amd import serve, workloads
from test_gpu_parity import rel_l2

pytestmark = pytest.mark.gpu


def requests(n, steps=(20, 40), seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for rid in range(n):
        p = workloads.scan_request(rng, 0.0, steps)
        p["n"] = [(32, 32, 32), (64, 32, 32)][rid % 2]
        out.append(serve.Request(rid, p, 0.02 * rid))
    return out


@pytest.mark.parametrize("prec,tol", [(ma.FP32_SPLIT, 2e-5), (ma.FP64, 1e-12)])
def test_results_same_alone_and_in_the_mix(prec, tol):
    reqs = requests(6)
    srv = serve.Server(prec, budget_gb=100.0, slice_steps=7, full_results=True)
    for r in reqs:
        srv.submit(r)
    rep = srv.serve()
    assert rep["requests"] == 6 and rep["rejected"] == 0
    assert rep["p50_s"] is not None and rep["p99_s"] >= rep["p50_s"]
    for req in reqs[:3]:
        state = workloads.foil_request(**req.params)
        alone = serve.start_run(state, prec, graphs=False)
        alone.advance(req.params["steps"])
        alone.sync()
        E, H = alone.fields()
        assert rel_l2(req.result["E"], E) <= tol and rel_l2(req.result["H"], H) <= tol
        assert req.result["markers"] == alone.marker_count()
        assert req.result["time"] == pytest.approx(alone.time, rel=1e-14)
        alone.close()


def test_admission_respects_budget():
    reqs = requests(5, steps=(10, 20), seed=9)
    states = [workloads.foil_request(**r.params) for r in reqs]
    sizes = [serve.footprint(s, ma.FP32_SPLIT) for s in states]
    biggest = max(sizes)
    budget = 1.5 * biggest / serve.GB
    srv = serve.Server(ma.FP32_SPLIT, budget_gb=budget, slice_steps=5)
    for r in reqs:
        srv.submit(r)
    rep = srv.serve()
    assert rep["requests"] == 5 and rep["rejected"] == 0
    assert rep["peak_resident_gb"] <= budget + 1e-9
    # an oversize request is answered with an error instead of waiting
    big = serve.Request(99, dict(reqs[int(np.argmax(sizes))].params), 0.0)
    srv2 = serve.Server(ma.FP32_SPLIT, budget_gb=0.5 * biggest / serve.GB)
    srv2.submit(big)
    rep2 = srv2.serve()
    assert rep2["rejected"] == 1 and "error" in big.result


@pytest.mark.parametrize("prec,tol", [(ma.FP32_SPLIT, 2e-5), (ma.FP64, 1e-12)])
def test_stream_with_a_large_request_same_alone_and_in_the_mix(prec, tol):
    """The serving configuration's mix: Poisson arrivals of scan requests and
    one 256x128x128 request at 16 markers per cell (the large tenth of the
    stream), admitted against the 240 GB budget while the others run; the
    large request and two small ones give the result they give alone."""
    rng = np.random.default_rng(11)
    reqs, t = [], 0.0
    for rid in range(6):
        t += rng.exponential(0.25)
        p = workloads.scan_request(rng, 1.0 if rid == 2 else 0.0, (20, 40))
        reqs.append(serve.Request(rid, p, t))
    assert tuple(reqs[2].params["n"]) == (256, 128, 128) and reqs[2].params["ppc"] == 16
    srv = serve.Server(prec, budget_gb=240.0, slice_steps=7, full_results=True)
    for r in reqs:
        srv.submit(r)
    rep = srv.serve()
    assert rep["requests"] == 6 and rep["rejected"] == 0
    print("serve mix with a large request: p50 %.2f s p99 %.2f s, %.3e marker-pushes/s"
          % (rep["p50_s"], rep["p99_s"], rep["marker_pushes_per_s"]))
    for req in (reqs[2], reqs[0], reqs[5]):
        state = workloads.foil_request(**req.params)
        alone = serve.start_run(state, prec, graphs=False)
        alone.advance(req.params["steps"])
        alone.sync()
        E, H = alone.fields()
        assert rel_l2(req.result["E"], E) <= tol and rel_l2(req.result["H"], H) <= tol
        assert req.result["markers"] == alone.marker_count()
        assert req.result["time"] == pytest.approx(alone.time, rel=1e-14)
        alone.close()
