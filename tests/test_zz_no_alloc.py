"""mpic_probe_sample_at, the probe sample that allocates nothing, against
mpic_set_probes + mpic_probe_sample on the same fields: two calls in a row,
the second with fewer probes, in both precisions, bit for bit (the same
kernel reads the same fields; only where the nodes and samples live
differs).  The scratch keeps the first call's contents, so the second call
also shows that nothing stale comes back."""
import numpy as np
import pytest

import mandor3_amd as ma
from oracle_lib import load_fixture

pytestmark = pytest.mark.gpu


def _run(prec):
    fx = load_fixture("laser_mur")
    r = ma.Run(fx["n"], fx["h"], fx["L"], fx["tau"][0], fx["bc"], 0.0, precision=prec,
               max_markers=fx["M0"].size)
    r.set_fields(fx["E0"], fx["H0"])
    r.load_markers(fx["M0"])
    r.add_laser(fx["laser_params"], fx["laser_map"])
    r.init()
    return r, tuple(int(v) for v in fx["n"])


def _nodes(rng, n, count):
    return np.stack([rng.integers(0, n[a], count) for a in range(3)], -1).astype(np.int32)


@pytest.mark.parametrize("prec", [ma.FP64, ma.FP32_SPLIT])
def test_probe_sample_at_equals_set_probes_and_sample(prec):
    r, n = _run(prec)
    rng = np.random.default_rng(7 + prec)
    for _ in range(6):
        r.step(diag=False)
    first = _nodes(rng, n, 3000)
    r.set_probes(first)
    want = r.probe_sample()
    got = r.probe_sample_at(first)
    assert got.shape == (3000, 6)
    assert np.all(np.isfinite(got)) and np.abs(got).max() > 0
    assert np.array_equal(got, want)
    for _ in range(3):
        r.step(diag=False)
    second = _nodes(rng, n, 417)
    got2 = r.probe_sample_at(second)
    r.set_probes(second)
    want2 = r.probe_sample()
    assert got2.shape == (417, 6)
    assert np.array_equal(got2, want2)
    assert not np.array_equal(got2, got[:417])
    # the run's own probe set is untouched by the no-allocation call
    r.set_probes(first)
    before = r.probe_sample()
    r.probe_sample_at(second)
    assert np.array_equal(r.probe_sample(), before)
    r.close()


def test_probe_sample_at_refuses_what_the_scratch_cannot_hold():
    r, n = _run(ma.FP64)
    too_many = np.zeros((ma.PROBE_AT_MAX + 1, 3), dtype=np.int32)
    with pytest.raises(ma.MpicError, match="scratch limit"):
        r.probe_sample_at(too_many)
    outside = np.array([[0, 0, 0], [n[0], 0, 0]], dtype=np.int32)
    with pytest.raises(ma.MpicError, match="outside"):
        r.probe_sample_at(outside)
    # the largest admitted call works and matches the allocating path
    full = _nodes(np.random.default_rng(3), n, ma.PROBE_AT_MAX)
    got = r.probe_sample_at(full)
    r.set_probes(full)
    assert np.array_equal(got, r.probe_sample())
    assert r.probe_sample_at(np.zeros((0, 3), dtype=np.int32)).shape == (0, 6)
    r.close()
