"""Graph replay, ensembles and the autotuner against the stage-by-stage
path (which the parity tests pin to the oracle).

  - replayed graphs in vacuum with the laser are bit-exact with the oracle
    (the per-step laser table is evaluated with the same host arithmetic);
  - replay with markers equals stage calls within the J summation order;
  - runs of different sizes advanced together in one ensemble give the
    results they give alone;
  - the autotuner's choice of launch shape leaves results unchanged.
"""
import numpy as np
import pytest

import mandor3_amd as ma
from oracle_lib import load_fixture, tfsf_arrays
from test_gpu_parity import by_id, cells, oracle, random_fields, rel_l2

pytestmark = pytest.mark.gpu


def make(fx, precision, markers=True, E=None, H=None, graphs=False):
    r = ma.Run(fx["n"], fx["h"], fx["L"], fx["tau"][0], fx["bc"], 0.0, precision=precision,
               max_markers=max(fx["M0"].size, 1))
    r.set_fields(fx["E0"] if E is None else E, fx["H0"] if H is None else H)
    if markers:
        r.load_markers(fx["M0"])
    if "laser_params" in fx:
        r.add_laser(fx["laser_params"], fx["laser_map"])
    if "spot_params" in fx:
        r.add_gauss_spot(fx["spot_params"])
    if "tfsf_reg" in fx:
        r.add_tfsf(*tfsf_arrays(fx))
    if "vsp" in fx:
        r.set_vsp(float(fx["vsp"][0]))
    r.init()
    if graphs:
        r.set_graphs(True)
    return r


def test_replay_vacuum_laser_bit_exact():
    fx = load_fixture("laser_mur")
    E0, H0 = random_fields(fx, 11)
    g = make(fx, ma.FP64, markers=False, E=E0, H=H0, graphs=True)
    o = oracle(fx, markers=False, E=E0, H=H0)
    for chunk in (1, 4, 7):
        g.advance(chunk)
        for _ in range(chunk):
            o.step()
        Eg, Hg = g.fields()
        Eo, Ho = o.fields()
        assert np.array_equal(Eg, Eo) and np.array_equal(Hg, Ho), chunk
        assert g.time == o.time


@pytest.mark.parametrize("prec", [ma.FP64, ma.FP32_SPLIT])
@pytest.mark.parametrize("case", ["laser_mur", "mixed", "spot", "tfsf"])
def test_replay_equals_stage_calls(case, prec):
    fx = load_fixture(case)
    a = make(fx, prec)
    b = make(fx, prec, graphs=True)
    for _ in range(13):
        a.step(diag=False)
    b.advance(5)
    b.advance(8)
    a.sync()
    b.sync()
    tol = 1e-12 if prec == ma.FP64 else 2e-5
    Ea, Ha = a.fields()
    Eb, Hb = b.fields()
    assert rel_l2(Eb, Ea) <= tol and rel_l2(Hb, Ha) <= tol
    Ma, Mb = by_id(a.markers()), by_id(b.markers())
    assert np.array_equal(Ma["id"], Mb["id"])
    if prec == ma.FP64:
        assert np.array_equal(cells(Ma, fx), cells(Mb, fx))
    assert np.max(np.abs(Ma["x"] - Mb["x"])) <= (1e-12 if prec == ma.FP64 else 1e-5) * fx["h"][0]
    assert a.time == b.time


@pytest.fixture(params=["fused", "streams"])
def batch_mode(request, monkeypatch):
    """mpic_batch_advance's two forms: the fused ensemble (one launch per
    stage for all runs, the default) and per-run graph replays on per-run
    streams (MPIC_BATCH_STREAMS=1)."""
    if request.param == "streams":
        monkeypatch.setenv("MPIC_BATCH_STREAMS", "1")
    else:
        monkeypatch.delenv("MPIC_BATCH_STREAMS", raising=False)
    return request.param


ALL_CASES = ("periodic", "laser_mur", "mixed", "spot", "tfsf", "tfsf_back", "vsp")


@pytest.mark.parametrize("prec", [ma.FP64, ma.FP32_SPLIT])
def test_ensemble_equals_alone(prec, batch_mode):
    """Every fixture box (periodic, Mur with the laser, mirror, the hard
    spot, TFSF forward and backward, VSP: steps of different launch
    sequences) advanced together gives what it gives alone; the second run
    starts one step ahead (its markers sit in the other buffer), and the
    ensemble is advanced in two calls (the fused step is reused)."""
    fxs = [load_fixture(c) for c in ALL_CASES]
    alone = [make(fx, prec) for fx in fxs]
    mix = [make(fx, prec) for fx in fxs]
    mix[1].step(diag=False)
    for q, r in enumerate(alone):
        for _ in range(9 + (q == 1)):
            r.step(diag=False)
        r.sync()
    ma.batch_advance(mix, 4)
    ma.batch_advance(mix, 5)
    tol = 1e-12 if prec == ma.FP64 else 2e-5
    for fx, a, b in zip(fxs, alone, mix):
        b.sync()
        Ea, Ha = a.fields()
        Eb, Hb = b.fields()
        assert rel_l2(Eb, Ea) <= tol and rel_l2(Hb, Ha) <= tol
        Ma, Mb = by_id(a.markers()), by_id(b.markers())
        assert np.array_equal(Ma["id"], Mb["id"])
        assert b.time == a.time


def test_autotune_is_result_neutral():
    fx = load_fixture("laser_mur")
    a = make(fx, ma.FP32_SPLIT)
    b = make(fx, ma.FP32_SPLIT)
    blocks, waves, item = b.autotune(steps=2)
    assert blocks > 0 and waves in (1, 2, 4) and item in (0, 1024, 2048)
    done = int(round(b.time / fx["tau"][0]))
    for _ in range(done + 4):
        a.step(diag=False)
    b.set_graphs(True)
    b.advance(4)
    a.sync()
    b.sync()
    Ea, Ha = a.fields()
    Eb, Hb = b.fields()
    assert rel_l2(Eb, Ea) <= 2e-5 and rel_l2(Hb, Ha) <= 2e-5
    assert a.time == pytest.approx(b.time, rel=1e-14)


@pytest.mark.parametrize("prec", [ma.FP64, ma.FP32_SPLIT])
@pytest.mark.parametrize("graphs", [False, True])
def test_diag_history_equals_stepwise_diagnostics(prec, graphs):
    """The energy history the loop body keeps on the device (em_energy and
    plasma_temperature of every asynchronous step, mpic_diag_history) equals
    the rows mpic_step(diag) hands out step by step and, in the FP64 form,
    the oracle's: Time exactly, energies to the summation order."""
    fx = load_fixture("laser_mur")
    a = make(fx, prec)
    b = make(fx, prec, graphs=graphs)
    o = oracle(fx)
    rows_a, rows_o = [], []
    for _ in range(14):
        rows_a.append(a.step(diag=True))
        rows_o.append(o.step()[0])
    b.advance(6)
    b.step(diag=False)
    b.advance(7)
    h = b.diag_history()
    assert h.shape == (14, 6)
    ra, ro = np.array(rows_a), np.array(rows_o)
    assert np.array_equal(h[:, 0], ra[:, 0])
    tol = 1e-12 if prec == ma.FP64 else 2e-5
    np.testing.assert_allclose(h[:, 1:], ra[:, 1:], rtol=tol, atol=1e-300)
    if prec == ma.FP64:
        np.testing.assert_allclose(h[:, 1:], ro[:, 1:], rtol=1e-10, atol=1e-300)
    # staged steps keep no row; the ring keeps the last 1024 rows
    b.stage("em_h_half")
    for s in ("push_deposit", "em_h_step", "em_e_start", "j_finish", "em_e_finish"):
        b.stage(s)
    b.advance(2)
    h2 = b.diag_history(4)
    assert h2.shape == (4, 6)
    assert np.array_equal(h2[:2, 0], ra[12:14, 0]) and h2[2, 0] > ra[13, 0] + 1.5 * fx["tau"][0]


@pytest.mark.parametrize("prec", [ma.FP64, ma.FP32_SPLIT])
def test_runs_set_up_on_another_thread_while_an_ensemble_advances(prec):
    """The serving pattern: one thread creates, loads, initialises, replays
    and closes runs while another advances an ensemble whose membership
    changes every call (so its fused graphs are rebuilt meanwhile); both
    sides give what they give alone."""
    import threading
    cases = ("laser_mur", "periodic", "mixed")
    fxs = [load_fixture(c) for c in cases]
    alone = [make(fx, prec) for fx in fxs]
    mix = [make(fx, prec) for fx in fxs]
    side_fx = load_fixture("spot")
    side_ref = make(side_fx, prec)
    for _ in range(6):
        side_ref.step(diag=False)
    side_ref.sync()
    E_ref, H_ref = side_ref.fields()
    side_ref.close()
    errors, side_out = [], []

    def side():
        try:
            for _ in range(4):
                r = make(side_fx, prec, graphs=True)
                r.advance(6)
                r.sync()
                side_out.append(r.fields())
                r.close()
        except BaseException as e:          # reported by the main thread
            errors.append(e)

    t = threading.Thread(target=side)
    t.start()
    done = [0, 0, 0]
    for s in range(12):
        members = [0, 1, 2] if s % 2 == 0 else [0, 2]
        ma.batch_advance([mix[q] for q in members], 2)
        for q in members:
            done[q] += 2
        mix[0].sync()
    t.join()
    assert not errors, errors[0]
    tol = 1e-12 if prec == ma.FP64 else 2e-5
    for a, b, n in zip(alone, mix, done):
        for _ in range(n):
            a.step(diag=False)
        a.sync()
        b.sync()
        Ea, Ha = a.fields()
        Eb, Hb = b.fields()
        assert rel_l2(Eb, Ea) <= tol and rel_l2(Hb, Ha) <= tol
        assert np.array_equal(by_id(a.markers())["id"], by_id(b.markers())["id"])
        assert b.time == a.time
    assert len(side_out) == 4
    for E, H in side_out:
        assert rel_l2(E, E_ref) <= tol and rel_l2(H, H_ref) <= tol
