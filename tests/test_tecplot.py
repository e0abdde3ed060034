"""Tecplot records and probes against a reference core.out run.

tests/golden/tecplot_periodic.npz (tests/golden/make_tecplot_fixture.py)
holds what the reference wrote for the 'periodic' deck over 20 steps with a
Tecplot record every 5 steps, a checkpoint every 10 and 17 probes.

CPU: the oracle, started from the same record 0, pins when each record is
taken and what it holds: tec_E of the step that starts at t is the node
mean of E at t (before the push), tec_J and tec_rho are J after the fold
and the charge density after the push, the E half of every probe sample is
E at the start of the step, and the probe times are the step start times —
all bit for bit.  run_probes.cfg parsing gives the reference's probe nodes.

GPU: mandor_core, started from that record 0, writes every Tecplot payload,
probe sample and descriptor; payloads match the reference to the tolerance
of the GPU's J summation order, descriptors byte for byte.
"""
import ctypes
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import records
from oracle_lib import Oracle
from test_host import SETUP, host  # noqa: F401  (fixture)

HERE = os.path.dirname(os.path.abspath(__file__))
FX = os.path.join(HERE, "golden", "tecplot_periodic.npz")
CORE = os.path.join(os.path.dirname(HERE), "mandor3_amd", "mandor_core")
N = (8, 7, 10)
STEPS, TEC = 20, 5


def node_E(E, n):
    """Mean of the two values along each component's axis on nodes [0, n]
    (E given on nodes [0, n] as 3 interleaved doubles; the +1 neighbour of
    node n comes from the ghost layer, so the test keeps interior values)."""
    F = E.reshape(n[0] + 1, n[1] + 1, n[2] + 1, 3)
    out = np.full(F.shape, np.nan)
    out[:-1, :, :, 0] = 0.5 * (F[:-1, :, :, 0] + F[1:, :, :, 0])
    out[:, :-1, :, 1] = 0.5 * (F[:, :-1, :, 1] + F[:, 1:, :, 1])
    out[:, :, :-1, 2] = 0.5 * (F[:, :, :-1, 2] + F[:, :, 1:, 2])
    return out


def node_J(J, n):
    """The same on the J block of nodes [-2, n+2]: every node of [0, n]."""
    F = J.reshape(n[0] + 5, n[1] + 5, n[2] + 5, 3)
    s = slice(2, -2)
    out = np.empty((n[0] + 1, n[1] + 1, n[2] + 1, 3))
    out[..., 0] = 0.5 * (F[2:-2, s, s, 0] + F[3:-1, s, s, 0])
    out[..., 1] = 0.5 * (F[s, 2:-2, s, 1] + F[s, 3:-1, s, 1])
    out[..., 2] = 0.5 * (F[s, s, 2:-2, 2] + F[s, s, 3:-1, 2])
    return out


def probe_nodes(fx):
    return [tuple(int(v) for v in str(name)[6:-4].split("_")) for name in fx["probe_names"]]


def probe_samples(fx, name):
    raw = fx["file/" + name].tobytes()
    assert raw[:6] == b"probes"
    return np.frombuffer(raw[20:], "<f8").reshape(-1, 7)


def start_oracle(fx):
    """The oracle on the periodic box (domain of tests/golden/periodic.npz,
    the same deck) from the reference run's record 0: zero fields."""
    from oracle_lib import load_fixture
    base = load_fixture("periodic")
    o = Oracle(base["n"], base["h"], base["L"], base["tau"][0], base["bc"], 0.0)
    o.set_fields(np.zeros_like(base["E0"]), np.zeros_like(base["H0"]))
    o.set_markers(fx["M0"])
    o.init()
    return o


def test_oracle_pins_records_and_probe_timing():
    fx = np.load(FX)
    o = start_oracle(fx)
    nodes = probe_nodes(fx)
    samples = {str(nm): probe_samples(fx, str(nm)) for nm in fx["probe_names"]}
    n = N
    for t in range(1, STEPS + 1):
        t0 = o.time
        E, _ = o.fields()
        nE = node_E(E, n)
        for nm, (i, j, k) in zip(fx["probe_names"], nodes):
            row = samples[str(nm)][t - 1]
            assert row[0] == t0
            np.testing.assert_array_equal(row[1:4], nE[i, j, k])
        o.step()
        if t % TEC == 0:
            r = t // TEC
            tE = fx["tec_E_%d" % r].reshape(nE.shape)
            ok = ~np.isnan(nE)
            np.testing.assert_array_equal(tE[ok], nE[ok])
            np.testing.assert_array_equal(fx["tec_J_%d" % r].reshape(nE.shape), node_J(o.J(), n))
            rho = o.rho().reshape(n[0] + 5, n[1] + 5, n[2] + 5)[2:-2, 2:-2, 2:-2]
            np.testing.assert_array_equal(fx["tec_rho_%d" % r], rho.reshape(-1))
            assert struct.unpack("<dii", fx["tec_txt_%d" % r].tobytes())[0] == t0


def test_probe_nodes_from_run_probes_cfg(host, tmp_path):
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_fixtures
    cfg = make_fixtures.CASES["periodic"]
    wd = str(tmp_path)
    open(os.path.join(wd, "case.input"), "w").write(make_fixtures.deck(cfg["n"], cfg["L"], cfg["bc"],
                                                                       cfg["species"]))
    subprocess.run([SETUP, "case.input"], cwd=wd, check=True, stdout=subprocess.PIPE)
    fx = np.load(FX)
    open(os.path.join(wd, "run_probes.cfg"), "w").write(str(fx["probes_cfg"]))

    class Dom(ctypes.Structure):
        _fields_ = [("h", ctypes.c_double * 3), ("L", ctypes.c_double * 3), ("tau", ctypes.c_double),
                    ("i", ctypes.c_int * 12)]

    class Units(ctypes.Structure):
        _fields_ = [("v%d" % q, ctypes.c_double) for q in range(15)]

    d, u = Dom(), Units()
    b = os.path.join(wd, "binData").encode()
    assert host.mh_domain_load(b + b"/domain.bin", ctypes.byref(d)) == 0
    assert host.mh_units_load(b + b"/units.bin", ctypes.byref(u)) == 0
    ijk = ctypes.POINTER(ctypes.c_int)()
    n = ctypes.c_int()
    assert host.mh_probes_parse(os.path.join(wd, "run_probes.cfg").encode(), ctypes.byref(d),
                                ctypes.byref(u), ctypes.byref(ijk), ctypes.byref(n)) == 0
    got = sorted(tuple(ijk[3 * p + a] for a in range(3)) for p in range(n.value))
    assert got == sorted(probe_nodes(fx))
    assert len(set(got)) == len(got)


@pytest.mark.gpu
def test_core_writes_reference_tecplot_and_probes(host, tmp_path):  # noqa: F811
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_fixtures
    fx = np.load(FX)
    cfg = make_fixtures.CASES["periodic"]
    wd = str(tmp_path / "run")
    os.makedirs(wd)
    open(os.path.join(wd, "case.input"), "w").write(make_fixtures.deck(cfg["n"], cfg["L"], cfg["bc"],
                                                                       cfg["species"]))
    subprocess.run([SETUP, "case.input"], cwd=wd, check=True, stdout=subprocess.PIPE)
    # the reference run's own record 0 (its momenta were drawn by rand())
    M0 = fx["M0"]
    with open(os.path.join(wd, "binData", "plasma_000000_000.bin"), "wb") as f:
        f.write(np.array([M0.size], "<i8").tobytes())
        f.write(M0.tobytes())
    open(os.path.join(wd, "run_mandor.cfg"), "w").write(str(fx["run_cfg"]))
    open(os.path.join(wd, "run_probes.cfg"), "w").write(str(fx["probes_cfg"]))
    print("mandor_core tecplot/probes in", wd, flush=True)
    r = subprocess.run([CORE], cwd=wd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)
    assert r.returncode == 0, r.stdout.decode()
    b = os.path.join(wd, "binData")

    def rel(a, ref):
        return np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-300)

    for rr in range(1, STEPS // TEC + 1):
        for q in ("E", "H", "J", "rho"):
            raw = open(os.path.join(b, "tec_%s_%06d(000).bin" % (q, rr)), "rb").read()
            hdr = struct.unpack("<6i", raw[:24])
            assert hdr == (0, 0, 0) + N
            got = np.frombuffer(raw[24 + 88:], "<f8")
            ref = fx["tec_%s_%d" % (q, rr)]
            assert got.size == ref.size
            assert rel(got, ref) <= 1e-10, (q, rr, rel(got, ref))
        assert open(os.path.join(b, "tec_%06d.txt" % rr), "rb").read() == fx["tec_txt_%d" % rr].tobytes()
    for k in ("sys_000001.txt", "sys_000002.txt", "fileMap.N", "tecplot.N", "sys.N", "probe_globals.bin"):
        assert open(os.path.join(b, k), "rb").read() == fx["file/" + k].tobytes(), k
    names = sorted(f for f in os.listdir(b) if f.startswith("probe_") and f != "probe_globals.bin")
    assert names == list(fx["probe_names"])
    for nm in names:
        raw = open(os.path.join(b, nm), "rb").read()
        ref = fx["file/" + nm].tobytes()
        assert raw[:20] == ref[:20]
        got = np.frombuffer(raw[20:], "<f8").reshape(-1, 7)
        want = np.frombuffer(ref[20:], "<f8").reshape(-1, 7)
        assert got.shape == want.shape == (STEPS, 7)
        np.testing.assert_array_equal(got[:, 0], want[:, 0])
        assert rel(got[:, 1:], want[:, 1:]) <= 1e-10, nm
    assert open(os.path.join(wd, "output", "probe_locations.dat")).read() == str(fx["probe_locations"])
