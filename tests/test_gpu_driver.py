"""The C drivers end to end on the GPU: mandor_setup writes record 0,
mandor_core runs the main loop through libmpic, and the records it writes
are compared with the oracle started from the same record 0.

Also: resuming with the continue flag reproduces the uninterrupted run,
tmp/stop.flag stops the run at the next check with a final checkpoint
(source/core/main.c:121-189,238-257,444-490).
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import records
from oracle_lib import Oracle
from test_host import DECK, SETUP, host  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CORE = os.path.join(ROOT, "mandor3_amd", "mandor_core")
N = (8, 7, 10)


def cfg(steps, chp, cont=0, stop=1000):
    return "@ %d steps\n@ %d checkpoint\n@ -100 tecplot\n@ %d continue\n@ %d stop check\n@ -100 spectr\n" % (
        steps, chp, cont, stop)


def setup_dir(path, deck):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "case.input"), "w") as f:
        f.write(deck)
    r = subprocess.run([SETUP, "case.input"], cwd=path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    assert r.returncode == 0, r.stdout.decode()


def core(path, *args):
    print("mandor_core in", path, args, flush=True)
    r = subprocess.run([CORE] + list(args), cwd=path, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=120)
    out = r.stdout.decode()
    assert r.returncode == 0, out
    return out


def oracle_from_record0(path, n, laser=None):
    d = records.domain(os.path.join(path, "binData", "domain.bin"))
    r0 = records.record(path, 0, n)
    bc = list(d["bc_min"]) + list(d["bc_max"])
    o = Oracle(n, d["h"], d["L"], d["tau"], bc, r0["time"])
    o.set_fields(r0["E"], r0["H"])
    o.set_markers(r0["M"])
    if laser is not None:
        o.add_laser(*laser)
    o.init()
    return o, d


def by_id(M):
    return M[np.argsort(M["id"], kind="stable")]


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def compare(rec, o, n, h):
    E, H = o.fields()
    assert rel_l2(rec["E"], E) <= 1e-10
    assert rel_l2(rec["H"], H) <= 1e-10
    Mg, Mo = by_id(rec["M"]), by_id(o.markers())
    assert np.array_equal(Mg["id"], Mo["id"])
    for a, k in enumerate("xyz"):
        assert np.array_equal(np.floor(Mg[k] / h[a] + 128.0) - 128, np.floor(Mo[k] / h[a] + 128.0) - 128)
        assert np.max(np.abs(Mg[k] - Mo[k])) <= 1e-10 * h[a]


def test_core_matches_oracle(tmp_path):
    wd = str(tmp_path / "run")
    setup_dir(wd, DECK)
    open(os.path.join(wd, "run_mandor.cfg"), "w").write(cfg(20, 10))
    o, d = oracle_from_record0(wd, N)
    core(wd)
    diag = []
    for s in range(20):
        dg, _ = o.step()
        diag.append(dg)
        if s + 1 in (10, 20):
            rec = records.record(wd, (s + 1) // 10, N)
            assert rec["time"] == pytest.approx(o.time, rel=1e-13)
            compare(rec, o, N, d["h"])
    rows = records.diag(os.path.join(wd, "binData", "diag.dat"))
    ref = np.array(diag)[1:]                     # t = 0 is not written
    assert rows.shape == (19, 10)
    assert np.allclose(rows[:, :6], ref, rtol=2e-6, atol=1e-300)
    assert not os.path.exists(os.path.join(wd, "tmp", "stop.flag"))


def test_continue_reproduces_uninterrupted_run(tmp_path):
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    setup_dir(a, DECK)
    shutil.copytree(a, b)
    open(os.path.join(a, "run_mandor.cfg"), "w").write(cfg(20, 10))
    core(a)
    open(os.path.join(b, "run_mandor.cfg"), "w").write(cfg(10, 10))
    core(b)
    open(os.path.join(b, "run_mandor.cfg"), "w").write(cfg(10, 10, cont=1))
    core(b)
    ra, rb = records.record(a, 2, N), records.record(b, 2, N)
    assert rb["time"] == pytest.approx(ra["time"], rel=1e-14)
    assert rel_l2(rb["E"], ra["E"]) <= 1e-10 and rel_l2(rb["H"], ra["H"]) <= 1e-10
    Ma, Mb = by_id(ra["M"]), by_id(rb["M"])
    assert np.array_equal(Ma["id"], Mb["id"])
    assert np.max(np.abs(Ma["x"] - Mb["x"])) <= 1e-12
    # diag.dat of the resumed run continues the first one
    assert records.diag(os.path.join(b, "binData", "diag.dat")).shape == (19, 10)


def test_stop_flag(tmp_path):
    wd = str(tmp_path / "run")
    setup_dir(wd, DECK)
    open(os.path.join(wd, "run_mandor.cfg"), "w").write(cfg(20, 100, stop=5))
    open(os.path.join(wd, "tmp", "stop.flag"), "w").write("1 0\n")
    core(wd)
    d = records.domain(os.path.join(wd, "binData", "domain.bin"))
    assert "@ 2" in open(os.path.join(wd, "binData", "sys.N")).read()
    rec = records.record(wd, 1, N)
    assert rec["time"] == pytest.approx(5 * d["tau"], rel=1e-13)


MUR_DECK = """[units]
@ 1.0   lambda [micron]
[mesh]
@ 24    imax
@ 16    jmax
@ 16    kmax
@ 1.2   Lx
@ 0.8   Ly
@ 0.8   Lz
@ -0.45 tau
[boundaryConditions]
@ 2 xMin
@ 2 xMax
@ 2 yMin
@ 2 yMax
@ 2 zMin
@ 2 zMax
[plasma]
@ 2 nx
@ 2 ny
@ 1 nz
@ 1 layers
@ boxes  a slab 0.8 <= x <= 0.95
> 0.875 0.4 0.4  centre
> 0.075 0 0  half-axis
> 0 1 0  half-axis
> 0 0 1  half-axis
[DF:uniform]
@ -3.0 rho
@ -1.0 q/M
@ 0 ux
@ 0 uy
@ 0 uz
@ 0.02 dux
@ 0.02 duy
@ 0.02 duz
[laser]
@ 1.5 a0
@ 0.25 waist
@ 0.7 xf
@ 0.4 yf
@ 0.4 zf
@ 4.0 fwhm
@ 3.0 T0
@ 0.0 phi0
"""


def test_core_mur_laser_matches_oracle(tmp_path, host):  # noqa: F811
    wd = str(tmp_path / "run")
    setup_dir(wd, MUR_DECK)
    open(os.path.join(wd, "run_mandor.cfg"), "w").write(cfg(20, 20))
    n = (24, 16, 16)
    # the laser as mandor_core reads it (mh_laser_load, pinned in test_host)
    from test_host import Ranmar  # noqa: F401
    raw = open(os.path.join(wd, ".EM_sources/laser/mirror_00"), "rb").read()
    assert raw.startswith(b"Mirror focused laser beam #0")

    class Laser(ctypes.Structure):
        _fields_ = [("A", ctypes.c_double), ("T0", ctypes.c_double), ("delay", ctypes.c_double),
                    ("dt", ctypes.c_double), ("phi0", ctypes.c_double), ("forward", ctypes.c_int),
                    ("plane", ctypes.c_int), ("nodes", ctypes.c_long),
                    ("map", ctypes.POINTER(ctypes.c_double))]

    class Domain(ctypes.Structure):
        _fields_ = [("h", ctypes.c_double * 3), ("L", ctypes.c_double * 3), ("tau", ctypes.c_double),
                    ("i", ctypes.c_int * 12)]

    class Units(ctypes.Structure):
        _fields_ = [("v%d" % q, ctypes.c_double) for q in range(15)]

    dd, uu, ll = Domain(), Units(), Laser()
    b = os.path.join(wd, "binData").encode()
    assert host.mh_domain_load(b + b"/domain.bin", ctypes.byref(dd)) == 0
    assert host.mh_units_load(b + b"/units.bin", ctypes.byref(uu)) == 0
    assert host.mh_laser_load(os.path.join(wd, ".EM_sources/laser/mirror_00").encode(),
                              ctypes.byref(dd), ctypes.byref(uu), ctypes.byref(ll)) == 0
    lmap = np.ctypeslib.as_array(ll.map, shape=(8 * ll.nodes,)).copy()
    params = (ll.A, ll.T0, ll.delay, ll.dt, ll.phi0, ll.forward, ll.plane)
    host.mh_laser_free(ctypes.byref(ll))
    o, d = oracle_from_record0(wd, n, laser=(params, lmap))
    core(wd)
    for _ in range(20):
        o.step()
    rec = records.record(wd, 1, n)
    compare(rec, o, n, d["h"])
    # the laser is in the box: the plasma alone (|rho| = 3, u ~ 0.02) leaves
    # fields of order 1e-5 after 20 steps, the pulse's leading edge ~0.5
    assert np.abs(rec["E"]).max() > 1e-2


def test_core_tfsf_play_matches_oracle(tmp_path):
    """mandor_core plays a TFSF recording from .EM_sources/TFSF (forward,
    15 frames, so the interface switches off within the 20 steps) and its
    records match the oracle playing the same frames."""
    import tfsf_files
    wd = str(tmp_path / "run")
    setup_dir(wd, MUR_DECK.split("[laser]")[0])
    open(os.path.join(wd, "run_mandor.cfg"), "w").write(cfg(20, 20))
    n = (24, 16, 16)
    d = records.domain(os.path.join(wd, "binData", "domain.bin"))
    reg, nf = (7, 16, 6, 10, 6, 10), 15
    dataE, dataH = tfsf_files.write_dir(os.path.join(wd, ".EM_sources", "TFSF"), reg, 0, nf,
                                        d["h"], d["tau"], seed=31, scale=0.5)
    r0 = records.record(wd, 0, n)
    bc = list(d["bc_min"]) + list(d["bc_max"])
    o = Oracle(n, d["h"], d["L"], d["tau"], bc, r0["time"])
    o.set_fields(r0["E"], r0["H"])
    o.set_markers(r0["M"])
    o.add_tfsf(reg, 0, nf, dataE, dataH)
    o.init()
    core(wd)
    for _ in range(20):
        o.step()
    rec = records.record(wd, 1, n)
    compare(rec, o, n, d["h"])
    assert np.abs(rec["H"]).max() > 1e-2          # the interface's fields are in the box


TAG_DECK = MUR_DECK.split("[plasma]")[0] + """[DF:foil]
@ 0.8    x0 [micron]
@ 0.4    y0
@ 0.4    z0
@ 1      lx
@ 0.3    ly
@ 0      lz
@ 0.15   width [micron]
> 0.02   shift
@ exponential  profile
@ -2     n1 [n_c]
@ -8     n2 [n_c]
@ -1     q
@ 1      M
@ 2      nx
@ 1      ny
@ 2      nz
[gradient]
> 0.7 0 0   point 1
> 0.9 0 0   point 2
> 1.0 0.5 0 linear
[plasma]
@ 1      nx
@ 1      ny
@ 1      nz
@ 4      layers
@ boxes  a slab
> 0.8 0.4 0.4   centre
> 0.1 0 0       half-axis
> 0 1 0         half-axis
> 0 0 1         half-axis
[DF:Maxwell]
@ 0.005  omega_pe^2
@ 5.4466e-4  q/M
@ 0.001  ux
@ 0.001  uy
@ 0.001  uz
@ 1      tune energy
@ 1      uniform weight
[gaussSpot]
@ -1.0   a0
@ 1.0    omega
@ 0.6    X [micron]
@ 0.4    Y
@ 0.4    Z
@ 0.2    width [micron]
@ 3.0    duration [fs]
@ 2.0    front offset [fs]
@ 1      Ey
@ 0      Ez
"""


def test_core_setup_tags_match_oracle(tmp_path, host):  # noqa: F811
    """A deck of the reference's deterministic tags ([DF:foil] with an
    exponential profile, [gradient], a [plasma] slab with a Maxwell quiet
    start, a [gaussSpot] hard source) set up by mandor_setup and run by
    mandor_core matches the oracle started from the same record 0."""
    wd = str(tmp_path / "run")
    setup_dir(wd, TAG_DECK)
    open(os.path.join(wd, "run_mandor.cfg"), "w").write(cfg(20, 20))
    n = (24, 16, 16)
    r0 = records.record(wd, 0, n)
    M = r0["M"]
    assert (M["qDivM"] == np.float32(-1.0)).sum() > 0 and (M["qDivM"] == np.float32(5.4466e-4)).sum() > 0
    spot = [float(ln.split()[1]) for ln in open(os.path.join(wd, ".gaussSpot.cfg")) if ln.startswith("@")]
    assert len(spot) == 11
    d = records.domain(os.path.join(wd, "binData", "domain.bin"))
    bc = list(d["bc_min"]) + list(d["bc_max"])
    o = Oracle(n, d["h"], d["L"], d["tau"], bc, r0["time"])
    o.set_fields(r0["E"], r0["H"])
    o.set_markers(M)
    o.add_gauss_spot(spot)
    o.init()
    core(wd)
    for _ in range(20):
        o.step()
    rec = records.record(wd, 1, n)
    compare(rec, o, n, d["h"])
    assert np.abs(rec["E"]).max() > 1e-2          # the spot's field is in the box
