"""The C host side (mandor3_amd/host): RANMAR, the setup loader and the
record formats, on the CPU.

RANMAR is pinned by the known answer of source/misc_ranmar.c:36-56 (seeds
1802/9373, draws 20001-20006 times 2^24).  The setup loader is checked by
reading its files back with tests/records.py, which follows the reference
formats and reads the reference's own records in the fixture script.
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

import records

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mandor3_amd")
HOST_LIB = os.path.join(PKG, "libmandor_host.so")
SETUP = os.path.join(PKG, "mandor_setup")


@pytest.fixture(scope="module")
def host():
    if not (os.path.exists(HOST_LIB) and os.path.exists(SETUP)):
        from mandor3_amd import build
        build.build_host()
    return ctypes.CDLL(HOST_LIB)


class Ranmar(ctypes.Structure):
    _fields_ = [("u", ctypes.c_double * 97), ("c", ctypes.c_double), ("cd", ctypes.c_double),
                ("cm", ctypes.c_double), ("i97", ctypes.c_int), ("j97", ctypes.c_int)]


def ranmar_stream(host, ij, kl, n):
    r = Ranmar()
    assert host.mh_ranmar_init(ctypes.byref(r), ij, kl) == 0
    out = np.zeros(n)
    host.mh_ranmar(ctypes.byref(r), out.ctypes.data_as(ctypes.c_void_p), n)
    return out


def test_ranmar_known_answer(host):
    v = ranmar_stream(host, 1802, 9373, 20006)
    assert np.array_equal(4096.0 * 4096.0 * v[20000:],
                          [6533892.0, 14220222.0, 7275067.0, 6172232.0, 8354498.0, 10633180.0])
    assert v.min() >= 0.0 and v.max() < 1.0
    # 24-bit values: exact multiples of 2^-24
    assert np.all(v * 2.0 ** 24 == np.floor(v * 2.0 ** 24))


def test_ranmar_normal_is_box_muller_of_the_stream(host):
    """mh_ranmar_normal (the workloads' thermal momenta) is the Box-Muller
    transform of the RANMAR stream's consecutive pairs, r cos then r sin,
    and the workloads' seed pair covers the generator's seed range."""
    from mandor3_amd import rng as mrng
    u = ranmar_stream(host, 1802, 9373, 2000)
    r = Ranmar()
    assert host.mh_ranmar_init(ctypes.byref(r), 1802, 9373) == 0
    z = np.zeros(1999)
    host.mh_ranmar_normal.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                      ctypes.c_double]
    host.mh_ranmar_normal(ctypes.byref(r), z.ctypes.data_as(ctypes.c_void_p), z.size, 0.5)
    rad = 0.5 * np.sqrt(-2.0 * np.log1p(-u[0::2]))
    ang = 2.0 * np.pi * u[1::2]
    want = np.empty(2000)
    want[0::2], want[1::2] = rad * np.cos(ang), rad * np.sin(ang)
    np.testing.assert_allclose(z, want[:1999], rtol=1e-13, atol=1e-15)
    g = mrng.Ranmar(12345).normal(400000, 0.01)
    assert abs(g.mean()) < 1e-4 and abs(g.std() - 0.01) < 1e-4
    assert mrng.seed_pair(0) == (0, 0)
    assert mrng.seed_pair(31329 * 30082 - 1) == (31328, 30081)
    assert mrng.seed_pair(31329) == (0, 1)


def test_ranmar_rejects_bad_seeds(host):
    r = Ranmar()
    assert host.mh_ranmar_init(ctypes.byref(r), 31329, 0) != 0
    assert host.mh_ranmar_init(ctypes.byref(r), 0, 30082) != 0


DECK = """[units]
@ 1.0      lambda [micron]
[mesh]
@ 8        imax
@ 7        jmax
@ 10       kmax
@ 0.4      Lx [micron]
@ 0.35     Ly [micron]
@ 0.5      Lz [micron]
@ -0.45    tau (Courant fraction)
[boundaryConditions]
@ 0        xMin
@ 0        xMax
@ 0        yMin
@ 0        yMax
@ 0        zMin
@ 0        zMax
[RANMAR]
@ 1802     ij
@ 9373     kl
[plasma]
@ 2        nx
@ 1        ny
@ 2        nz
@ 1        layers
[DF:uniform]
@ -1.0     rho [n_c]
@ -1.0     q/M
@ 0.1      ux
@ 0.0      uy
@ -0.05    uz
@ 0.2      dux
@ 0.2      duy
@ 0.2      duz
[plasma]
@ 1        nx
@ 1        ny
@ 1        nz
@ 1        layers
> random   in-cell positions from RANMAR
@ boxes    a slab 0.1 <= x <= 0.3
> 0.2 0.175 0.25   centre
> 0.1 0 0          half-axis
> 0 1 0            half-axis
> 0 0 1            half-axis
[DF:thermal]
@ 1.0      rho [n_c]
@ 0.000544662   q/M
@ 0.0      ux
@ 0.0      uy
@ 0.0      uz
@ 0.01     vth
"""


def run_setup(tmp_path, deck):
    (tmp_path / "case.input").write_text(deck)
    r = subprocess.run([SETUP, "case.input"], cwd=tmp_path, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT)
    assert r.returncode == 0, r.stdout.decode()
    return tmp_path


def test_setup_record_zero(host, tmp_path):
    wd = run_setup(tmp_path, DECK)
    b = wd / "binData"
    d = records.domain(b / "domain.bin")
    n = (8, 7, 10)
    assert tuple(d["dmn_max"]) == n and tuple(d["dmn_min"]) == (0, 0, 0)
    # lengths in lambda (1 micron), h = L / n, tau = 0.45 tau_max
    assert np.allclose(d["L"], [0.4, 0.35, 0.5], rtol=1e-15)
    assert np.allclose(d["h"], 0.05, rtol=1e-14)
    assert math.isclose(d["tau"], 0.45 / math.sqrt(3 / 0.05 ** 2), rel_tol=1e-14)
    u = records.units(b / "units.bin")
    assert u[0] == pytest.approx(1e-4) and u[2] == pytest.approx(1e-4 / 2.9979e10)
    E, reg = records.mesh(b / "sys_E_000000(000).bin", n)
    assert tuple(reg) == (0, 0, 0) + n and not E.any()
    M = records.markers(b / "plasma_000000_000.bin")
    ne = 8 * 7 * 10 * 4
    ni = 4 * 7 * 10              # random points of the cells 2..5 lie in the slab
    assert M.size == ne + ni
    assert np.array_equal(M["id"], np.arange(1, M.size + 1))
    # plasma objects are written newest first (source/setup/plasma.c:93-124)
    i, e = M[:ni], M[ni:]
    # charge per marker: rho * rho_crit / ppc, rho_crit = 4 pi^2 up to rounding
    assert np.allclose(e["rho"], -4 * math.pi ** 2 / 4, rtol=1e-6)
    assert np.allclose(i["rho"], 4 * math.pi ** 2, rtol=1e-6)
    assert np.all(e["qDivM"] == np.float32(-1.0))
    # electrons on the (i + 1/2)/n lattice of their cell
    fx = e["x"] / 0.05 - np.floor(e["x"] / 0.05)
    assert np.allclose(np.sort(np.unique(np.round(fx, 12))), [0.25, 0.75])
    fy = e["y"] / 0.05 - np.floor(e["y"] / 0.05)
    assert np.allclose(fy, 0.5)
    # ions: random in-cell positions inside the slab
    assert i["x"].min() >= 0.1 - 1e-12 and i["x"].max() < 0.3
    # uniform momenta: <u> + (r - 1/2) du with r the RANMAR stream, 3 draws
    # per marker in creation order
    r = ranmar_stream(host, 1802, 9373, 3 * ne).reshape(ne, 3)
    assert np.array_equal(e["vx"], 0.1 + (r[:, 0] - 0.5) * 0.2)
    assert np.array_equal(e["vz"], -0.05 + (r[:, 2] - 0.5) * 0.2)
    # thermal ions: zero-mean Gaussian of width vth
    assert abs(i["vx"].mean()) < 0.01 * 0.2
    assert 0.8 * 0.01 < i["vx"].std() < 1.2 * 0.01
    t, cpus, fmap = records.sys_info(b / "sys_000000.txt")
    assert (t, cpus, fmap) == (0.0, 1, 1)          # file map 0 is the Tecplot record 0's
    assert "@ 1" in (b / "sys.N").read_text()
    assert (b / "fileMap_000000.txt").read_text().startswith("1 cpus total.")


def test_setup_laser_source(host, tmp_path):
    deck = DECK.split("[RANMAR]")[0].replace("@ 8        imax", "@ 24       imax").replace(
        "@ 7        jmax", "@ 16       jmax").replace("@ 10       kmax", "@ 16       kmax").replace(
        "@ 0.4      Lx", "@ 1.2      Lx").replace("@ 0.35     Ly", "@ 0.8      Ly").replace(
        "@ 0.5      Lz", "@ 0.8      Lz")
    deck += "[laser]\n@ 1.5 a0\n@ 0.25 waist\n@ 0.7 xf\n@ 0.4 yf\n@ 0.4 zf\n@ 4.0 fwhm\n@ 10.0 T0\n@ 0.3 phi0\n"
    wd = run_setup(tmp_path, deck)

    class Domain(ctypes.Structure):
        _fields_ = [("h", ctypes.c_double * 3), ("L", ctypes.c_double * 3), ("tau", ctypes.c_double),
                    ("dmn_min", ctypes.c_int * 3), ("dmn_max", ctypes.c_int * 3),
                    ("bc_min", ctypes.c_int * 3), ("bc_max", ctypes.c_int * 3)]

    class Units(ctypes.Structure):
        _fields_ = [(k, ctypes.c_double) for k in
                    ("lambda_", "v", "t", "omega", "E", "S", "A", "n", "rho", "ncr", "debye",
                     "micron", "femtosecond", "wpe", "Tpe")]

    class Laser(ctypes.Structure):
        _fields_ = [("A", ctypes.c_double), ("T0", ctypes.c_double), ("delay", ctypes.c_double),
                    ("dt", ctypes.c_double), ("phi0", ctypes.c_double), ("forward", ctypes.c_int),
                    ("plane", ctypes.c_int), ("nodes", ctypes.c_long),
                    ("map", ctypes.POINTER(ctypes.c_double))]

    d, u, l = Domain(), Units(), Laser()
    b = str(wd / "binData").encode()
    assert host.mh_domain_load(b + b"/domain.bin", ctypes.byref(d)) == 0
    assert host.mh_units_load(b + b"/units.bin", ctypes.byref(u)) == 0
    p = str(wd / ".EM_sources/laser/mirror_00").encode()
    assert host.mh_laser_load(p, ctypes.byref(d), ctypes.byref(u), ctypes.byref(l)) == 0
    t0 = 1e-4 / 2.9979e10
    # conversions of source/scpic/em_mirror.c:296-330
    assert l.A == pytest.approx(2 * math.pi * 1.5, rel=1e-15)
    assert l.delay == pytest.approx(0.7 - 5 * 0.05, rel=1e-14)
    assert l.dt == pytest.approx(4.0e-15 / t0 * math.sqrt(2 / math.log(2)), rel=1e-14)
    assert l.T0 == pytest.approx(10.0e-15 / t0, rel=1e-14)
    assert (l.forward, l.plane, l.nodes) == (1, 5, (16 + 1 - 10) ** 2)
    m = np.ctypeslib.as_array(l.map, shape=(2, l.nodes, 4))
    assert np.isfinite(m).all()
    # E map (x = 4h): peak |field| = w0/w(x) exp(-r^2/w^2) at the node
    # nearest the beam axis (y nodes at (j - 1/2)h, so r = h/2 there)
    w0, zr, h = 0.25, math.pi * 0.25 ** 2, 0.05
    w = w0 * math.sqrt(1 + ((4 * h - 0.7) / zr) ** 2)
    amp = np.hypot(m[0, :, 0], m[0, :, 1])
    assert amp.max() == pytest.approx(w0 / w * math.exp(-(h / 2) ** 2 / w ** 2), rel=1e-12)
    # y polarisation: the E map has no z component
    assert not m[0, :, 2:].any()
    host.mh_laser_free(ctypes.byref(l))


def test_diag_file_format(host, tmp_path):
    p = str(tmp_path / "diag.dat").encode()
    pts = np.array([[0.0, 1, 2, 3, 4, 5], [0.5, 1.5, 2.5, 0.1, 0.2, 0.3]], dtype=float)
    assert host.mh_diag_append(p, pts.ctypes.data_as(ctypes.c_void_p), 2,
                               ctypes.c_double(0.1), 1) == 0
    assert host.mh_diag_append(p, pts[1:].ctypes.data_as(ctypes.c_void_p), 1,
                               ctypes.c_double(0.1), 0) == 0
    text = (tmp_path / "diag.dat").read_text().splitlines()
    assert text[0].startswith("variables = t, W_E, W_M")
    rows = records.diag(tmp_path / "diag.dat")
    assert rows.shape == (2, 10)            # t = 0 < tau/2 is skipped
    assert np.allclose(rows[0], [0.5, 1.5, 2.5, 0.1, 0.2, 0.3, 0.5, 0.6, 4.0, 4.6])


def test_run_cfg_rejects_zero(host, tmp_path):
    class RunCfg(ctypes.Structure):
        _fields_ = [(k, ctypes.c_int) for k in ("a", "b", "c", "d", "e", "f")]
    good = tmp_path / "good.cfg"
    good.write_text("@ 20 steps\n@ 10 chp\n@ -100 tec\n@ 1 cont\n@ 5 stop\n@ -100 spec\n")
    rc = RunCfg()
    assert host.mh_runcfg_load(str(good).encode(), ctypes.byref(rc)) == 0
    assert (rc.a, rc.b, rc.d, rc.e) == (20, 10, 1, 5)
    bad = tmp_path / "bad.cfg"
    bad.write_text("@ 20\n@ 0\n@ -100\n@ 0\n@ 5\n@ -100\n")
    assert host.mh_runcfg_load(str(bad).encode(), ctypes.byref(rc)) != 0


class _Domain(ctypes.Structure):
    _fields_ = [("h", ctypes.c_double * 3), ("L", ctypes.c_double * 3), ("tau", ctypes.c_double),
                ("i", ctypes.c_int * 12)]


class _Tfsf(ctypes.Structure):
    _fields_ = [("reg", ctypes.c_int * 6), ("mode", ctypes.c_int), ("nframes", ctypes.c_int),
                ("vol", ctypes.c_long * 12), ("dataE", ctypes.c_void_p * 12),
                ("dataH", ctypes.c_void_p * 12)]


def _tfsf_load(host, d, dom):
    host.mh_tfsf_load.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
    t = _Tfsf()
    rc = host.mh_tfsf_load(d.encode(), ctypes.byref(dom), ctypes.byref(t))
    return rc, t


def test_tfsf_play_files(host, tmp_path):
    """mh_tfsf_load reads the play-mode directory core.out reads: the
    frames of every face slot (E faces get the H recordings and vice versa),
    opened faces dropped, consistency checks enforced."""
    import tfsf_files
    dom = _Domain()
    dom.h[:] = (0.05, 0.05, 0.05)
    dom.tau = 0.0129903810567666
    reg, nf = (7, 16, 6, 13, 6, 11), 9
    d = str(tmp_path / "TFSF")
    dataE, dataH = tfsf_files.write_dir(d, reg, 2, nf, dom.h[:], dom.tau, opened=(0, 1, 0, 0, 1, 0))
    rc, t = _tfsf_load(host, d, dom)
    assert rc == 0
    assert tuple(t.reg) == reg and t.mode == 2 and t.nframes == nf
    vols = tfsf_files.face_nodes(reg)
    assert list(t.vol) == vols
    for f in range(12):
        if dataE[f] is None:
            assert not t.dataE[f] and not t.dataH[f]
            continue
        e = np.ctypeslib.as_array((ctypes.c_double * (vols[f] * nf)).from_address(t.dataE[f]))
        h = np.ctypeslib.as_array((ctypes.c_double * (vols[f] * nf)).from_address(t.dataH[f]))
        assert np.array_equal(e, dataE[f]) and np.array_equal(h, dataH[f])
    host.mh_tfsf_free(ctypes.byref(t))
    # a different time step, an uncompiled recording and a missing frame
    dom2 = _Domain()
    dom2.h[:] = dom.h[:]
    dom2.tau = dom.tau * (1 + 1e-9)
    assert _tfsf_load(host, d, dom2)[0] != 0
    open(os.path.join(d, "TFSF_defrag.map"), "w").write("x\n")
    assert _tfsf_load(host, d, dom)[0] != 0
    os.remove(os.path.join(d, "TFSF_defrag.map"))
    np.array([nf + 1], "<i4").tofile(os.path.join(d, "lastFrameTime.bin"))
    assert _tfsf_load(host, d, dom)[0] != 0
    # no TFSF.cfg: no interface
    rc, t = _tfsf_load(host, str(tmp_path / "none"), dom)
    assert rc == 0 and t.mode == -1
