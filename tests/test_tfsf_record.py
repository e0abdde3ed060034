"""TFSF record mode and TFSFdefrag against the reference.

tests/golden/tfsf_record.npz (tests/golden/make_tfsf_record_fixture.py)
holds what the reference core.out wrote while its TFSF interface recorded
20 steps of a hard Gaussian spot inside the interface box (the 24 fragment
files, TFSF_defrag.map, consistencyCheck.dat, lastFrameTime.bin) and what
the reference TFSFdefrag made of them (the 24 joined files).

CPU: the host library writes the map and consistencyCheck.dat byte for
byte, writes fragment files byte for byte from the reference's frames, and
joins the reference's fragments into the reference's joined files byte for
byte (removing fragments and map as TFSFdefrag does).

GPU: mandor_core, from the same record 0, records all 24 faces (headers byte
for byte, frames to the tolerance of the GPU's J summation order), and
'mandor_core --tfsf-defrag' joins them.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from test_host import SETUP, host  # noqa: F401  (fixture)

HERE = os.path.dirname(os.path.abspath(__file__))
FX = os.path.join(HERE, "golden", "tfsf_record.npz")
CORE = os.path.join(os.path.dirname(HERE), "mandor3_amd", "mandor_core")
REG = (7, 16, 6, 13, 6, 11)
E_NAMES = ("Ey_xMin", "Ez_xMin", "Ey_xMax", "Ez_xMax", "Ez_yMin", "Ex_yMin",
           "Ez_yMax", "Ex_yMax", "Ex_zMin", "Ey_zMin", "Ex_zMax", "Ey_zMax")
H_NAMES = ("Hz_xMin", "Hy_xMin", "Hz_xMax", "Hy_xMax", "Hx_yMin", "Hz_yMin",
           "Hx_yMax", "Hz_yMax", "Hy_zMin", "Hx_zMin", "Hy_zMax", "Hx_zMax")
FACES = E_NAMES + H_NAMES


class Dom(ctypes.Structure):
    _fields_ = [("h", ctypes.c_double * 3), ("L", ctypes.c_double * 3), ("tau", ctypes.c_double),
                ("i", ctypes.c_int * 12)]


def setup_case(wd, fx):
    os.makedirs(wd, exist_ok=True)
    open(os.path.join(wd, "case.input"), "w").write(str(fx["deck"]))
    r = subprocess.run([SETUP, "case.input"], cwd=wd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    assert r.returncode == 0, r.stdout.decode()
    return os.path.join(wd, ".EM_sources", "TFSF")


def read(path):
    return open(path, "rb").read()


def chdir_call(wd, fn, *args):
    old = os.getcwd()
    os.chdir(wd)
    try:
        return fn(*args)
    finally:
        os.chdir(old)


def test_record_begin_writes_map_and_consistency(host, tmp_path):
    fx = np.load(FX)
    wd = str(tmp_path)
    t = setup_case(wd, fx)
    d = Dom()
    assert host.mh_domain_load(os.path.join(wd, "binData", "domain.bin").encode(), ctypes.byref(d)) == 0
    reg = (ctypes.c_int * 6)(*REG)
    assert chdir_call(wd, host.mh_tfsf_record_begin, b".EM_sources/TFSF", reg, ctypes.byref(d)) == 0
    assert read(os.path.join(t, "TFSF_defrag.map")) == fx["rec/TFSF_defrag.map"].tobytes()
    assert read(os.path.join(t, "consistencyCheck.dat")) == fx["rec/consistencyCheck.dat"].tobytes()


def test_record_save_writes_reference_fragments(host, tmp_path):
    fx = np.load(FX)
    nf = int(np.frombuffer(fx["rec/lastFrameTime.bin"].tobytes(), "<i4")[0])
    data = np.concatenate([np.frombuffer(fx["rec/" + f].tobytes()[44:], "<f8") for f in FACES])
    t = str(tmp_path)
    reg = (ctypes.c_int * 6)(*REG)
    assert host.mh_tfsf_record_save(t.encode(), reg, nf, data.ctypes.data_as(ctypes.c_void_p)) == 0
    for f in FACES:
        assert read(os.path.join(t, f + "_000.bin")) == fx["rec/" + f].tobytes(), f
    assert read(os.path.join(t, "lastFrameTime.bin")) == fx["rec/lastFrameTime.bin"].tobytes()


def test_defrag_joins_reference_fragments(host, tmp_path):
    fx = np.load(FX)
    wd = str(tmp_path)
    t = setup_case(wd, fx)                       # TFSF.cfg and filter.cfg of the deck
    for f in FACES:
        open(os.path.join(t, f + "_000.bin"), "wb").write(fx["rec/" + f].tobytes())
    for f in ("TFSF_defrag.map", "consistencyCheck.dat", "lastFrameTime.bin"):
        open(os.path.join(t, f), "wb").write(fx["rec/" + f].tobytes())
    assert chdir_call(wd, host.mh_tfsf_defrag, b".EM_sources/TFSF") == 0
    for f in FACES:
        assert read(os.path.join(t, f + "_99999.bin")) == fx["join/" + f].tobytes(), f
    assert read(os.path.join(t, "lastFrameTime.bin")) == fx["join/lastFrameTime.bin"].tobytes()
    assert sorted(os.listdir(t)) == list(fx["join/left"])


@pytest.mark.gpu
def test_core_records_the_interface(host, tmp_path):  # noqa: F811
    fx = np.load(FX)
    wd = str(tmp_path / "run")
    t = setup_case(wd, fx)
    M0 = fx["M0"]
    with open(os.path.join(wd, "binData", "plasma_000000_000.bin"), "wb") as f:
        f.write(np.array([M0.size], "<i8").tobytes())
        f.write(M0.tobytes())
    open(os.path.join(wd, "run_mandor.cfg"), "w").write(str(fx["run_cfg"]))
    print("mandor_core TFSF record in", wd, flush=True)
    r = subprocess.run([CORE], cwd=wd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)
    assert r.returncode == 0, r.stdout.decode()
    for f in ("TFSF_defrag.map", "consistencyCheck.dat", "lastFrameTime.bin"):
        assert read(os.path.join(t, f)) == fx["rec/" + f].tobytes(), f

    def rel(a, ref):
        return np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-300)

    for f in FACES:
        got, ref = read(os.path.join(t, f + "_000.bin")), fx["rec/" + f].tobytes()
        assert got[:44] == ref[:44] and len(got) == len(ref), f
        a, b = np.frombuffer(got[44:], "<f8"), np.frombuffer(ref[44:], "<f8")
        assert np.abs(b).max() > 0
        assert rel(a, b) <= 1e-10, (f, rel(a, b))
    print("mandor_core --tfsf-defrag", flush=True)
    r = subprocess.run([CORE, "--tfsf-defrag"], cwd=wd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       timeout=60)
    assert r.returncode == 0, r.stdout.decode()
    assert sorted(os.listdir(t)) == list(fx["join/left"])
    for f in FACES:
        got, ref = read(os.path.join(t, f + "_99999.bin")), fx["join/" + f].tobytes()
        assert got[:44] == ref[:44] and len(got) == len(ref), f
        assert rel(np.frombuffer(got[44:], "<f8"), np.frombuffer(ref[44:], "<f8")) <= 1e-10, f
