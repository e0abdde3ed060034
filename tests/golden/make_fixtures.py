"""Writes the golden fixtures of the PIC time step from the reference build.

Runs in the development container only (the reference tree and its MPI
build are not on the GPU host).  For every case it

  1. runs the reference ``setup.out`` on an ``*.input`` deck (tag layout of
     source/setup/tag.c:16-54) to make binData record 0,
  2. optionally drops a focused-laser map into ``.EM_sources/laser/mirror_00``
     (layout of source/scpic/em_mirror.c:392-420; the map itself is an
     analytic Gaussian beam written by ``laser_map`` below, so the
     reference's injection algebra runs on a map we also hand the oracle),
  3. runs the reference ``core.out`` with a checkpoint every step
     (run_mandor.cfg, source/core/main.c:205-220) and the MC_GAUSS_TEST
     build for the Gauss/continuity residuals,
  4. stores record 0 in full, SHA-256 digests of E, H and the marker array of
     every later record, the final record in full, diag.dat and
     gauss_test.dat, as ``tests/golden/<case>.npz``.

The fixtures are data (inputs and expected outputs); no reference source is
copied.  Usage, from the repository root:

    make -f oracle/ref.mk && python tests/golden/make_fixtures.py
"""
import hashlib
import math
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "oracle", "_ref")
STEPS = 20

MARKER_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
                         ("vx", "<f8"), ("vy", "<f8"), ("vz", "<f8"),
                         ("rho", "<f4"), ("qDivM", "<f4"), ("id", "<i4"),
                         ("pad", "<i4")])


def deck(n, L, bc, species, tau=-0.45):
    """An *.input deck: [units] [mesh] [boundaryConditions] then plasma."""
    lines = ["[units]", "@ 1.0  lambda [micron]", "[mesh]"]
    lines += ["@ %d  cells" % v for v in n]
    lines += ["@ %.17g  L [micron]" % v for v in L]
    lines += ["@ %g  tau" % tau, "[boundaryConditions]"]
    for a in range(3):          # xMin xMax yMin yMax zMin zMax
        lines += ["@ %d  min" % bc[a], "@ %d  max" % bc[3 + a]]
    for sp in species:
        lines += ["[plasma]"] + ["@ %d  sub" % v for v in sp["sub"]]
        if "box" in sp:
            c, hw = sp["box"]
            lines += ["@ boxes",
                      "> %.17g %.17g %.17g center" % tuple(c),
                      "> %.17g 0 0  dir1" % hw[0],
                      "> 0 %.17g 0  dir2" % hw[1],
                      "> 0 0 %.17g  dir3" % hw[2]]
        lines += ["[DF:uniform]", "@ %.17g rho" % sp["rho"],
                  "@ %.17g q/M" % sp["qm"]]
        lines += ["@ %.17g  u" % v for v in sp["u"]]
        lines += ["@ %.17g  du" % v for v in sp["du"]]
    return "\n".join(lines) + "\n"


CASES = {
    # Periodic box, unequal steps per axis (h = 0.05, 0.06, 0.045), hot
    # electrons crossing up to three cell faces per step, cold ions.
    "periodic": dict(
        n=(8, 7, 10), L=(0.4, 0.42, 0.45), bc=(0, 0, 0, 0, 0, 0),
        species=[dict(sub=(2, 2, 1, 1), rho=-1.0, qm=-1.0,
                      u=(0.3, -0.2, 0.1), du=(2.0, 1.5, 1.8)),
                 dict(sub=(1, 1, 1, 2), rho=1.0, qm=5.4466e-4,
                      u=(0, 0, 0), du=(0, 0, 0))],
        laser=None),
    # Mur on all faces, focused-laser soft source at i = 5, electrons filling
    # the box (the open walls delete markers within one cell) and a denser
    # slab with ions.
    "laser_mur": dict(
        n=(24, 16, 16), L=(1.2, 0.8, 0.8), bc=(2, 2, 2, 2, 2, 2),
        species=[dict(sub=(1, 1, 1, 1), rho=-0.5, qm=-1.0,
                      u=(0, 0, 0), du=(0.6, 0.6, 0.6)),
                 dict(sub=(2, 1, 1, 1), rho=-4.0, qm=-1.0,
                      u=(0, 0, 0), du=(0.2, 0.2, 0.2),
                      box=((0.8, 0.4, 0.4), (0.1, 0.3, 0.3))),
                 dict(sub=(1, 1, 1, 2), rho=4.0, qm=5.4466e-4,
                      u=(0, 0, 0), du=(0, 0, 0),
                      box=((0.8, 0.4, 0.4), (0.1, 0.3, 0.3)))],
        laser=dict(a0=1.5, waist=0.25, focus=(0.7, 0.4, 0.4), dt_fs=4.0,
                   T0_fs=2.0, phi0=0.3)),
    # Mur box with the hard Gaussian spot source (source/core/em_gaussSpot.c:
    # 129-156) on the plane i = 8, both polarisations, and electrons.
    "spot": dict(
        n=(16, 12, 12), L=(0.8, 0.6, 0.6), bc=(2, 2, 2, 2, 2, 2),
        species=[dict(sub=(1, 1, 1, 1), rho=-0.5, qm=-1.0,
                      u=(0, 0, 0), du=(0.4, 0.4, 0.4))],
        laser=None,
        spot=dict(ampl=2.0, frequency=1.0, x=0.4, y=0.31, z=0.28, width=0.1,
                  duration=0.12, front=0.15, plateau=0.05, Ey=1, Ez=-1)),
    # Mur box with a TFSF interface played forward from synthetic face
    # recordings (source/core/em_TFSF.c); 15 frames, so the interface
    # switches itself off after 13 of the 20 steps.
    "tfsf": dict(
        n=(24, 20, 18), L=(1.2, 1.0, 0.9), bc=(2, 2, 2, 2, 2, 2),
        species=[dict(sub=(1, 1, 1, 1), rho=-0.5, qm=-1.0,
                      u=(0, 0, 0), du=(0.4, 0.4, 0.4))],
        laser=None,
        tfsf=dict(reg=(7, 16, 6, 13, 6, 11), mode=0, frames=15, opened=(0, 0, 0, 0, 0, 0),
                  seed=101)),
    # The same box played backward (sources of E subtracted), z faces opened.
    "tfsf_back": dict(
        n=(24, 20, 18), L=(1.2, 1.0, 0.9), bc=(2, 2, 2, 2, 2, 2),
        species=[dict(sub=(1, 1, 1, 1), rho=-0.5, qm=-1.0,
                      u=(0, 0, 0), du=(0.4, 0.4, 0.4))],
        laser=None,
        tfsf=dict(reg=(7, 16, 6, 13, 6, 11), mode=2, frames=40, opened=(0, 0, 0, 0, 1, 1),
                  seed=202)),
    # x periodic, y mirror, z Mur: every wall kind and their edges.
    "mixed": dict(
        n=(10, 12, 9), L=(0.5, 0.6, 0.405), bc=(0, 1, 2, 0, 1, 2),
        species=[dict(sub=(1, 2, 1, 1), rho=-1.5, qm=-1.0,
                      u=(0.1, 0.2, -0.3), du=(1.6, 1.6, 1.6)),
                 dict(sub=(1, 1, 1, 1), rho=1.5, qm=5.4466e-4,
                      u=(0, 0, 0), du=(0.01, 0.01, 0.01))],
        laser=None),
    # The mixed box with the VSP (beta, alpha, beta) convolution of J and rho
    # at alpha = 0.8 (core.out's --VSP:weight option, source/core/main.c:
    # 314-320, source/core/plasma_VSP.c).
    "vsp": dict(
        n=(10, 12, 9), L=(0.5, 0.6, 0.405), bc=(0, 1, 2, 0, 1, 2),
        species=[dict(sub=(1, 2, 1, 1), rho=-1.5, qm=-1.0,
                      u=(0.1, 0.2, -0.3), du=(1.6, 1.6, 1.6)),
                 dict(sub=(1, 1, 1, 1), rho=1.5, qm=5.4466e-4,
                      u=(0, 0, 0), du=(0.01, 0.01, 0.01))],
        laser=None, args=["--VSP:weight(0.8)"], vsp=0.8),
}


def laser_map(n, h, tau, lz):
    """Tangential incident-field map of a y-polarised paraxial Gaussian beam
    for the one-plane soft source (E map drives H at plane I, H map drives E
    at plane I-1; signs as the soft-source algebra of
    source/scpic/em_mirror.h:56-96 requires).  Records {y_cos, y_sin, z_cos,
    z_sin} for j, k in [5, n-5] (k fastest)."""
    I = 5
    k0 = 2 * math.pi
    w0 = lz["waist"]
    xf, yf, zf = lz["focus"]
    zr = math.pi * w0 * w0
    js = np.arange(5, n[1] - 5 + 1)
    ks = np.arange(5, n[2] - 5 + 1)
    jj, kk = np.meshgrid(js, ks, indexing="ij")

    def beam(x, y, z):
        xi = x - xf
        w = w0 * math.sqrt(1 + (xi / zr) ** 2)
        invR = xi / (xi * xi + zr * zr)
        gouy = math.atan2(xi, zr)
        r2 = (y - yf) ** 2 + (z - zf) ** 2
        amp = (w0 / w) * np.exp(-r2 / (w * w))
        # phase relative to the plane x = I*h (the reference's 'delay'
        # is the travel time from that plane to the focus)
        th = -k0 * r2 * invR / 2 + gouy - k0 * (x - I * h[0])
        return amp * np.cos(th), -amp * np.sin(th)

    mapE = np.zeros(jj.shape + (4,))
    mapH = np.zeros(jj.shape + (4,))
    c, s = beam((I - 1) * h[0], (jj - 0.5) * h[1], kk * h[2])      # Ey^
    mapE[..., 0], mapE[..., 1] = c, s
    c, s = beam((I - 0.5) * h[0], (jj - 0.5) * h[1], kk * h[2])    # Hz^
    mapH[..., 0], mapH[..., 1] = c, s          # Ez[I-1] += +c1*Hy^ (zero)
    mapH[..., 2], mapH[..., 3] = -c, -s        # Ey[I-1] += -c1*Hz^
    # Hy[I] += -c1*Ez^ (zero for y polarisation), Hz[I] += +c1*Ey^
    return np.ascontiguousarray(np.stack([mapE.reshape(-1, 4),
                                          mapH.reshape(-1, 4)]))


# Face file names in slot order axis*4 + top*2 + comp (tfsfE / tfsfH of
# source/core/em_TFSF.c:82-123).
TFSF_E_NAMES = ("Ey_xMin", "Ez_xMin", "Ey_xMax", "Ez_xMax", "Ez_yMin", "Ex_yMin",
                "Ez_yMax", "Ex_yMax", "Ex_zMin", "Ey_zMin", "Ex_zMax", "Ey_zMax")
TFSF_H_NAMES = ("Hz_xMin", "Hy_xMin", "Hz_xMax", "Hy_xMax", "Hx_yMin", "Hz_yMin",
                "Hx_yMax", "Hz_yMax", "Hy_zMin", "Hx_zMin", "Hy_zMax", "Hx_zMax")


def tfsf_faces(reg):
    """Node boxes (lo, hi) of the 12 E faces and the 12 H faces of an
    interface box reg = (imin, imax, jmin, jmax, kmin, kmax), following
    TFSF_init (source/core/em_TFSF.c:500-537)."""
    side = ((reg[0], reg[2], reg[4]), (reg[1], reg[3], reg[5]))
    fe, fh = [], []
    for ea in range(3):
        for top in range(2):
            for comp in range(2):
                lo, hi = list(side[0]), list(side[1])
                lo[ea] = hi[ea] = side[top][ea]
                c = (ea + 1 + comp) % 3
                lo[c] += 1
                fe.append((tuple(lo), tuple(hi)))
                lo[ea] += top
                hi[ea] += top
                fh.append((tuple(lo), tuple(hi)))
    return fe, fh


def box_volume(b):
    lo, hi = b
    return int(np.prod([hi[a] - lo[a] + 1 for a in range(3)]))


def write_tfsf(wd, cfg, h, tau):
    """Synthetic recordings in the play layout (reg_t header of 11 int32,
    then frames x face nodes doubles, source/core/em_TFSF.c:630-659) and the
    control files core.out reads in play mode (:400-470).  Returns the
    arrays handed to the oracle and the GPU."""
    t = cfg["tfsf"]
    reg, nf = t["reg"], t["frames"]
    rng = np.random.default_rng(t["seed"])
    fe, fh = tfsf_faces(reg)
    d = os.path.join(wd, ".EM_sources/TFSF")
    dataE, dataH = [], []
    opened = t["opened"]                    # xMin xMax yMin yMax zMin zMax
    for f in range(12):
        ea, top = f // 4, (f // 2) % 2
        vol = box_volume(fe[f])
        # a smooth face pattern times a random amplitude per frame
        a = rng.standard_normal(nf)[:, None] * 0.05
        b = rng.standard_normal(vol)[None, :] * 0.5 + 1.0
        e_src = (a * b).reshape(-1)                         # played onto E face f
        a = rng.standard_normal(nf)[:, None] * 0.05
        b = rng.standard_normal(vol)[None, :] * 0.5 + 1.0
        h_src = (a * b).reshape(-1)                         # played onto H face f
        off = opened[2 * ea + top]
        dataE.append(None if off else e_src)
        dataH.append(None if off else h_src)
        for name, box, arr in ((TFSF_H_NAMES[f], fe[f], e_src), (TFSF_E_NAMES[f], fh[f], h_src)):
            hdr = np.zeros(11, "<i4")
            hdr[0:3], hdr[3:6] = box[0], box[1]
            with open(os.path.join(d, name + "_99999.bin"), "wb") as fp:
                fp.write(hdr.tobytes())
                fp.write(arr.astype("<f8").tobytes())
    # every value line carries a comment: cfg_readDouble's "@ %le %*[^\\n] "
    # (source/misc_cfgReader.c:62) would otherwise swallow the next line
    lines = ["@ %d    bounding box" % v for v in reg] + ["@ %d    mode" % t["mode"]]
    open(os.path.join(d, "TFSF.cfg"), "w").write("\n".join(lines) + "\n")
    lines = ["@ %d    interface size" % (reg[2 * a + 1] - reg[2 * a]) for a in range(3)]
    lines += ["@ %.17e    time step" % tau] + ["@ %.17e    spatial step" % v for v in h]
    open(os.path.join(d, "consistencyCheck.dat"), "w").write("\n".join(lines) + "\n")
    np.array([nf], "<i4").tofile(os.path.join(d, "lastFrameTime.bin"))
    if any(opened):
        lines = ["@ %d    face opened" % v for v in opened]
        open(os.path.join(d, "openFaces.cfg"), "w").write("\n".join(lines) + "\n")
    vols = np.array([box_volume(b) for b in fe], dtype=np.int64)
    return dict(tfsf_reg=np.array(reg, dtype=np.int32), tfsf_mode=np.array([t["mode"]]),
                tfsf_frames=np.array([nf]), tfsf_vol=vols,
                tfsf_on=np.array([0 if x is None else 1 for x in dataE], dtype=np.int32),
                tfsf_E=np.concatenate([x if x is not None else np.zeros(nf * v)
                                       for x, v in zip(dataE, vols)]),
                tfsf_H=np.concatenate([x if x is not None else np.zeros(nf * v)
                                       for x, v in zip(dataH, vols)]))


def read_mesh(path, n):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:24], "<i4")
    assert tuple(hdr) == (0, 0, 0) + tuple(n), hdr
    data = np.frombuffer(raw[24 + 88:], "<f8")
    assert data.size == 3 * np.prod([v + 1 for v in n])
    return data.copy()


def read_markers(path):
    raw = open(path, "rb").read()
    N = int(np.frombuffer(raw[:8], "<i8")[0])
    m = np.frombuffer(raw[8:], MARKER_DTYPE)
    assert m.size == N
    return m.copy()


def marker_digest(m):
    b = np.ascontiguousarray(m).view(np.uint8).reshape(-1, 64)[:, :60]
    return hashlib.sha256(b.tobytes()).hexdigest()


def run(cmd, cwd, env):
    r = subprocess.run(cmd, cwd=cwd, env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=600)
    if r.returncode != 0:
        sys.stdout.write(r.stdout.decode(errors="replace")[-4000:])
        raise RuntimeError("%s failed (%d)" % (cmd, r.returncode))


def make_case(name, cfg, workroot):
    env = dict(os.environ, PATH="/opt/conda/bin:" + os.environ["PATH"],
               MPICH_CC="gcc")
    out = {}
    wd = os.path.join(workroot, name + "_plain")
    os.makedirs(wd)
    os.symlink(os.path.join(REF, "source"), os.path.join(wd, "source"))
    for d in ("binData", "output/logs", "tmp", ".EM_sources/TFSF",
              ".EM_sources/laser"):
        os.makedirs(os.path.join(wd, d))
    open(os.path.join(wd, "case.input"), "w").write(
        deck(cfg["n"], cfg["L"], cfg["bc"], cfg["species"]))
    open(os.path.join(wd, "run_mandor.cfg"), "w").write(
        "@ %d\n@ 1\n@ -100\n@ 0\n@ 1000\n@ -100\n" % STEPS)
    run(["mpirun", "-np", "1", os.path.join(BIN, "setup.out"),
         "case.input", "create"], wd, env)
    dom = open(os.path.join(wd, "binData/domain.bin"), "rb").read()
    d7 = np.frombuffer(dom[:56], "<f8")
    i12 = np.frombuffer(dom[56:104], "<i4")
    n = tuple(int(v) for v in i12[3:6])
    assert n == tuple(cfg["n"])
    units = np.frombuffer(open(os.path.join(wd, "binData/units.bin"),
                               "rb").read(), "<f8")
    lz = cfg["laser"]
    if lz is not None:
        hh = (d7[0], d7[1], d7[2])
        lmap = laser_map(n, hh, d7[6], lz)
        fp = open(os.path.join(wd, ".EM_sources/laser/mirror_00"), "wb")
        hdr = ["Mirror focused laser beam #0\n",
               "---------------------------------------\n",
               "@ %.17e   A\n" % (-lz["a0"]),
               "@ 0.9   f-number\n",
               "@ %.17e   xf\n" % lz["focus"][0],
               "@ %.17e   yf\n" % lz["focus"][1],
               "@ %.17e   zf\n" % lz["focus"][2],
               "@ %.17e   dt\n" % lz["dt_fs"],
               "@ %.17e   T0\n" % lz["T0_fs"],
               "@ %.17e   phi0\n" % lz["phi0"],
               "@ -1   rratio\n",
               "@ 1       direction\n",
               "---------------------------------------\n"]
        fp.write("".join(hdr).encode())
        fp.write(lmap.tobytes())
        fp.close()
        # core.out's conversions (source/scpic/em_mirror.c:288-322)
        r0, t0 = units[0], units[2]
        xf = lz["focus"][0] * 1e-4 / r0
        dt = lz["dt_fs"] * 1e-15 / t0
        dt *= math.sqrt(2.0 / math.log(2))
        T0 = lz["T0_fs"] * 1e-15 / t0
        out["laser_params"] = np.array(
            [-2 * math.pi * -lz["a0"], T0, xf - 5 * d7[0], dt,
             lz["phi0"], 1.0, 5.0])
        out["laser_map"] = lmap.reshape(-1)
    sp = cfg.get("spot")
    if sp is not None:
        # .gaussSpot.cfg in code units, read by gaussSpot_init
        # (source/core/em_gaussSpot.c:42-55)
        keys = ("ampl", "frequency", "x", "y", "z", "width", "duration", "front", "plateau")
        lines = ["@ %.17e   %s" % (sp[k], k) for k in keys]
        lines += ["@ %d   Ey" % sp["Ey"], "@ %d   Ez" % sp["Ez"]]
        open(os.path.join(wd, ".gaussSpot.cfg"), "w").write("\n".join(lines) + "\n")
        out["spot_params"] = np.array([sp[k] for k in keys] + [sp["Ey"], sp["Ez"]], dtype=float)
    if cfg.get("tfsf") is not None:
        out.update(write_tfsf(wd, cfg, d7[:3], d7[6]))
    if cfg.get("vsp") is not None:
        out["vsp"] = np.array([cfg["vsp"]])
    # setup.out seeds rand() from the clock (source/misc_partition.c:255):
    # the Gauss build runs on a copy of the same record 0.
    wg = os.path.join(workroot, name + "_gauss")
    shutil.copytree(wd, wg, symlinks=True)
    for variant, exe, wdir in (("plain", "core.out", wd),
                               ("gauss", "core_gauss.out", wg)):
        wd = wdir
        rec0 = dict(E=read_mesh(os.path.join(wd, "binData/sys_E_000000(000).bin"), n),
                    H=read_mesh(os.path.join(wd, "binData/sys_H_000000(000).bin"), n),
                    M=read_markers(os.path.join(wd, "binData/plasma_000000_000.bin")))
        run(["mpirun", "-np", "1", os.path.join(BIN, exe)] + list(cfg.get("args", [])), wd, env)
        if variant == "plain":
            out["h"] = d7[:3].copy()
            out["L"] = d7[3:6].copy()
            out["tau"] = np.array([d7[6]])
            out["n"] = np.array(n, dtype=np.int32)
            out["bc"] = i12[6:12].astype(np.int32).copy()
            out["E0"], out["H0"], out["M0"] = rec0["E"], rec0["H"], rec0["M"]
            dE, dH, dM, times = [], [], [], []
            for r in range(1, STEPS + 1):
                E = read_mesh(os.path.join(wd, "binData/sys_E_%06d(000).bin" % r), n)
                H = read_mesh(os.path.join(wd, "binData/sys_H_%06d(000).bin" % r), n)
                M = read_markers(os.path.join(wd, "binData/plasma_%06d_000.bin" % r))
                dE.append(hashlib.sha256(E.tobytes()).hexdigest())
                dH.append(hashlib.sha256(H.tobytes()).hexdigest())
                dM.append(marker_digest(M))
                t = open(os.path.join(wd, "binData/sys_%06d.txt" % r)).readline()
                times.append(float(t.split()[1]))
                if r in (1, STEPS):
                    out["E%d" % r], out["H%d" % r], out["M%d" % r] = E, H, M
            out["digest_E"] = np.array(dE)
            out["digest_H"] = np.array(dH)
            out["digest_M"] = np.array(dM)
            out["times"] = np.array(times)
            diag = np.loadtxt(os.path.join(wd, "binData/diag.dat"), skiprows=2)
            out["diag"] = diag
        else:
            g = np.loadtxt(os.path.join(wd, "output/gauss_test.dat"), skiprows=1)
            out["gauss"] = g
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("%s: %d markers, %d steps, %d bytes" % (
        name, out["M0"].size, STEPS, os.path.getsize(os.path.join(HERE, name + ".npz"))))


def main():
    work = tempfile.mkdtemp(prefix="mandor_fix_")
    try:
        for name in (sys.argv[1:] or list(CASES)):
            make_case(name, CASES[name], work)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main()
