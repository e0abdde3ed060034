"""TFSF recordings in the play layout core.out reads (test input writer).

The face geometry follows TFSF_init (source/core/em_TFSF.c:500-537); the
files are the defragmented recordings (a reg_t header of 11 int32, then
frames x face nodes doubles, :630-659) plus the control files of the play
mode (TFSF.cfg :151-180, consistencyCheck.dat and lastFrameTime.bin
:430-462, openFaces.cfg :339-371).  Values are synthetic.
"""
import os

import numpy as np

E_NAMES = ("Ey_xMin", "Ez_xMin", "Ey_xMax", "Ez_xMax", "Ez_yMin", "Ex_yMin",
           "Ez_yMax", "Ex_yMax", "Ex_zMin", "Ey_zMin", "Ex_zMax", "Ey_zMax")
H_NAMES = ("Hz_xMin", "Hy_xMin", "Hz_xMax", "Hy_xMax", "Hx_yMin", "Hz_yMin",
           "Hx_yMax", "Hz_yMax", "Hy_zMin", "Hx_zMin", "Hy_zMax", "Hx_zMax")


def face_nodes(reg):
    """Nodes of each of the 12 face slots (E and H faces of a slot agree)."""
    out = []
    for ea in range(3):
        ep, eq = (ea + 1) % 3, (ea + 2) % 3
        for top in range(2):
            for comp in range(2):
                lo = [reg[0], reg[2], reg[4]]
                hi = [reg[1], reg[3], reg[5]]
                lo[(ea + 1 + comp) % 3] += 1
                out.append((hi[ep] - lo[ep] + 1) * (hi[eq] - lo[eq] + 1))
    return out


def write_dir(d, reg, mode, nframes, h, tau, opened=(0,) * 6, seed=7, scale=0.05):
    """Writes a play-mode .EM_sources/TFSF directory; returns (dataE, dataH),
    the arrays the interface plays onto E / H face slot f (None: opened)."""
    os.makedirs(d, exist_ok=True)
    rng = np.random.default_rng(seed)
    vols = face_nodes(reg)
    dataE, dataH = [], []
    for f in range(12):
        e = (rng.standard_normal(nframes)[:, None] * scale *
             (1 + 0.5 * rng.standard_normal(vols[f]))[None, :]).reshape(-1)
        hh = (rng.standard_normal(nframes)[:, None] * scale *
              (1 + 0.5 * rng.standard_normal(vols[f]))[None, :]).reshape(-1)
        for name, arr in ((H_NAMES[f], e), (E_NAMES[f], hh)):
            with open(os.path.join(d, name + "_99999.bin"), "wb") as fp:
                fp.write(np.zeros(11, "<i4").tobytes())
                fp.write(arr.astype("<f8").tobytes())
        off = opened[2 * (f // 4) + (f // 2) % 2]
        dataE.append(None if off else e)
        dataH.append(None if off else hh)
    lines = ["@ %d    box" % v for v in reg] + ["@ %d    mode" % mode]
    open(os.path.join(d, "TFSF.cfg"), "w").write("\n".join(lines) + "\n")
    lines = ["@ %d    size" % (reg[2 * a + 1] - reg[2 * a]) for a in range(3)]
    lines += ["@ %.17e    tau" % tau] + ["@ %.17e    h" % v for v in h]
    open(os.path.join(d, "consistencyCheck.dat"), "w").write("\n".join(lines) + "\n")
    np.array([nframes], "<i4").tofile(os.path.join(d, "lastFrameTime.bin"))
    if any(opened):
        lines = ["@ %d    opened" % v for v in opened]
        open(os.path.join(d, "openFaces.cfg"), "w").write("\n".join(lines) + "\n")
    return dataE, dataH
