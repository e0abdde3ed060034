"""Writes tests/golden/tfsf_record.npz: what the reference writes when a
TFSF interface records (source/core/em_TFSF.c:273-330, 573-624, 866-890)
and what TFSFdefrag (source/TFSFdefrag/main.c) makes of it.

Runs in the development container only.  The reference setup.out makes
record 0 of DECK (an all-Mur box, a hard Gaussian spot inside the interface
box, electrons); core.out runs 20 steps recording the interface; then
TFSFdefrag joins the fragments.  Stored, as data: the marker record 0 (its
[DF:uniform] momenta come from the clock-seeded rand(), so a replay starts
from it), the 24 fragment files <face>_000.bin, TFSF_defrag.map,
consistencyCheck.dat and lastFrameTime.bin as core.out left them, and the
24 joined files <face>_99999.bin and lastFrameTime.bin after TFSFdefrag.

Usage, from the repository root:

    make -f oracle/ref.mk && python tests/golden/make_tfsf_record_fixture.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_fixtures as mf  # noqa: E402

STEPS = 20
N, L, BC = (24, 20, 18), (1.2, 1.0, 0.9), (2, 2, 2, 2, 2, 2)
REG = (7, 16, 6, 13, 6, 11)
SPECIES = [dict(sub=(1, 1, 1, 1), rho=-0.5, qm=-1.0, u=(0, 0, 0), du=(0.4, 0.4, 0.4))]


def deck():
    text = mf.deck(N, L, BC, SPECIES)
    text += "\n".join(["[gaussSpot]", "@ -2.0  a0", "@ 1.0  omega", "@ 0.5  X", "@ 0.45  Y", "@ 0.4  Z",
                       "@ 0.2  width", "@ 5.0  duration", "@ 5.0  front offset", "@ 1  Ey", "@ 1  Ez",
                       "[TFSF]"] + ["@ %d  box" % v for v in REG] + ["@ record  mode"]) + "\n"
    return text


FACES = mf.TFSF_E_NAMES + mf.TFSF_H_NAMES


def main():
    env = dict(os.environ, PATH="/opt/conda/bin:" + os.environ["PATH"], MPICH_CC="gcc")
    out = {"deck": np.array(deck()),
           "run_cfg": np.array("@ %d steps\n@ 100 checkpoint\n@ -100 tecplot\n@ 0 continue\n"
                               "@ 1000 stop check\n@ -100 spectr\n" % STEPS)}
    with tempfile.TemporaryDirectory() as root:
        wd = os.path.join(root, "rec")
        os.makedirs(wd)
        os.symlink(os.path.join(mf.REF, "source"), os.path.join(wd, "source"))
        for d in ("binData", "output/logs", "tmp", ".EM_sources/TFSF", ".EM_sources/laser"):
            os.makedirs(os.path.join(wd, d))
        open(os.path.join(wd, "case.input"), "w").write(str(out["deck"]))
        open(os.path.join(wd, "run_mandor.cfg"), "w").write(str(out["run_cfg"]))
        mf.run(["mpirun", "-np", "1", os.path.join(mf.BIN, "setup.out"), "case.input", "create"], wd, env)
        out["M0"] = mf.read_markers(os.path.join(wd, "binData/plasma_000000_000.bin"))
        mf.run(["mpirun", "-np", "1", os.path.join(mf.BIN, "core.out")], wd, env)
        t = os.path.join(wd, ".EM_sources/TFSF")
        for f in FACES:
            out["rec/" + f] = np.frombuffer(open(os.path.join(t, f + "_000.bin"), "rb").read(), np.uint8)
        for f in ("TFSF_defrag.map", "consistencyCheck.dat", "lastFrameTime.bin"):
            out["rec/" + f] = np.frombuffer(open(os.path.join(t, f), "rb").read(), np.uint8)
        mf.run([os.path.join(mf.BIN, "TFSFdefrag.out")], wd, env)
        for f in FACES:
            out["join/" + f] = np.frombuffer(open(os.path.join(t, f + "_99999.bin"), "rb").read(), np.uint8)
        out["join/lastFrameTime.bin"] = np.frombuffer(open(os.path.join(t, "lastFrameTime.bin"), "rb").read(),
                                                      np.uint8)
        out["join/left"] = np.array(sorted(os.listdir(t)))
    np.savez_compressed(os.path.join(HERE, "tfsf_record.npz"), **out)


if __name__ == "__main__":
    main()
