"""Fairness gate of the CPU baseline: the oracle's single-core speed against
the reference's own CPU path on the same input.

Runs in the development container only (it needs the reference tree and the
reference build of oracle/ref.mk in oracle/_ref/).  On box32.input (32^3
periodic box, 2x2x2 markers per cell per species, uniform momenta; the probe
input of the reference timings) it

  1. makes record 0 with the reference setup.out,
  2. times the reference core.out (one MPI rank) over STEPS steps from its
     own "main loop have taken" line (source/core/main.c:480),
  3. loads the same record 0 into the oracle and times STEPS full steps,

and prints both rates in marker-pushes/s and their ratio.  The gate holds
when the oracle is within +-25 % of the reference.

    make -f oracle/ref.mk && python oracle/fairness.py
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = "/root/reference"
BIN = os.path.join(HERE, "_ref")
STEPS = 50

MARKER_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
                         ("vx", "<f8"), ("vy", "<f8"), ("vz", "<f8"),
                         ("rho", "<f4"), ("qDivM", "<f4"), ("id", "<i4"),
                         ("pad", "<i4")])

DECK = """[units]
@ 1.0\t\tlambda [micron]
[mesh]
@ 32\t\timax
@ 32\t\tjmax
@ 32\t\tkmax
@ 1.6\t\tLx [micron]
@ 1.6\t\tLy [micron]
@ 1.6\t\tLz [micron]
@ -0.45\t\ttau (Courant fraction)
[boundaryConditions]
@ 0\t\txMin
@ 0\t\txMax
@ 0\t\tyMin
@ 0\t\tyMax
@ 0\t\tzMin
@ 0\t\tzMax
[plasma]
@ 2\t\tnx
@ 2\t\tny
@ 2\t\tnz
@ 1\t\tlayers
[DF:uniform]
@ -0.1\t\trho
@ -1.0\t\tq/M
@ 0.0\t\tux
@ 0.0\t\tuy
@ 0.0\t\tuz
@ 0.17\t\tdUx
@ 0.17\t\tdUy
@ 0.17\t\tdUz
[plasma]
@ 2\t\tnx
@ 2\t\tny
@ 2\t\tnz
@ 1\t\tlayers
[DF:uniform]
@ 0.1\t\trho
@ 5.4466e-4\tq/M
@ 0.0\t\tux
@ 0.0\t\tuy
@ 0.0\t\tuz
@ 0.0\t\tdUx
@ 0.0\t\tdUy
@ 0.0\t\tdUz
"""


def run(cmd, cwd, env):
    r = subprocess.run(cmd, cwd=cwd, env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=1200)
    if r.returncode != 0:
        sys.stdout.write(r.stdout.decode(errors="replace")[-4000:])
        raise RuntimeError("%s failed (%d)" % (cmd, r.returncode))
    return r.stdout.decode(errors="replace")


def read_mesh(path, n):
    raw = open(path, "rb").read()
    data = np.frombuffer(raw[24 + 88:], "<f8")
    assert data.size == 3 * np.prod([v + 1 for v in n])
    return data.copy()


def main():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    env = dict(os.environ, PATH="/opt/conda/bin:" + os.environ["PATH"], MPICH_CC="gcc")
    wd = tempfile.mkdtemp(prefix="mandor_fair_")
    try:
        os.symlink(os.path.join(REF, "source"), os.path.join(wd, "source"))
        for d in ("binData", "output/logs", "tmp", ".EM_sources/TFSF", ".EM_sources/laser"):
            os.makedirs(os.path.join(wd, d))
        open(os.path.join(wd, "box32.input"), "w").write(DECK)
        open(os.path.join(wd, "run_mandor.cfg"), "w").write(
            "@ %d\n@ 2000\n@ -100\n@ 0\n@ 20\n@ -100\n" % STEPS)
        run(["mpirun", "-np", "1", os.path.join(BIN, "setup.out"), "box32.input", "create"],
            wd, env)
        dom = open(os.path.join(wd, "binData/domain.bin"), "rb").read()
        d7 = np.frombuffer(dom[:56], "<f8")
        i12 = np.frombuffer(dom[56:104], "<i4")
        n = tuple(int(v) for v in i12[3:6])
        E0 = read_mesh(os.path.join(wd, "binData/sys_E_000000(000).bin"), n)
        H0 = read_mesh(os.path.join(wd, "binData/sys_H_000000(000).bin"), n)
        raw = open(os.path.join(wd, "binData/plasma_000000_000.bin"), "rb").read()
        M0 = np.frombuffer(raw[8:], MARKER_DTYPE).copy()
        t0 = time.perf_counter()
        log = run(["mpirun", "-np", "1", os.path.join(BIN, "core.out")], wd, env)
        wall_ref = time.perf_counter() - t0
        logs = log
        for f in os.listdir(os.path.join(wd, "output/logs")):
            logs += open(os.path.join(wd, "output/logs", f), errors="replace").read()
        m = re.search(r"main loop have taken\s+([0-9.eE+-]+)", logs)
        loop_ref = float(m.group(1)) if m else wall_ref
        o = Oracle(n, d7[:3], d7[3:6], d7[6], i12[6:12], 0.0)
        o.set_fields(E0, H0)
        o.set_markers(M0)
        o.init()
        nm = o.marker_count()
        t0 = time.perf_counter()
        for _ in range(STEPS):
            o.step()
        loop_or = time.perf_counter() - t0
        o.close()
        ref = M0.size * STEPS / loop_ref
        ora = nm * STEPS / loop_or
        print("box32.input, %d markers, %d steps, one core:" % (M0.size, STEPS))
        print("  reference core.out  %.3e marker-pushes/s (main loop %.2f s%s)"
              % (ref, loop_ref, "" if m else ", process wall"))
        print("  oracle              %.3e marker-pushes/s (%.2f s)" % (ora, loop_or))
        print("  oracle / reference  %.3f  (gate: 0.75 .. 1.25)" % (ora / ref))
    finally:
        shutil.rmtree(wd, ignore_errors=True)


if __name__ == "__main__":
    main()
