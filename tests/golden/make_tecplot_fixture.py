"""Writes tests/golden/tecplot_periodic.npz: the Tecplot records, probe files
and record descriptors of a reference core.out run.

Runs in the development container only.  The reference setup.out makes
record 0 of the 'periodic' time-step fixture deck (tests/golden/
make_fixtures.py); core.out then runs 20 steps with a checkpoint every 10
steps, a Tecplot record every 5 (source/core/main.c:386-435) and the probes
of PROBES (source/core/diag_probe.c:171-383).  Stored, as data:

  * the marker record 0 the run started from (its [DF:uniform] momenta come
    from the wall-clock-seeded rand(), so a replay must start from it),
  * per Tecplot record r = 1..4: the payloads of tec_E/H/J (node values,
    3 doubles per node of [0, n]) and tec_rho, and tec_%06d.txt,
  * every probe file (header and samples), probe_globals.bin and
    output/probe_locations.dat,
  * sys_000001.txt, sys_000002.txt, fileMap.N and tecplot.N.

Usage, from the repository root:

    make -f oracle/ref.mk && python tests/golden/make_tecplot_fixture.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_fixtures as mf  # noqa: E402

STEPS, FULL, TEC = 20, 10, 5
RUN_CFG = "@ %d steps\n@ %d checkpoint\n@ %d tecplot\n@ 0 continue\n@ 1000 stop check\n@ -100 spectr\n" % (
    STEPS, FULL, TEC)
# a point, a line with one inner point, a 3 x 2 rectangle, a 2 x 2 x 2 volume
# sharing nodes with the others (repeats are dropped), then the terminator
PROBES = """@ point       one probe
> 0.1         x [micron]
> 0.12        y
> 0.2         z
@ line        a line
> 0.0         x1
> 0.0         y1
> 0.0         z1
> 0.35        x2
> 0.3         y2
> 0.4         z2
> 1           inner points
@ rectangle   a rectangle
> 0.05        origin
> 0.3
> 0.1
> 0.3         vertex along side 1
> 0.3
> 0.1
> 0.05        vertex along side 2
> 0.3
> 0.4
> 1           inner points along side 1
> 0           along side 2
@ volume      a box of probes
> 0.1         origin
> 0.12
> 0.2
> 0.2         vertex 1
> 0.12
> 0.2
> 0.1         vertex 2
> 0.24
> 0.2
> 0.1         vertex 3
> 0.12
> 0.3
> 0           inner points
> 0
> 0
@ end         terminator
"""


def main():
    cfg = mf.CASES["periodic"]
    env = dict(os.environ, PATH="/opt/conda/bin:" + os.environ["PATH"], MPICH_CC="gcc")
    out = {"run_cfg": np.array(RUN_CFG), "probes_cfg": np.array(PROBES)}
    with tempfile.TemporaryDirectory() as root:
        wd = os.path.join(root, "periodic")
        os.makedirs(wd)
        os.symlink(os.path.join(mf.REF, "source"), os.path.join(wd, "source"))
        for d in ("binData", "output/logs", "tmp", ".EM_sources/TFSF", ".EM_sources/laser"):
            os.makedirs(os.path.join(wd, d))
        open(os.path.join(wd, "case.input"), "w").write(
            mf.deck(cfg["n"], cfg["L"], cfg["bc"], cfg["species"]))
        open(os.path.join(wd, "run_mandor.cfg"), "w").write(RUN_CFG)
        open(os.path.join(wd, "run_probes.cfg"), "w").write(PROBES)
        mf.run(["mpirun", "-np", "1", os.path.join(mf.BIN, "setup.out"), "case.input", "create"], wd, env)
        out["M0"] = mf.read_markers(os.path.join(wd, "binData/plasma_000000_000.bin"))
        mf.run(["mpirun", "-np", "1", os.path.join(mf.BIN, "core.out")], wd, env)
        b = os.path.join(wd, "binData")
        for r in range(1, STEPS // TEC + 1):
            for q in "EHJ":
                raw = open(os.path.join(b, "tec_%s_%06d(000).bin" % (q, r)), "rb").read()
                out["tec_%s_%d" % (q, r)] = np.frombuffer(raw[24 + 88:], "<f8")
            raw = open(os.path.join(b, "tec_rho_%06d(000).bin" % r), "rb").read()
            out["tec_rho_%d" % r] = np.frombuffer(raw[24 + 88:], "<f8")
            out["tec_txt_%d" % r] = np.frombuffer(open(os.path.join(b, "tec_%06d.txt" % r), "rb").read(),
                                                  np.uint8)
        names = sorted(f for f in os.listdir(b) if f.startswith("probe_") and f != "probe_globals.bin")
        out["probe_names"] = np.array(names)
        for f in names + ["probe_globals.bin"]:
            out["file/" + f] = np.frombuffer(open(os.path.join(b, f), "rb").read(), np.uint8)
        for f in ("sys_000001.txt", "sys_000002.txt", "fileMap.N", "tecplot.N", "sys.N"):
            out["file/" + f] = np.frombuffer(open(os.path.join(b, f), "rb").read(), np.uint8)
        out["probe_locations"] = np.array(open(os.path.join(wd, "output/probe_locations.dat")).read())
    np.savez_compressed(os.path.join(HERE, "tecplot_periodic.npz"), **out)
    print(len(names), "probes")


if __name__ == "__main__":
    main()
