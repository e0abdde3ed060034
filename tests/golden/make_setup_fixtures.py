"""Writes tests/golden/setup_fixtures.npz: record 0 and the source control
files that the reference ``setup.out`` writes for a set of decks covering
its deterministic tags.

Runs in the development container only (the reference build is not on the
GPU host).  For every deck it runs ``mpirun -np 1 setup.out case.input
create`` and keeps:

  * the deck text,
  * domain.bin and units.bin (raw bytes),
  * the marker records of plasma_000000_000.bin without the padding word
    (the reference leaves it uninitialised),
  * the '@' values of .gaussSpot.cfg, .EM_sources/TFSF/TFSF.cfg, filter.cfg
    and openFaces.cfg where the deck writes them,
  * the record descriptors (sys_000000.txt, sys.N, file maps, partition
    files, tec_000000.txt, tecplot.N, tracer.inf) and the charge density of
    the Tecplot record 0.

[DF:uniform] momenta are drawn with rand() seeded from the wall clock
(source/misc_partition.c:255), so the decks use it only with zero spread,
where the momenta are the mean exactly.  The fixtures are data; no reference
source is copied.  Usage, from the repository root:

    make -f oracle/ref.mk && python tests/golden/make_setup_fixtures.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "oracle", "_ref")

MARKER_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
                         ("vx", "<f8"), ("vy", "<f8"), ("vz", "<f8"),
                         ("rho", "<f4"), ("qDivM", "<f4"), ("id", "<i4"),
                         ("pad", "<i4")])


def head(lam, n, L, bc, tau=-0.45):
    lines = ["Deck of the setup fixtures.", "[units]", "@ %r  lambda or -n_cr" % lam, "[mesh]"]
    lines += ["@ %d  cells" % v for v in n]
    lines += ["@ %r  L [micron]" % v for v in L]
    lines += ["@ %r  tau" % tau, "[boundaryConditions]"]
    lines += ["@ %d  bc" % v for v in bc]          # xMin xMax yMin yMax zMin zMax
    return lines


def uniform(rho, qm, u):
    """[DF:uniform] with zero spread: deterministic momenta."""
    return (["[DF:uniform]", "@ %r  rho" % rho, "@ %r  q/M" % qm]
            + ["@ %r  u" % v for v in u] + ["@ 0  du"] * 3)


def plasma(sub, shapes=()):
    lines = ["[plasma]"] + ["@ %d  sampling" % v for v in sub]
    for kind, rows in shapes:
        lines.append("@ %s  shape" % kind)
        lines += ["> %s  row" % r for r in rows]
    return lines


DECKS = {
    # Lattice plasma cut by a slanted half-space polyhedron, then a second
    # species cut by two spheres (one kept inside, one cut out), and an
    # exponential gradient on the newest object.  ([DF:Maxwell] cannot be
    # run: tag_plasma.c:109 stores the pattern places with the unsequenced
    # 'cell[p].vx = p++', which gcc and clang both evaluate increment first,
    # leaving NaN places that crash tag_maxwell.c:162.)
    "shapes": head(1.0, (10, 8, 6), (0.5, 0.4, 0.3), (0, 0, 0, 0, 0, 0))
    + plasma((2, 1, 1, 4), [("planes", ["0.25 0.2 0.15  point", "1 0.5 -0.3  normal", "0.02  shift",
                                        "0.1 0.1 0.1  point", "-1 0 0  normal", "0  shift"])])
    + uniform(-0.7, -1.0, (0.05, 0.03, 0.04))
    + plasma((1, 2, 1, 2), [("spheres", ["0.25 0.2 0.15 0.17  keep inside",
                                         "0.3 0.25 0.1 -0.06  cut out"])])
    + uniform(0.7, 5.4466e-4, (0.002, 0.003, 0.001))
    + ["[gradient]", "> 0.1 0 0  point 1", "> 0.4 0.1 0  point 2", "> 1.0 3.0 1  exponential"],

    # Two foils, linear with a shift along a tilted normal and exponential
    # in n_c units along x, 5-cell margins on the Mur x walls only; a
    # linear gradient on the second; then a lattice cut by cylinders and a
    # box, with zero-spread uniform momenta.
    "foil": head(0.8, (24, 8, 8), (0.96, 0.32, 0.32), (2, 2, 0, 0, 0, 0))
    + ["[DF:foil]", "@ 0.4  x0", "@ 0.16  y0", "@ 0.16  z0", "@ 1  lx", "@ 0.2  ly", "@ -0.1  lz",
       "@ 0.12  width", "> 0.02  shift", "@ linear  profile", "@ 3e21  n1", "@ 9e21  n2",
       "@ -1  q", "@ 1  M", "@ 2  nx", "@ 2  ny", "@ 1  nz"]
    + ["[DF:foil]", "@ 0.6  x0", "@ 0  y0", "@ 0  z0", "@ 1  lx", "@ 0  ly", "@ 0  lz",
       "@ 0.08  width", "@ exponential  profile", "@ -2  n1", "@ -30  n2",
       "@ 1  q", "@ 1836  M", "@ 1  nx", "@ 1  ny", "@ 2  nz"]
    + ["[gradient]", "> 0.55 0 0  point 1", "> 0.65 0 0  point 2", "> 0.5 2 0  linear"]
    + plasma((1, 1, 2, 1), [("cylinders", ["0.5 0.16 0.16  axis point", "0 0 1  axis", "0.1  radius",
                                           "0.5 0.1 0.1  axis point", "1 1 0  axis", "-0.03  radius"]),
                            ("boxes", ["0.5 0.16 0.16  centre", "0.3 0 0  half-axis",
                                       "0 0.2 0.05  half-axis", "0 0 0.25  half-axis"])])
    + uniform(-2.0, -1.0, (0.1, -0.2, 0.05)),

    # Two cold beams tuned to a mode, units set by the critical density.
    "two_stream": head(-1.1e21, (12, 10, 8), (0.6, 0.5, 0.4), (0, 0, 0, 0, 0, 0), tau=-0.5)
    + ["[DF:two-stream]", "@ 0.05  growth rate", "@ -1.0  q/M", "@ 6  Nx", "@ 5  Ny", "@ 4  Nz",
       "@ 1  mx", "@ 1  my", "@ 0  mz"],

    # Control files of the sources: a hard Gaussian spot with both optional
    # lines, a TFSF box played backward with a filter block, and two opened
    # faces; one lattice species.
    "sources": head(1.0, (24, 20, 18), (1.2, 1.0, 0.9), (2, 2, 2, 2, 2, 2))
    + ["[gaussSpot]", "@ 3e18  I", "@ 1.2  omega", "@ 0.5  X", "@ 0.45  Y", "@ 0.4  Z",
       "@ 0.2  width", "@ 2.5  duration", "@ 4.0  front", "> 1.5  plateau", "@ 1  Ey", "@ -3  Ez",
       "> 0.9  TFSF position"]
    + ["[TFSF]", "@ 7  imin", "@ 16  imax", "@ 6  jmin", "@ 13  jmax", "@ 6  kmin", "@ 11  kmax",
       "@ playBackward  mode", "> new  blend", "> 3  I0", "> 12  I1", "> 2  shift"]
    + ["[TFSF: open faces]", "> xMin  face", "> zMax  face"]
    + plasma((1, 1, 1, 1), [("boxes", ["0.6 0.5 0.45  centre", "0.2 0 0  half-axis",
                                       "0 0.2 0  half-axis", "0 0 0.2  half-axis"])])
    + uniform(-0.5, -1.0, (0.0, 0.0, 0.3)),
}

CONTROL = {"gaussSpot": ".gaussSpot.cfg", "TFSF": ".EM_sources/TFSF/TFSF.cfg",
           "filter": ".EM_sources/TFSF/filter.cfg", "openFaces": ".EM_sources/TFSF/openFaces.cfg"}


RECORD_FILES = ("sys_000000.txt", "sys.N", "sys_prtn_000000.bin", "fileMap_000000.txt",
                "fileMap_000001.txt", "fileMap.N", "fileMap_start.txt", "tec_000000.txt",
                "tec_prtn_000000.bin", "tecplot.N", "tracer.inf")


def at_values(path):
    """The first word of every '@' line of a control file."""
    return [ln[1:].split()[0] for ln in open(path).read().splitlines() if ln.startswith("@")]


def run_case(name, text, root):
    wd = os.path.join(root, name)
    os.makedirs(wd)
    os.symlink(os.path.join(REF, "source"), os.path.join(wd, "source"))
    for d in ("binData", "output/logs", "tmp", ".EM_sources/TFSF", ".EM_sources/laser"):
        os.makedirs(os.path.join(wd, d))
    open(os.path.join(wd, "case.input"), "w").write(text)
    env = dict(os.environ, PATH="/opt/conda/bin:" + os.environ["PATH"], MPICH_CC="gcc")
    r = subprocess.run(["mpirun", "-np", "1", os.path.join(BIN, "setup.out"), "case.input", "create"],
                       cwd=wd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=600)
    if r.returncode != 0:
        sys.stdout.write(r.stdout.decode(errors="replace")[-4000:])
        raise RuntimeError("setup.out failed on " + name)
    raw = open(os.path.join(wd, "binData/plasma_000000_000.bin"), "rb").read()
    n = int(np.frombuffer(raw[:8], "<i8")[0])
    m = np.frombuffer(raw[8:], MARKER_DTYPE)
    assert m.size == n
    out = {name + "/deck": np.array(text),
           name + "/domain": np.frombuffer(open(os.path.join(wd, "binData/domain.bin"), "rb").read(), np.uint8),
           name + "/units": np.frombuffer(open(os.path.join(wd, "binData/units.bin"), "rb").read(), np.uint8)}
    for k in MARKER_DTYPE.names[:-1]:
        out["%s/M/%s" % (name, k)] = np.ascontiguousarray(m[k])
    for k, p in CONTROL.items():
        p = os.path.join(wd, p)
        if os.path.exists(p):
            out["%s/%s" % (name, k)] = np.array(at_values(p))
    # record 0 descriptors and the Tecplot record 0 (rho payload: the mesh_t
    # header holds pointers, so only the region and the data are kept)
    for f in RECORD_FILES:
        out["%s/file/%s" % (name, f)] = np.frombuffer(open(os.path.join(wd, "binData", f), "rb").read(),
                                                      np.uint8)
    raw = open(os.path.join(wd, "binData", "tec_rho_000000(000).bin"), "rb").read()
    out[name + "/tec_rho"] = np.frombuffer(raw[24 + 88:], "<f8")
    return out


def main():
    out = {}
    with tempfile.TemporaryDirectory() as root:
        for name, lines in DECKS.items():
            out.update(run_case(name, "\n".join(lines) + "\n", root))
            print(name, out[name + "/M/x"].size, "markers")
    np.savez_compressed(os.path.join(HERE, "setup_fixtures.npz"), **out)


if __name__ == "__main__":
    main()
