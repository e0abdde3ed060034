"""Readers of the reference's binData records, written from the formats
(not from this build's C code), for checking the C host side:

  domain.bin   source/misc_parameters.c:98-110
  units.bin    source/misc_units.c:39-61
  sys_E/H      source/type_mesh.c:285-311 (24 B region, 88 B mesh_t, nodes)
  plasma       source/core/plasma_IO.c:273-287 (int64 N, marker_t records)
  sys_%06d.txt source/IO_sys.c:113-118
"""
import os

import numpy as np

from oracle_lib import MARKER_DTYPE


def domain(path):
    raw = open(path, "rb").read()
    assert len(raw) == 7 * 8 + 12 * 4
    d7 = np.frombuffer(raw[:56], "<f8")
    i12 = np.frombuffer(raw[56:], "<i4")
    return dict(h=d7[0:3].copy(), L=d7[3:6].copy(), tau=float(d7[6]),
                dmn_min=i12[0:3].copy(), dmn_max=i12[3:6].copy(),
                bc_min=i12[6:9].copy(), bc_max=i12[9:12].copy())


def units(path):
    v = np.frombuffer(open(path, "rb").read(), "<f8")
    assert v.size == 15
    return v.copy()


def mesh(path, n):
    """Node block [0, n] (xyz interleaved, k fastest) of a one-rank record."""
    raw = open(path, "rb").read()
    reg = np.frombuffer(raw[:24], "<i4")
    head = np.frombuffer(raw[24:24 + 40], "<i4")      # imin..kmax, type, widths, size
    assert head[6] == 1, "vector mesh expected"
    lo, hi = reg[:3], reg[3:]
    shape = tuple(int(v) for v in hi - lo + 1) + (3,)
    data = np.frombuffer(raw[24 + 88:], "<f8").reshape(shape)
    out = np.zeros((n[0] + 1, n[1] + 1, n[2] + 1, 3))
    out[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = data
    return out.reshape(-1), reg.copy()


def markers(path):
    raw = open(path, "rb").read()
    N = int(np.frombuffer(raw[:8], "<i8")[0])
    assert (len(raw) - 8) == 64 * N, "64-byte tracer records expected"
    return np.frombuffer(raw[8:], MARKER_DTYPE).copy()


def sys_info(path):
    vals = [l.split()[1] for l in open(path) if l.startswith("@")]
    return float(vals[0]), int(vals[1]), int(vals[2])


def diag(path):
    rows = [l.split() for l in open(path).read().splitlines()[2:] if l.strip()]
    return np.array(rows, dtype=float)


def record(workdir, rec, n):
    b = os.path.join(workdir, "binData")
    E, _ = mesh(os.path.join(b, "sys_E_%06d(000).bin" % rec), n)
    H, _ = mesh(os.path.join(b, "sys_H_%06d(000).bin" % rec), n)
    M = markers(os.path.join(b, "plasma_%06d_000.bin" % rec))
    t, cpus, fmap = sys_info(os.path.join(b, "sys_%06d.txt" % rec))
    return dict(E=E, H=H, M=M, time=t)
