"""ctypes binding of the CPU oracle (oracle/libopic.so) for the tests.

The oracle is test infrastructure: it is the checker the HIP product is
compared against, never part of the product path.
"""
import ctypes
import hashlib
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "libopic.so")

MARKER_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
                         ("vx", "<f8"), ("vy", "<f8"), ("vz", "<f8"),
                         ("rho", "<f4"), ("qDivM", "<f4"), ("id", "<i4"),
                         ("pad", "<i4")])

_lib = None


def build():
    """Compiles the oracle with the flags its bit-exactness relies on."""
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "-f", "Makefile"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.opic_create.restype = P
        L.opic_create.argtypes = [P, P, P, D, P, D]
        L.opic_destroy.argtypes = [P]
        L.opic_set_fields.argtypes = [P, P, P]
        L.opic_get_fields.argtypes = [P, P, P]
        L.opic_set_markers.argtypes = [P, P, ctypes.c_long, I]
        L.opic_marker_count.restype = ctypes.c_long
        L.opic_marker_count.argtypes = [P]
        L.opic_get_markers.argtypes = [P, P, I]
        L.opic_add_laser.argtypes = [P, D, D, D, D, D, I, I, P]
        L.opic_add_gauss_spot.argtypes = [P, P]
        L.opic_add_tfsf.argtypes = [P, P, I, I, P, P]
        L.opic_init.argtypes = [P]
        L.opic_step.argtypes = [P, P, P]
        L.opic_time.restype = D
        L.opic_time.argtypes = [P]
        L.opic_get_J.argtypes = [P, P]
        L.opic_rho.argtypes = [P, P]
        L.opic_set_vsp.argtypes = [P, D]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One oracle run: whole domain, FP64, reference arithmetic."""

    def __init__(self, n, h, L, tau, bc, time=0.0):
        self.n = tuple(int(v) for v in n)
        self._n = np.array(self.n, dtype=np.int32)
        self._h = np.array(h, dtype=np.float64)
        self._L = np.array(L, dtype=np.float64)
        self._bc = np.array(bc, dtype=np.int32)
        self.tau = float(tau)
        self.s = lib().opic_create(_ptr(self._n), _ptr(self._h), _ptr(self._L),
                                   self.tau, _ptr(self._bc), float(time))

    def close(self):
        if self.s:
            lib().opic_destroy(self.s)
            self.s = None

    def __del__(self):
        self.close()

    @property
    def nodes(self):
        return int(np.prod([v + 1 for v in self.n]))

    def set_fields(self, E, H):
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(-1)
        H = np.ascontiguousarray(H, dtype=np.float64).reshape(-1)
        assert E.size == 3 * self.nodes and H.size == 3 * self.nodes
        lib().opic_set_fields(self.s, _ptr(E), _ptr(H))

    def fields(self):
        E = np.empty(3 * self.nodes)
        H = np.empty(3 * self.nodes)
        lib().opic_get_fields(self.s, _ptr(E), _ptr(H))
        return E, H

    def set_markers(self, M):
        M = np.ascontiguousarray(M, dtype=MARKER_DTYPE)
        assert lib().opic_set_markers(self.s, _ptr(M), M.size, 64) == 0

    def marker_count(self):
        return int(lib().opic_marker_count(self.s))

    def markers(self):
        n = lib().opic_marker_count(self.s)
        M = np.zeros(n, dtype=MARKER_DTYPE)
        lib().opic_get_markers(self.s, _ptr(M), 64)
        return M

    def add_laser(self, params, lmap):
        A, T0, delay, dt, phi0, fwd, plane = params
        lmap = np.ascontiguousarray(lmap, dtype=np.float64)
        assert lib().opic_add_laser(self.s, A, T0, delay, dt, phi0, int(fwd),
                                    int(plane), _ptr(lmap)) == 0

    def add_gauss_spot(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        assert p.size == 11
        assert lib().opic_add_gauss_spot(self.s, _ptr(p)) == 0

    def add_tfsf(self, reg, mode, nframes, dataE, dataH):
        """dataE / dataH: 12 arrays (or None for an opened face) of the frames
        played onto E / H face slot f."""
        self._tfsf = [np.ascontiguousarray(x, dtype=np.float64) if x is not None else None
                      for x in list(dataE) + list(dataH)]
        arr = (ctypes.c_void_p * 12)
        pe = arr(*[x.ctypes.data if x is not None else None for x in self._tfsf[:12]])
        ph = arr(*[x.ctypes.data if x is not None else None for x in self._tfsf[12:]])
        r = np.ascontiguousarray(reg, dtype=np.int32)
        assert lib().opic_add_tfsf(self.s, _ptr(r), int(mode), int(nframes), pe, ph) == 0

    def set_vsp(self, alpha):
        assert lib().opic_set_vsp(self.s, float(alpha)) == 0

    def init(self):
        rc = lib().opic_init(self.s)
        assert rc == 0, rc

    def step(self, gauss=False):
        d = np.zeros(6)
        g = np.zeros(2)
        assert lib().opic_step(self.s, _ptr(d), _ptr(g) if gauss else None) == 0
        return d, (g if gauss else None)

    def J(self):
        J = np.empty(3 * int(np.prod([v + 5 for v in self.n])))
        lib().opic_get_J(self.s, _ptr(J))
        return J

    def rho(self):
        r = np.empty(int(np.prod([v + 5 for v in self.n])))
        lib().opic_rho(self.s, _ptr(r))
        return r

    @property
    def time(self):
        return lib().opic_time(self.s)


def marker_digest(M):
    b = np.ascontiguousarray(M).view(np.uint8).reshape(-1, 64)[:, :60]
    return hashlib.sha256(b.tobytes()).hexdigest()


def field_digest(F):
    return hashlib.sha256(np.ascontiguousarray(F).tobytes()).hexdigest()


def load_fixture(name):
    path = os.path.join(ROOT, "tests", "golden", name + ".npz")
    return dict(np.load(path, allow_pickle=False))


def tfsf_arrays(fx):
    """(reg, mode, nframes, dataE, dataH) of a fixture's TFSF recordings,
    split per face slot; opened faces are None."""
    nf = int(fx["tfsf_frames"][0])
    vols = fx["tfsf_vol"]
    offs = np.concatenate([[0], np.cumsum(vols * nf)])
    on = fx["tfsf_on"]
    dataE = [fx["tfsf_E"][offs[f]:offs[f + 1]] if on[f] else None for f in range(12)]
    dataH = [fx["tfsf_H"][offs[f]:offs[f + 1]] if on[f] else None for f in range(12)]
    return fx["tfsf_reg"], int(fx["tfsf_mode"][0]), nf, dataE, dataH


def oracle_from_fixture(fx):
    o = Oracle(fx["n"], fx["h"], fx["L"], fx["tau"][0], fx["bc"], 0.0)
    o.set_fields(fx["E0"], fx["H0"])
    o.set_markers(fx["M0"])
    if "laser_params" in fx:
        o.add_laser(fx["laser_params"], fx["laser_map"])
    if "spot_params" in fx:
        o.add_gauss_spot(fx["spot_params"])
    if "tfsf_reg" in fx:
        o.add_tfsf(*tfsf_arrays(fx))
    if "vsp" in fx:
        o.set_vsp(float(fx["vsp"][0]))
    o.init()
    return o
