"""RANMAR streams of the synthetic workloads.

The configurations draw their random numbers from RANMAR, the generator of
the reference (source/misc_ranmar.c:78-146), seeded by one (ij, kl) pair per
run.  The generator itself is the C host library's (mh_ranmar_init /
mh_ranmar in mandor3_amd/host/mandor_host.c, pinned by the reference's known
answer); this module only derives the seed pair from a run seed and turns
the uniform stream into the deviates a workload needs.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(PKG, "libmandor_host.so")

IJ_MAX, KL_MAX = 31328, 30081
_CHUNK = 1 << 24                 # doubles per call (the C length is an int)

_host = None


def _lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError("libmandor_host.so is not built (python -m mandor3_amd.build)")
        L = ctypes.CDLL(HOST_LIB)
        L.mh_ranmar_init.restype = ctypes.c_int
        L.mh_ranmar_init.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.mh_ranmar.restype = None
        L.mh_ranmar.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.mh_ranmar_normal.restype = None
        L.mh_ranmar_normal.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                       ctypes.c_double]
        _host = L
    return _host


class _State(ctypes.Structure):
    """mh_ranmar_t: the 97 lagged values, the arithmetic sequence, indices."""
    _fields_ = [("u", ctypes.c_double * 97), ("c", ctypes.c_double), ("cd", ctypes.c_double),
                ("cm", ctypes.c_double), ("i97", ctypes.c_int), ("j97", ctypes.c_int)]


def seed_pair(seed):
    """The (ij, kl) pair of a run seed: ij = seed mod 31329, kl = (seed div
    31329) mod 30082 -- every pair the generator accepts is reachable."""
    seed = int(seed)
    if seed < 0:
        raise ValueError("run seeds are non-negative")
    return seed % (IJ_MAX + 1), (seed // (IJ_MAX + 1)) % (KL_MAX + 1)


class Ranmar:
    """One RANMAR stream; uniform() continues the stream call after call."""

    def __init__(self, seed):
        self.ij, self.kl = seed_pair(seed)
        self._s = _State()
        if _lib().mh_ranmar_init(ctypes.byref(self._s), self.ij, self.kl) != 0:
            raise ValueError("RANMAR seeds (%d, %d) out of range" % (self.ij, self.kl))

    def uniform(self, n):
        """n doubles in [0, 1) with 24-bit mantissas."""
        out = np.empty(int(n))
        for a in range(0, out.size, _CHUNK):
            b = min(a + _CHUNK, out.size)
            _lib().mh_ranmar(ctypes.byref(self._s),
                             out[a:].ctypes.data_as(ctypes.c_void_p), b - a)
        return out

    def normal(self, n, sigma=1.0):
        """n deviates of N(0, sigma^2): the Box-Muller transform of pairs of
        the stream's uniforms (G. E. P. Box and M. E. Muller, 1958), done by
        mh_ranmar_normal."""
        out = np.empty(int(n))
        _lib().mh_ranmar_normal(ctypes.byref(self._s), out.ctypes.data_as(ctypes.c_void_p),
                                out.size, float(sigma))
        return out
