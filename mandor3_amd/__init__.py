"""mandor3_amd -- MI355X-native PIC time step of Mandor behind a C ABI.

The compute path is libmpic.so (HIP, gfx950; see include/mpic.h).  This
module is a thin ctypes binding used by the tests and the benchmark; the
host driver of the product is C (mandor3_amd/host).  There is no CPU
fallback: if libmpic.so is missing or no HIP device is present, the calls
raise.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmpic.so")

BC_PERIODIC, BC_MIRROR, BC_OPEN = 0, 1, 2
PROBE_AT_MAX = 16384          # MPIC_PROBE_AT_MAX of include/mpic.h
FP64, FP32_SPLIT = 0, 1

MARKER_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
                         ("vx", "<f8"), ("vy", "<f8"), ("vz", "<f8"),
                         ("rho", "<f4"), ("qDivM", "<f4"), ("id", "<i4"),
                         ("pad", "<i4")])
# the reference's marker_t without the tracer id (source/type_marker.h:12-25)
MARKER56_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
                           ("vx", "<f8"), ("vy", "<f8"), ("vz", "<f8"),
                           ("rho", "<f4"), ("qDivM", "<f4")])


class Config(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("h", ctypes.c_double * 3),
                ("L", ctypes.c_double * 3), ("tau", ctypes.c_double),
                ("bc", ctypes.c_int * 6), ("time", ctypes.c_double),
                ("precision", ctypes.c_int), ("max_markers", ctypes.c_long),
                ("device", ctypes.c_int)]


_P, _I, _D, _LG = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
# (restype, argtypes) of every entry point of include/mpic.h
SIGNATURES = {
    "mpic_last_error": (ctypes.c_char_p, []),
    "mpic_device_count": (_I, []),
    "mpic_create": (_I, [ctypes.POINTER(Config), ctypes.POINTER(_P)]),
    "mpic_destroy": (_I, [_P]),
    "mpic_footprint": (_I, [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_size_t)]),
    "mpic_set_fields": (_I, [_P, _P, _P]),
    "mpic_get_fields": (_I, [_P, _P, _P]),
    "mpic_get_J": (_I, [_P, _P]),
    "mpic_markers_load": (_I, [_P, _P, _LG, _I]),
    "mpic_markers_append": (_I, [_P, _P, _LG, _I]),
    "mpic_markers_save": (_I, [_P, _P, ctypes.POINTER(_LG), _I]),
    "mpic_marker_count": (_LG, [_P]),
    "mpic_add_laser": (_I, [_P, _D, _D, _D, _D, _D, _I, _I, _P]),
    "mpic_add_gauss_spot": (_I, [_P, _P]),
    "mpic_add_tfsf": (_I, [_P, _P, _I, _I, _P, _P]),
    "mpic_set_vsp": (_I, [_P, _D]),
    "mpic_init": (_I, [_P]),
    "mpic_em_h_half": (_I, [_P]),
    "mpic_em_energy": (_I, [_P, _P, _P]),
    "mpic_push_deposit": (_I, [_P]),
    "mpic_temperature": (_I, [_P, _P, _P, _P]),
    "mpic_em_h_step": (_I, [_P]),
    "mpic_em_e_start": (_I, [_P]),
    "mpic_j_finish": (_I, [_P]),
    "mpic_em_e_finish": (_I, [_P]),
    "mpic_step": (_I, [_P, _P]),
    "mpic_advance": (_I, [_P, _I]),
    "mpic_diag_history": (_I, [_P, _P, _I, _P]),
    "mpic_set_graphs": (_I, [_P, _I]),
    "mpic_batch_advance": (_I, [_P, _I, _I]),
    "mpic_autotune": (_I, [_P, _I, _P]),
    "mpic_sync": (_I, [_P]),
    "mpic_time": (_D, [_P]),
    "mpic_rho": (_I, [_P, _P]),
    "mpic_node_fields": (_I, [_P, _I, _P]),
    "mpic_tfsf_record": (_I, [_P, _P, _I]),
    "mpic_tfsf_recorded": (_I, [_P, _P, _P, _P]),
    "mpic_set_probes": (_I, [_P, _P, _I]),
    "mpic_probe_sample": (_I, [_P, _P]),
    "mpic_probe_sample_at": (_I, [_P, _P, _I, _P]),
    "mpic_gauss_before": (_I, [_P]),
    "mpic_gauss_after": (_I, [_P, _P]),
    "mpic_sort": (_I, [_P, _P]),
    "mpic_marker_cells": (_I, [_P, _P, ctypes.POINTER(_LG)]),
    "mpic_last_timing": (_I, [_P, _P, _P, _P]),
}


class MpicError(RuntimeError):
    pass


_lib = None


def lib():
    """Loads libmpic.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MpicError("libmpic.so is not built (python -m mandor3_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc, what):
    if rc != 0:
        msg = lib().mpic_last_error()
        raise MpicError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


def batch_advance(runs, nsteps):
    """Advances an ensemble of runs together, step by step across the runs,
    each replaying its captured step on its own stream."""
    arr = (ctypes.c_void_p * len(runs))(*[r.handle.value for r in runs])
    _check(lib().mpic_batch_advance(arr, len(runs), int(nsteps)), "mpic_batch_advance")


def device_count():
    return lib().mpic_device_count()


class Run:
    """One simulation resident on the device (the reference's process state
    of source/core/main.c, per run)."""

    def __init__(self, n, h, L, tau, bc, time=0.0, precision=FP64,
                 max_markers=1 << 20, device=0):
        c = Config()
        for a in range(3):
            c.n[a] = int(n[a])
            c.h[a] = float(h[a])
            c.L[a] = float(L[a])
        for b in range(6):
            c.bc[b] = int(bc[b])
        c.tau = float(tau)
        c.time = float(time)
        c.precision = int(precision)
        c.max_markers = int(max_markers)
        c.device = int(device)
        self.cfg = c
        self.n = tuple(int(v) for v in n)
        self.h = tuple(float(v) for v in h)
        self.tau = float(tau)
        self.handle = ctypes.c_void_p()
        _check(lib().mpic_create(ctypes.byref(c), ctypes.byref(self.handle)),
               "mpic_create")

    def close(self):
        if self.handle:
            lib().mpic_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nodes(self):
        return int(np.prod([v + 1 for v in self.n]))

    def set_fields(self, E, H):
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(-1)
        H = np.ascontiguousarray(H, dtype=np.float64).reshape(-1)
        if E.size != 3 * self.nodes or H.size != 3 * self.nodes:
            raise ValueError("field blocks must hold 3*(n+1)^3 doubles")
        _check(lib().mpic_set_fields(self.handle, _ptr(E), _ptr(H)), "mpic_set_fields")

    def fields(self):
        E = np.empty(3 * self.nodes)
        H = np.empty(3 * self.nodes)
        _check(lib().mpic_get_fields(self.handle, _ptr(E), _ptr(H)), "mpic_get_fields")
        return E, H

    def J(self):
        J = np.empty(3 * int(np.prod([v + 5 for v in self.n])))
        _check(lib().mpic_get_J(self.handle, _ptr(J)), "mpic_get_J")
        return J

    def load_markers(self, M, recsize=64):
        """64-byte records (with id) or 56-byte records (no id: the run
        numbers the markers 1..N in record order)."""
        dt = MARKER_DTYPE if recsize == 64 else MARKER56_DTYPE
        if M.dtype != dt:
            A = np.zeros(M.shape, dtype=dt)
            for f in dt.names:
                if f in M.dtype.names:
                    A[f] = M[f]
            M = A
        M = np.ascontiguousarray(M)
        _check(lib().mpic_markers_load(self.handle, _ptr(M), M.size, recsize),
               "mpic_markers_load")

    def append_markers(self, M):
        """Appends records after those loaded (before init), as plasma_load
        reads a record buffer by buffer."""
        M = np.ascontiguousarray(M, dtype=MARKER_DTYPE)
        _check(lib().mpic_markers_append(self.handle, _ptr(M), M.size, 64),
               "mpic_markers_append")

    def markers(self, recsize=64):
        cap = max(int(lib().mpic_marker_count(self.handle)), 1)
        M = np.zeros(cap, dtype=MARKER_DTYPE if recsize == 64 else MARKER56_DTYPE)
        n = ctypes.c_long(cap)
        _check(lib().mpic_markers_save(self.handle, _ptr(M), ctypes.byref(n), recsize),
               "mpic_markers_save")
        return M[:n.value]

    def marker_count(self):
        return int(lib().mpic_marker_count(self.handle))

    def add_laser(self, params, lmap):
        A, T0, delay, dt, phi0, fwd, plane = params
        lmap = np.ascontiguousarray(lmap, dtype=np.float64)
        _check(lib().mpic_add_laser(self.handle, A, T0, delay, dt, phi0, int(fwd),
                                    int(plane), _ptr(lmap)), "mpic_add_laser")

    def set_vsp(self, alpha):
        """The VSP (beta, alpha, beta) convolution of J and rho (1: off)."""
        _check(lib().mpic_set_vsp(self.handle, float(alpha)), "mpic_set_vsp")

    def init(self):
        _check(lib().mpic_init(self.handle), "mpic_init")

    def step(self, diag=True):
        if not diag:
            _check(lib().mpic_step(self.handle, None), "mpic_step")
            return None
        d = np.zeros(6)
        _check(lib().mpic_step(self.handle, _ptr(d)), "mpic_step")
        return d

    def stage(self, name):
        _check(getattr(lib(), "mpic_" + name)(self.handle), "mpic_" + name)

    def advance(self, nsteps):
        _check(lib().mpic_advance(self.handle, int(nsteps)), "mpic_advance")

    def diag_history(self, max_rows=1024):
        """Rows {Time, W_E, W_M, WTx, WTy, WTz} of the last asynchronous
        steps (mpic_diag_history), oldest first."""
        rows = np.zeros((max_rows, 6))
        n = np.zeros(1, dtype=np.int32)
        _check(lib().mpic_diag_history(self.handle, _ptr(rows), int(max_rows), _ptr(n)),
               "mpic_diag_history")
        return rows[:int(n[0])]

    def add_gauss_spot(self, params):
        """Hard Gaussian spot: the 11 values of .gaussSpot.cfg in file order."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        if p.size != 11:
            raise ValueError("11 spot parameters expected")
        _check(lib().mpic_add_gauss_spot(self.handle, _ptr(p)), "mpic_add_gauss_spot")

    def add_tfsf(self, reg, mode, nframes, dataE, dataH):
        """TFSF interface playback: reg = (imin, imax, jmin, jmax, kmin, kmax),
        mode 0 forward / 2 backward, dataE / dataH = 12 arrays each (None for
        an opened face) of the frames played onto E / H face slot f."""
        self._tfsf = [np.ascontiguousarray(x, dtype=np.float64) if x is not None else None
                      for x in list(dataE) + list(dataH)]
        arr = ctypes.c_void_p * 12
        pe = arr(*[x.ctypes.data if x is not None else None for x in self._tfsf[:12]])
        ph = arr(*[x.ctypes.data if x is not None else None for x in self._tfsf[12:]])
        r = np.ascontiguousarray(reg, dtype=np.int32)
        _check(lib().mpic_add_tfsf(self.handle, _ptr(r), int(mode), int(nframes), pe, ph),
               "mpic_add_tfsf")

    def set_graphs(self, on=True):
        """Replay captured HIP graphs of the step in advance()."""
        _check(lib().mpic_set_graphs(self.handle, 1 if on else 0), "mpic_set_graphs")

    def autotune(self, steps=3):
        """Times `steps` real steps per candidate launch shape of the FP32
        push (markers per work item x wavefronts per workgroup x workgroups)
        and keeps the fastest; returns (workgroups, wavefronts per
        workgroup, markers per item, 0 = one item per tile)."""
        c = np.zeros(3, dtype=np.int32)
        _check(lib().mpic_autotune(self.handle, int(steps), _ptr(c)), "mpic_autotune")
        return int(c[0]), int(c[1]), int(c[2])

    def sync(self):
        _check(lib().mpic_sync(self.handle), "mpic_sync")

    def timing(self):
        """(push kernel, rest of the deposit, rest of the step) in ms summed
        over the steps of the last advance()."""
        a, b, c = np.zeros(1), np.zeros(1), np.zeros(1)
        _check(lib().mpic_last_timing(self.handle, _ptr(a), _ptr(b), _ptr(c)), "mpic_last_timing")
        return float(a[0]), float(b[0]), float(c[0])

    @property
    def time(self):
        return lib().mpic_time(self.handle)

    def rho(self):
        r = np.empty(int(np.prod([v + 5 for v in self.n])))
        _check(lib().mpic_rho(self.handle, _ptr(r)), "mpic_rho")
        return r

    def tfsf_record(self, reg, max_frames):
        r = np.ascontiguousarray(reg, dtype=np.int32)
        _check(lib().mpic_tfsf_record(self.handle, _ptr(r), int(max_frames)), "mpic_tfsf_record")

    def tfsf_recorded(self):
        """(frames, [24 arrays of shape (frames, face nodes)]): E faces then H faces."""
        nf = np.zeros(1, np.int32)
        vol = np.zeros(24, np.int64)
        _check(lib().mpic_tfsf_recorded(self.handle, _ptr(nf), _ptr(vol), None), "mpic_tfsf_recorded")
        out = np.empty(int(nf[0] * vol.sum()))
        _check(lib().mpic_tfsf_recorded(self.handle, _ptr(nf), _ptr(vol), _ptr(out)), "mpic_tfsf_recorded")
        faces, at = [], 0
        for v in vol:
            faces.append(out[at:at + nf[0] * v].reshape(nf[0], v))
            at += nf[0] * v
        return int(nf[0]), faces

    def node_fields(self, what):
        """Node values of E (0), H (1) or J (2) on the nodes [0, n], as the
        Tecplot records hold them: shape ((n0+1)(n1+1)(n2+1), 3)."""
        out = np.empty(3 * self.nodes)
        _check(lib().mpic_node_fields(self.handle, int(what), _ptr(out)), "mpic_node_fields")
        return out.reshape(-1, 3)

    def set_probes(self, ijk):
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        _check(lib().mpic_set_probes(self.handle, _ptr(ijk), ijk.shape[0]), "mpic_set_probes")
        self._nprobes = ijk.shape[0]

    def probe_sample(self):
        out = np.empty(6 * getattr(self, "_nprobes", 0))
        _check(lib().mpic_probe_sample(self.handle, _ptr(out)), "mpic_probe_sample")
        return out.reshape(-1, 6)

    def probe_sample_at(self, ijk):
        """Samples at the nodes ijk without allocating device memory
        (mpic_probe_sample_at: at most PROBE_AT_MAX probes)."""
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        out = np.empty(6 * ijk.shape[0])
        _check(lib().mpic_probe_sample_at(self.handle, _ptr(ijk), ijk.shape[0], _ptr(out)),
               "mpic_probe_sample_at")
        return out.reshape(-1, 6)

    def gauss_before(self):
        _check(lib().mpic_gauss_before(self.handle), "mpic_gauss_before")

    def gauss_after(self):
        g = np.zeros(2)
        _check(lib().mpic_gauss_after(self.handle, _ptr(g)), "mpic_gauss_after")
        return g

    def energy(self):
        a, b = np.zeros(1), np.zeros(1)
        _check(lib().mpic_em_energy(self.handle, _ptr(a), _ptr(b)), "mpic_em_energy")
        return float(a[0]), float(b[0])

    def temperature(self):
        w = np.zeros(3)
        _check(lib().mpic_temperature(self.handle, _ptr(w[0:]), _ptr(w[1:]), _ptr(w[2:])),
               "mpic_temperature")
        return w

    def cells(self):
        cap = max(self.marker_count(), 1)
        c = np.zeros(cap, dtype=np.int32)
        n = ctypes.c_long(cap)
        _check(lib().mpic_marker_cells(self.handle, _ptr(c), ctypes.byref(n)),
               "mpic_marker_cells")
        return c[:n.value]

    def sort(self, with_counts=False):
        counts = np.zeros(int(np.prod(self.n)), dtype=np.int32) if with_counts else None
        _check(lib().mpic_sort(self.handle, _ptr(counts) if with_counts else None),
               "mpic_sort")
        return counts
