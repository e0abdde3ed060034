"""The C-ABI library loads and exports every entry point include/mpic.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "mpic.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpic_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_header():
    import mandor3_amd
    lib = ctypes.CDLL(mandor3_amd.LIB_PATH)
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    """The binding's signature table names exactly the header's entry
    points, and the entry points that return something other than int are
    bound with their C return types."""
    import mandor3_amd
    sig = mandor3_amd.SIGNATURES
    names = set(declared())
    assert set(sig) == names, (sorted(names - set(sig)), sorted(set(sig) - names))
    assert sig["mpic_time"][0] is ctypes.c_double
    assert sig["mpic_marker_count"][0] is ctypes.c_long
    assert sig["mpic_last_error"][0] is ctypes.c_char_p
    lib = mandor3_amd.lib()
    for n in names:
        assert getattr(lib, n).restype is sig[n][0], n
        assert list(getattr(lib, n).argtypes) == list(sig[n][1]), n
    ret = {}
    txt = open(os.path.join(ROOT, "include", "mpic.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    for t, n in re.findall(r"^\s*(int|long|double|const char \*)\s*(mpic_[A-Za-z0-9_]+)\s*\(",
                           txt, flags=re.M):
        ret[n] = t.strip()
    assert set(ret) == names
    cmap = {"int": ctypes.c_int, "long": ctypes.c_long, "double": ctypes.c_double,
            "const char *": ctypes.c_char_p}
    for n, t in ret.items():
        assert sig[n][0] is cmap[t], (n, t)


def _config(n=(16, 16, 16), h=0.05, tau=0.01, precision=0, max_markers=1000):
    import mandor3_amd
    c = mandor3_amd.Config()
    for a in range(3):
        c.n[a] = n[a]
        c.h[a] = h
        c.L[a] = n[a] * h
    c.tau = tau
    for b in range(6):
        c.bc[b] = 0
    c.precision = precision
    c.max_markers = max_markers
    return c


def test_create_rejects_bad_configs_before_touching_a_device():
    """mpic_create's argument checks (the reference's ENSURE/DIE cases and
    the split form's limits) return MPIC_ERR_ARG with a message before any
    HIP call, so they run here; the marker entry points refuse a null run."""
    import mandor3_amd
    lib = mandor3_amd.lib()
    run = ctypes.c_void_p()
    cases = [
        (_config(n=(6, 16, 16)), ">= 7"),                          # core region
        (_config(tau=0.05), "unstable"),                           # Courant
        (_config(precision=7), "precision"),
        (_config(n=(2048, 2048, 1024), h=0.05, tau=0.01, precision=1), "2^29"),
        (_config(max_markers=1 << 31), "2^31"),
    ]
    for cfg, words in cases:
        rc = lib.mpic_create(ctypes.byref(cfg), ctypes.byref(run))
        assert rc == -1, (words, rc)
        assert words in lib.mpic_last_error().decode(), lib.mpic_last_error()
    assert lib.mpic_create(None, ctypes.byref(run)) == -1
    assert lib.mpic_markers_load(None, None, 0, 64) == -1
    assert lib.mpic_markers_append(None, None, 0, 64) == -1


def test_probe_scratch_limit_matches_the_header():
    """The binding's PROBE_AT_MAX is the header's MPIC_PROBE_AT_MAX, and the
    no-allocation probe call refuses a null run before touching a device."""
    import mandor3_amd
    txt = open(os.path.join(ROOT, "include", "mpic.h")).read()
    m = re.search(r"#define\s+MPIC_PROBE_AT_MAX\s+(\d+)", txt)
    assert m and int(m.group(1)) == mandor3_amd.PROBE_AT_MAX
    lib = mandor3_amd.lib()
    assert lib.mpic_probe_sample_at(None, None, 0, None) == -1
