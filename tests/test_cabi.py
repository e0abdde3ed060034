"""The C-ABI library loads and exports every entry point include/mpic.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "mpic.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpic_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header():
    import mandor3_amd
    lib = ctypes.CDLL(mandor3_amd.LIB_PATH)
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import mandor3_amd
    lib = mandor3_amd.lib()
    for n in declared():
        assert getattr(lib, n).restype is not None or n in (), n
