"""Builds libmpic.so (HIP, gfx950) in-tree from mandor3_amd/csrc.

    python -m mandor3_amd.build

Every HIP source is compiled for gfx950 with contraction off (the FP64
plain form reproduces the CPU arithmetic bit for bit), hardware FP atomics,
and packed FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmpic.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
          "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
          "-Wno-unused-result", "-Wno-unused-value"]
# per-file contraction mode: the FP64 plain path must not fuse mul+add
# (bit-exact with the reference); the FP32 split push may
FP_MODE = {"default": ["-ffp-contract=off"],
           "mpic_push32.hip": ["-ffp-contract=fast"]}


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src)[:-4] + ".o")
    flags = FP_MODE.get(os.path.basename(src), FP_MODE["default"])
    cmd = [HIPCC] + COMMON + flags + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return src, obj, r.returncode, r.stdout.decode(errors="replace")


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        res = list(ex.map(_compile, srcs))
    objs = []
    for src, obj, rc, log in res:
        # the host pass does not know the packed-fp32 feature: harmless
        lines = [l for l in log.splitlines() if "packed-fp32-ops" not in l
                 and l.strip() and "is not a recognized feature" not in l]
        if rc != 0:
            sys.stderr.write(log)
            raise RuntimeError("hipcc failed on " + src)
        if verbose and lines:
            sys.stderr.write("\n".join(lines) + "\n")
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    build_host()
    return LIB


HOST = os.path.join(PKG, "host")
HOST_LIB = os.path.join(PKG, "libmandor_host.so")
CORE_EXE = os.path.join(PKG, "mandor_core")
SETUP_EXE = os.path.join(PKG, "mandor_setup")
CC = os.environ.get("CC", "gcc")
CFLAGS = ["-O2", "-std=gnu99", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"),
          "-I" + HOST]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("failed: " + " ".join(cmd))


def build_host():
    """The C host side: libmandor_host.so (formats, RANMAR), mandor_setup
    and mandor_core (linked to libmpic.so, found next to it)."""
    src = os.path.join(HOST, "mandor_host.c")
    obj = os.path.join(OUT, "mandor_host.o")
    _run([CC] + CFLAGS + ["-fPIC", "-c", src, "-o", obj])
    _run([CC, "-shared", obj, "-lm", "-o", HOST_LIB])
    _run([CC] + CFLAGS + [os.path.join(HOST, "mandor_setup.c"), obj, "-lm", "-o", SETUP_EXE])
    _run([CC] + CFLAGS + [os.path.join(HOST, "mandor_core.c"), obj, "-L" + PKG, "-lmpic",
                          "-Wl,-rpath,$ORIGIN", "-lm", "-o", CORE_EXE])


if __name__ == "__main__":
    print(build(verbose=True))
