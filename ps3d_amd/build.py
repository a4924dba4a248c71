"""Builds the HIP library libps3d_gpu.so (gfx950) in-tree with hipcc, and the
Fortran host program ps3d_driver on top of it with amdflang.

Usage: python -m ps3d_amd.build [--verbose]
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libps3d_gpu.so")

SOURCES = ["zpass.hip", "zwave.hip", "linepass.hip", "closure.hip", "stats.hip", "ps3d_gpu.hip", "ncfile.cpp"]
HEADERS = ["common.hpp", "kernels.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PS3D_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # keep hipcc from forming packed FP32 instructions (device code is FP64 anyway)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"),
]


def _compile(src, verbose=False, extra=()):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + CFLAGS + list(extra) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)
    return obj


FC = os.environ.get("PS3D_FC", "/opt/rocm/bin/amdflang")
FSRC = os.path.join(HERE, "fortran")
DRIVER = os.path.join(LIBDIR, "ps3d_driver")


def build_driver(verbose=False):
    """The Fortran host program (ps3d_amd/fortran): the iso_c_binding module and
    the ps3d driver, linked against libps3d_gpu.so found next to it."""
    fb = os.path.join(BUILD, "fortran")
    os.makedirs(fb, exist_ok=True)
    objs = []
    for src in ("ps3d_gpu.f90", "ps3d_driver.f90"):
        obj = os.path.join(fb, os.path.splitext(src)[0] + ".o")
        cmd = [FC, "-O2", "-module-dir", fb, "-c", os.path.join(FSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("%s failed for %s:\n%s\n%s" % (FC, src, " ".join(cmd), r.stderr[-4000:]))
        if verbose and r.stderr.strip():
            sys.stderr.write(r.stderr)
        objs.append(obj)
    tmp = DRIVER + ".tmp"
    cmd = [FC, "-o", tmp] + objs + ["-L" + LIBDIR, "-lps3d_gpu", "-Wl,-rpath,$ORIGIN",
                                    "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, DRIVER)
    return DRIVER


def build(verbose=False, extra=()):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, extra), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, LIB)
    build_driver(verbose)
    return LIB


if __name__ == "__main__":
    v = "--verbose" in sys.argv
    extra = ["-Rpass-analysis=kernel-resource-usage"] if "--resource-usage" in sys.argv else []
    print(build(verbose=v or bool(extra), extra=extra))
