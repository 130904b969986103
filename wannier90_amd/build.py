"""Build libw90gpu.so (HIP, gfx950) and the Fortran host in-tree.

    python -m wannier90_amd.build            # everything
    python -m wannier90_amd.build --hip      # only the HIP library

Outputs go to wannier90_amd/lib/ (git-ignored, shipped with the tree).
"""
import argparse
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIBDIR = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("W90_FC", "/opt/rocm/bin/amdflang")
ARCH = os.environ.get("W90_ARCH", "gfx950")

# Packed FP32 instructions stay off for all device code.
HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
    f"-I{INCLUDE}",
]

HIP_SOURCES = ["w90gpu.hip", "w90dis.hip"]
CPP_SOURCES = ["kmesh.cpp", "w90io.cpp"]

LIB_NAME = "libw90gpu.so"
# the host-only part (k-mesh search, .mmn/.amn readers) as a library of its
# own without HIP code, for processes that must not start the GPU runtime
# (problem generators in worker processes)
HOST_LIB_NAME = "libw90io.so"


def _run(cmd, cwd=None):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip(verbose=True):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CPP_SOURCES:
        obj = os.path.join(objdir, src.replace(".cpp", ".o"))
        _run(["g++", "-O2", "-fPIC", "-std=c++17", "-pthread", "-Wall", f"-I{INCLUDE}", "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(LIBDIR, LIB_NAME)
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-pthread", "-o", tmp])
    os.replace(tmp, out)
    host = os.path.join(LIBDIR, HOST_LIB_NAME)
    _run(["g++", "-shared", "-fPIC", *objs[len(HIP_SOURCES):], "-pthread", "-o", host + ".tmp"])
    os.replace(host + ".tmp", host)
    return out


def build_fortran():
    """Fortran host: ISO_C_BINDING interface, library-interface mirror and
    the wannier90.x-style driver, linked against libw90gpu.so."""
    if not os.path.isdir(FSRC):
        return None
    srcs = sorted(f for f in os.listdir(FSRC) if f.endswith(".F90"))
    if not srcs:
        return None
    objdir = os.path.join(LIBDIR, "fobj")
    os.makedirs(objdir, exist_ok=True)
    order = [
        "w90gpu_c.F90", "w90gpu_io.F90", "w90gpu_library.F90", "w90gpu_c_interface.F90",
        "wannier90_gpu.F90",
    ]
    order = [f for f in order if f in srcs] + [f for f in srcs if f not in order]
    objs = []
    prog = None
    for f in order:
        if f == "wannier90_gpu.F90":
            prog = f
            continue
        obj = os.path.join(objdir, f.replace(".F90", ".o"))
        _run([FC, "-O2", "-fPIC", "-module-dir", objdir, "-I", objdir, "-c",
              os.path.join(FSRC, f), "-o", obj])
        objs.append(obj)
    lib = os.path.join(LIBDIR, "libw90host.so")
    rpath = "-Wl,-rpath,$ORIGIN"
    _run([FC, "-shared", *objs, f"-L{LIBDIR}", "-lw90gpu", rpath, "-o", lib])
    exe = None
    if prog:
        exe = os.path.join(LIBDIR, "wannier90_gpu.x")
        _run([FC, "-O2", "-module-dir", objdir, "-I", objdir, os.path.join(FSRC, prog),
              f"-L{LIBDIR}", "-lw90host", "-lw90gpu", rpath, "-o", exe])
    return exe


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--hip", action="store_true", help="only the HIP library")
    args = ap.parse_args(argv)
    build_hip()
    if not args.hip:
        build_fortran()


if __name__ == "__main__":
    main()
