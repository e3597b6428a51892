"""Build the gfx950 shared library monofonic_amd/libmfx.so in-tree with hipcc,
and the host-only output library monofonic_amd/libmfx_io.so with g++.

Every HIP/C++ source under monofonic_amd/csrc is compiled to an object in
monofonic_amd/_build (in parallel) and linked into libmfx.so.  Device code is
compiled for gfx950 only, with packed FP32 instructions disabled.
libmfx_io.so (monofonic_amd/io, include/mfx_io.h) links the HDF5 C library
found under HDF5_DIR (default /opt/conda).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmfx.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")
IO_SRC = os.path.join(HERE, "io", "mfx_io.cpp")
IO_LIB = os.path.join(HERE, "libmfx_io.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
ARCH = "gfx950"
CXXFLAGS = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-variable", "-Wno-unused-parameter"]
DEVFLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + CXXFLAGS + DEVFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 1)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, r in ex.map(_compile, srcs):
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError(f"hipcc failed on {os.path.basename(src)}")
            if verbose and (r.stderr.strip()):
                sys.stderr.write(r.stderr)
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libmfx.so failed")
    os.replace(tmp, LIB)
    build_io()
    return LIB


def build_io():
    """libmfx_io.so: the HDF5 output plugins (host code only)"""
    inc = os.path.join(HDF5_DIR, "include")
    lib = os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        raise RuntimeError(f"hdf5.h not found under {inc}: set HDF5_DIR")
    tmp = IO_LIB + ".tmp"
    cmd = [CXX, "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", f"-I{INCLUDE}", f"-I{inc}", IO_SRC, "-o", tmp,
           f"-L{lib}", "-lhdf5", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build of libmfx_io.so failed")
    os.replace(tmp, IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
