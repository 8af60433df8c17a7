"""Build the gfx950 HIP library `ccake_amd/libccake_amd.so` in-tree.

Every translation unit under csrc/ is compiled with hipcc for gfx950 only
and linked into one shared object whose exported symbols are exactly the
C-ABI of include/ccake_amd.h.  Packed FP32 instructions are disabled for
all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libccake_amd.so")
OBJ = os.path.join(HERE, "_build")
SOURCES = ["api.hip", "k_neighbors.hip", "k_smooth.hip", "k_eos.hip", "k_grad.hip", "k_eom.hip", "k_misc.hip", "k_diff.hip",
           "k_pairs32.hip", "k_milne3.hip"]
ARCH = os.environ.get("CCAKE_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
            "-I", os.path.join(ROOT, "include")]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "unused-command-line-argument" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{msg}")
    return obj, msg


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, msg in results:
        if verbose and msg.strip():
            print(msg)
    objs = [o for o, _ in results]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    build_host_programs()
    return OUT


HOST = os.path.join(HERE, "host")
BIN = os.path.join(HERE, "bin")
HOST_PROGRAMS = ["gubser_run", "ic_dump", "bsqhydro_run"]


HDF5_PREFIX = os.environ.get("CCAKE_HDF5_PREFIX", "/opt/conda")


HDF5_LIB = os.path.join(HERE, "libccake_hdf5.so")


def build_hdf5_shim():
    """libccake_hdf5.so: the HDF5 input and output of the C++ call surface
    (host/ccake_hdf5.cpp) on the HDF5 C library, which only this library
    links (with a run-time path to HDF5's directory): a program linking it
    keeps finding its C++ runtime in the system directories.  Returns the
    compile and link flags for the programs, or none when HDF5 is not
    installed."""
    inc, lib = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(os.path.join(lib, "libhdf5.so"))):
        return [], []
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-shared", "-fPIC", "-I", HOST, "-I", inc,
           os.path.join(HOST, "ccake_hdf5.cpp"), "-o", HDF5_LIB + ".tmp", os.path.join(lib, "libhdf5.so"),
           "-Wl,-rpath," + lib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed on ccake_hdf5.cpp:\n" + r.stdout + r.stderr)
    os.replace(HDF5_LIB + ".tmp", HDF5_LIB)
    return ["-DCCAKE_WITH_HDF5"], ["-lccake_hdf5"]


def build_host_programs():
    """C++ drivers of the host call surface (ccake_amd/host), linked against
    the library (and libccake_hdf5.so when HDF5 is installed) with an rpath
    to them."""
    os.makedirs(BIN, exist_ok=True)
    h5c, h5l = build_hdf5_shim()
    for prog in HOST_PROGRAMS:
        out = os.path.join(BIN, prog)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", HOST] + h5c + [
               os.path.join(HOST, prog + ".cpp"), "-o", out + ".tmp", "-L", HERE, "-lccake_amd"] + h5l + [
               "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed on {prog}:\n" + r.stdout + r.stderr)
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
