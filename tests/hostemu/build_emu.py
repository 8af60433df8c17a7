"""Build the library's HIP sources for the host (test infrastructure).

The kernels in ccake_amd/csrc are compiled with g++ against the device-model
emulation in tests/hostemu/hip/hip_runtime.h, with AddressSanitizer, into
tests/hostemu/_build/libccake_emu.so.  Device buffers are allocated with
their exact sizes (CCAKE_ALLOC_PAD=0), so every access past a buffer, and
every out-of-bounds access to a work-item's private arrays, is reported.

Run a program against it with
    LD_PRELOAD=$(g++ -print-file-name=libasan.so) \\
    CCAKE_AMD_LIBRARY=tests/hostemu/_build/libccake_emu.so python ...
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "ccake_amd", "csrc")
OUT_DIR = os.path.join(HERE, "_build")
SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def flags(asan=True):
    f = ["-std=c++20", "-O1", "-g", "-fPIC", "-pthread", "-ffp-contract=off", "-fno-omit-frame-pointer",
         "-DCCAKE_HOST_EMU", "-DCCAKE_ALLOC_PAD=0", "-Wno-unknown-pragmas", "-Wno-attributes",
         "-I", HERE, "-I", os.path.join(ROOT, "include")]
    if asan:
        f += ["-fsanitize=address"]
    return f


def _compile(src, asan):
    # objects carry the process id, so concurrent builds (pytest-xdist
    # workers) never link each other's half-written files
    obj = os.path.join(OUT_DIR, os.path.splitext(os.path.basename(src))[0]
                       + (".asan" if asan else "") + f".{os.getpid()}.o")
    path = src if os.path.isabs(src) else os.path.join(CSRC, src)
    cmd = ["g++"] + flags(asan) + ["-x", "c++", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build(asan=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, asan),
                           SOURCES + [os.path.join(HERE, "emu_runtime.cpp")]))
    out = os.path.join(OUT_DIR, "libccake_emu.so" if asan else "libccake_emu_noasan.so")
    tmp = f"{out}.{os.getpid()}.tmp"
    cmd = ["g++", "-shared", "-pthread"] + (["-fsanitize=address"] if asan else []) + ["-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    for o in objs:
        os.remove(o)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    # the same library under the product's name, for the C++ drivers of
    # ccake_amd/bin run with LD_LIBRARY_PATH pointing here (lib, or
    # lib_noasan for the build without AddressSanitizer)
    lib_dir = os.path.join(OUT_DIR, "lib" if asan else "lib_noasan")
    os.makedirs(lib_dir, exist_ok=True)
    ltmp = os.path.join(lib_dir, f"libccake_amd.so.{os.getpid()}.tmp")
    shutil.copyfile(tmp, ltmp)
    os.replace(ltmp, os.path.join(lib_dir, "libccake_amd.so"))
    os.replace(tmp, out)
    return out


def asan_runtime():
    return subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()


if __name__ == "__main__":
    print(build(asan="--no-asan" not in sys.argv))
