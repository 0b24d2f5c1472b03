"""Build libalebgk_hip.so (gfx950) in-tree with hipcc, and the host library.

    python -m alebgk4paper_amd.build        # or __graft_entry__.build()

Every HIP source is compiled to an object in parallel, then linked into
alebgk4paper_amd/libalebgk_hip.so. Packed FP32 instructions are disabled for
all device code (-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libalebgk_hip.so")
HOSTLIB = os.path.join(PKG, "libalebgk_host.so")
CLI = os.path.join(PKG, "bin", "alebgk_amd")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ALEBGK_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    return obj


def build_hip(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_host():
    """Host C++ library and CLI (cases, lattice, checkpoint, time loop)."""
    if not os.path.isdir(HOST):
        return None
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    if not srcs:
        return None
    lib_srcs = [s for s in srcs if os.path.basename(s) != "main.cpp"]
    # no -march and no contraction: setup arithmetic rounds as the reference's
    flags = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-ffp-contract=off",
             "-I", os.path.join(ROOT, "include"), "-I", HOST]
    tmp = HOSTLIB + ".tmp"
    _run(["g++"] + flags + ["-shared", "-o", tmp] + lib_srcs)
    os.replace(tmp, HOSTLIB)
    main = os.path.join(HOST, "main.cpp")
    if os.path.exists(main):
        os.makedirs(os.path.dirname(CLI), exist_ok=True)
        _run(["g++"] + flags + ["-o", CLI, main] + lib_srcs +
             ["-L", PKG, "-lalebgk_hip", f"-Wl,-rpath,$ORIGIN/..", "-L/opt/rocm/lib",
              "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    return HOSTLIB


def build(jobs=8):
    build_hip(jobs)
    build_host()
    return LIB


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print("built", LIB)
