#!/usr/bin/env python3
"""Builds tests/emu/_build/libcatabrafft_emu.so: the library's own sources (planner, C ABI,
kernels) compiled with g++ against the CPU emulation runtime in this folder.  TEST
INFRASTRUCTURE ONLY: loaded by tests/test_emu.py, never by the product."""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "catabrafft.jl_amd", "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libcatabrafft_emu.so")
FLAGS = ["-std=c++20", "-O1", "-fPIC", "-ffp-contract=off", "-w",
         "-I", os.path.join(HERE, "include"), "-I", CSRC, "-I", os.path.join(ROOT, "include")]


def sources():
    return ([os.path.join(HERE, "emu.cpp")] + sorted(glob.glob(os.path.join(CSRC, "*.cpp"))) +
            sorted(glob.glob(os.path.join(CSRC, "*.hip"))))


def deps():
    return (glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
            + [os.path.join(HERE, "include", "hip", "hip_runtime.h")])


def compile_one(src):
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    newest = max([os.path.getmtime(src)] + [os.path.getmtime(d) for d in deps()])
    if os.path.exists(obj) and os.path.getmtime(obj) >= newest:
        return obj
    tmp = f"{obj}.{os.getpid()}.tmp"
    cmd = ["g++"] + FLAGS + ["-x", "c++", "-c", src, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, obj)
    return obj


def build():
    os.makedirs(OUT, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(compile_one, sources()))
    if os.path.exists(LIB) and os.path.getmtime(LIB) >= max(os.path.getmtime(o) for o in objs):
        return LIB
    # link under a private name and rename: concurrent test workers never see a partial file
    tmp = f"{LIB}.{os.getpid()}.tmp"
    subprocess.run(["g++", "-shared", "-o", tmp] + objs + ["-lpthread"], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
