#!/usr/bin/env python3
"""Builds lib/libcatabrafft.so (the C-ABI HIP library) for gfx950 with hipcc.

    python build.py [--force] [--jobs N]

Sources are compiled in parallel into _build/ and linked into lib/libcatabrafft.so.  The radix
codelets (csrc/codelets.h) are regenerated first by gen/gen_codelets.py.  Without --force an
object is rebuilt only when a source or header is newer than it.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcatabrafft.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops keeps the compiler from emitting packed FP32 VALU instructions.
# -ffp-contract=off: no multiply-add fusion, so every kernel that computes a transform performs
# the same IEEE operations in the same order and gives the same bits (a transform alone or in a
# work list, pipelined or not), and so does the CPU emulation build of the same sources.
CFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC, "-I", os.path.join(ROOT, "include"),
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def headers() -> list[str]:
    return glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(ROOT, "include", "*.h"))


def _compile(src: str, force: bool) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    newest = max([os.path.getmtime(src)] + [os.path.getmtime(h) for h in headers()])
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= newest:
        return obj
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + CFLAGS + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{err}")
    if err.strip():
        sys.stderr.write(err + "\n")
    return obj


def build(force: bool = False, jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    subprocess.run([sys.executable, os.path.join(HERE, "gen", "gen_codelets.py"),
                    os.path.join(CSRC, "codelets.h")], check=True, capture_output=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the C one-transform-per-launch caller the serving benchmark compares against
    loop = os.path.join(LIBDIR, "libcfft_one_per_launch.so")
    cmd = ["gcc", "-O2", "-shared", "-fPIC", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "bench", "one_per_launch.c"), "-L", LIBDIR, "-lcatabrafft",
           "-Wl,-rpath,$ORIGIN", "-o", loop + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(loop + ".tmp", loop)
    # standalone measurement programs (no library): the access patterns of the four-step passes,
    # and a radix-16 stage on the matrix cores against the codelet
    for name in ("pattern_copy", "mfma_dft16"):
        probe = os.path.join(LIBDIR, name)
        cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", "-I", CSRC, "-x", "hip",
               os.path.join(HERE, "bench", name + ".cpp"), "-o", probe + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stderr}")
        os.replace(probe + ".tmp", probe)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    print(build(a.force, a.jobs))
