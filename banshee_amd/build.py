"""Builds libbanshee_hip.so (HIP kernels for gfx950 + host BVH builder + C ABI).

    python -m banshee_amd.build            # compile into banshee_amd/libbanshee_hip.so

Every HIP translation unit is compiled with -ffp-contract=off (the ray/triangle test
and the shading code must round exactly as the reference's x86-64 build does; fused
multiply-adds are written explicitly where wanted) and without packed-FP32
instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbanshee_hip.so")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["api.hip", "trace.hip", "render.hip", "diag.hip", "sh.hip", "guide.hip"]
CXX_SOURCES = ["bvh_build.cpp", "obj_import.cpp"]

ARCH = os.environ.get("BANSHEE_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def hip_flags():
    return [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-Wno-unused-result"]


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    hc = hipcc()
    jobs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        jobs.append([hc, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", o] + hip_flags())
    for src in CXX_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        jobs.append(["g++", "-c", os.path.join(CSRC, src), "-o", o, "-O2", "-fPIC", "-std=c++17",
                     "-ffp-contract=off", "-pthread"])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [os.path.join(OBJ, s + ".o") for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB + ".tmp"
    _run([hc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
