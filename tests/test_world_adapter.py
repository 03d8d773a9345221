"""The reference-side adapters (include/banshee_world.h): HipImageRenderer and HipAccelerator over
the C ABI, driven by the reference's own World (tests/integration/world_adapter_demo.cpp).

CPU: the header and the driver compile against the reference's headers (where the reference's
sources are present). GPU: the driver, built by oracle/Makefile, renders a World built with the
reference's API on the CPU (ImageRenderer + GiTracer) and the GPU (HipImageRenderer), queries
World::Intersect over HipAccelerator against the reference's Bvh ray for ray, and renders with the
reference's CPU tracer over HipAccelerator."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("BANSHEE_REF", "/root/reference/FireRays")
DEMO = os.path.join(ROOT, "oracle", "_ref", "world_adapter_demo")


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "Banshee")), reason="reference sources absent")
def test_adapters_compile_against_reference_headers():
    cmd = ["g++", "-std=c++11", "-O0", "-DNDEBUG", "-include", "memory", "-include", "cassert", "-w",
           "-I", os.path.join(REF, "Banshee"), "-I", os.path.join(ROOT, "include"), "-fsyntax-only",
           os.path.join(ROOT, "tests", "integration", "world_adapter_demo.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DEMO), reason="adapter demo not built (needs the reference's sources)")
def test_reference_world_drives_the_gpu_path():
    print("running", DEMO, flush=True)
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=180)
    line = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else ""
    print(line, r.stderr[-500:])
    res = json.loads(line)
    assert res["within_3se"] >= 0.99
    assert res["closest_same"] == res["rays"] and res["occluded_same"] == res["rays"]
    assert res["batch_same"] == res["rays"]
    assert res["cpu_tracer_gpu_accel_same"] >= 0.99
    assert r.returncode == 0 and res["ok"]
