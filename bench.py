"""Benchmark: plain wavefront GPU path tracing on the seeded terrain scene.

Workload (one "step" = one frame): heightfield 512x512 quads + 2000 icospheres of
3 subdivisions (3,084,292 triangles incl. 4 light triangles), per-object Lambert or
GGX microfacet, 2 quad area lights, 1920x1080, 64 spp (8x8 stratified), GiTracer depth 8,
seed 1. The scene is generated, its BVH built and uploaded before the timed region;
each step renders the full frame into a device accumulation buffer.

value = BVH queries (closest + any hit) per second over the timed steps, in Mrays/s.
The roofline object prices the dominant kernel (the wavefront trace kernel) by its
algorithmic bytes: B_ray = 32 (ray) + 16 (hit) + 64 * node visits + 48 * triangle
tests, with visits counted by the same kernel in counting mode on a 4-spp frame.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GI Mrays/s; time-to-target-RMSE guided vs plain PT, 1 MI355X"
HBM_PEAK_GBS = 8000.0
W, H, GRID, DEPTH, SEED = 1920, 1080, 8, 8, 1


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def cpu_baseline(sc_full_args) -> dict | None:
    """The reference's own renderer (oracle/_ref, compiled from its sources) when present,
    else the CPU restatement, on a bounded sample of the same workload: the same scene,
    spp and depth at 240x135 (1/64 of the pixels)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from banshee_amd.scenes import terrain_scene

    kind = "reference" if oracle_lib.available("ref") else ("port" if oracle_lib.available("cpu") else None)
    if kind is None:
        return None
    w, h = 240, 135
    sc = terrain_scene(width=w, height=h)
    o = oracle_lib.Oracle("ref" if kind == "reference" else "cpu")
    s = o.scene(sc)
    cores = os.cpu_count() or 1
    if kind == "reference":
        dt, rays = s.time_render(w, h, GRID, 1, DEPTH, SEED, True)   # MtImageRenderer, hardware_concurrency threads
    else:
        lib = o.lib
        import ctypes
        import numpy as np
        r = np.zeros(2, np.int64)
        dt = lib.cref_time_render(s.h, w, h, GRID, 1, DEPTH, SEED, cores, r.ctypes.data_as(ctypes.c_void_p))
        rays = r
    n = int(rays[0] + rays[1])
    return {"value": round(n / dt / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": kind,
            "sample": f"terrain scene (3.08M tris), {w}x{h} (1/64 of the 1920x1080 pixels), 64 spp, depth 8: "
                      f"{n} rays in {dt:.2f} s on {cores} threads of {cpu_model()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the path would shard by pixels, samples or probes with one final reduction)",
              file=sys.stderr)
        sys.exit(2)

    from banshee_amd import hip
    from banshee_amd.scenes import terrain_scene

    t0 = time.time()
    sc = terrain_scene(width=W, height=H)
    ctx = hip.Context(0)
    scene = hip.Scene(ctx, sc)
    info = scene.info()
    setup_s = time.time() - t0
    accum = ctx.alloc((H, W, 8), "float32")

    # counting frame (untimed): node visits and triangle tests of the same kernel
    scene.render(W, H, 2, DEPTH, SEED, accum=accum, download=False, count=True)
    ctx.sync()
    cst = scene.stats()
    crays = cst["rays_closest"] + cst["rays_any"]
    v_node = cst["node_visits"] / crays
    v_tri = cst["tri_tests"] / crays
    b_ray = 32 + 16 + 64 * v_node + 48 * v_tri

    for _ in range(args.warmup):
        scene.render(W, H, GRID, DEPTH, SEED, accum=accum, download=False)
    ctx.sync()
    rays = 0
    trace_ms = 0.0
    trace_launches = 0
    samples = 0
    t1 = time.perf_counter()
    for k in range(args.steps):
        scene.render(W, H, GRID, DEPTH, SEED + k, accum=accum, download=False)
        st = scene.stats()   # render() returns after its frame's events complete
        rays += st["rays_closest"] + st["rays_any"]
        trace_ms += st["trace_ms"]
        trace_launches += st["trace_launches"]
        samples += st["camera_samples"]
    ctx.sync()
    dt = time.perf_counter() - t1

    value = rays / dt / 1e6
    avg_launch_s = trace_ms / 1e3 / trace_launches
    bytes_per_launch = (rays / trace_launches) * b_ray
    achieved = bytes_per_launch / avg_launch_s / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain wavefront GPU path tracing, seeded terrain scene (heightfield 512x512 + 2000 "
                               "icospheres subdiv 3, GGX/Lambert 50/50, 2 quad area lights), 1920x1080, 64 spp, depth 8",
                   "tris": int(info["tris"]), "bvh_nodes": int(info["nodes"]), "bvh_depth": int(info["depth"]),
                   "samples_per_s": round(samples / dt, 1), "rays_per_step": rays // args.steps,
                   "setup_s": round(setup_s, 2), "bvh_build_ms": round(info["build_ms"], 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_trace_queue", "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                     "bytes_per_ray": round(b_ray, 1), "node_visits_per_ray": round(v_node, 2),
                     "tri_tests_per_ray": round(v_tri, 2), "trace_share": round(trace_ms / (dt * 1e3), 3)},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(None)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
