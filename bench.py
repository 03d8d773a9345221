"""Benchmark: plain wavefront GPU path tracing on the seeded terrain scene.

Workload (one "step" = one frame): heightfield 512x512 quads + 2000 icospheres of
3 subdivisions (3,084,292 triangles incl. 4 light triangles), per-object Lambert or
GGX microfacet, 2 quad area lights, 1920x1080, 64 spp (8x8 stratified), GiTracer depth 8,
seed 1. The scene is generated, its BVH built and uploaded before the timed region;
each step renders the full frame into a device accumulation buffer.

value = BVH queries (closest + any hit) per second over the timed steps, in Mrays/s.
The roofline object prices the dominant kernel (the wavefront trace kernel) by its
algorithmic bytes: B_ray = 32 (ray) + 16 (hit) + S_node * node visits + 48 * triangle
tests (S_node = 64 B for the binary layout, 128 B for the 4-wide one), with visits
counted by the same kernel in counting mode on a 4-spp frame.

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
MFMA_FP16_PEAK_TFLOPS = 2500.0
MLP_FLOP_PER_ROW = 2 * (96 * 64 + 3 * 64 * 64 + 64 * 48)   # one forward row of the guide network
# one training row: the forward, the backward products of layers 5..2 and every weight gradient
TRAIN_FLOP_PER_ROW = MLP_FLOP_PER_ROW + 2 * (48 * 64 + 3 * 64 * 64) + 2 * (96 * 64 + 3 * 64 * 64 + 64 * 48)
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


# Guided vs plain (quirk-free) path tracing: the occluded room, 500k triangles, 1024x1024, depth 8.
G_RES, G_DEPTH, G_TARGET = 1024, 8, 0.02
G_PASS_GRID, G_CHECK = 2, 1               # 4 spp per pass; RMSE estimated after every pass (outside the clock)
G_ALPHA, G_WARM, G_LR, G_BATCH = 0.5, 4, 5e-3, 1 << 18
G_STEPS, G_STEPS_LATE, G_LATE = 10, 0, 8   # Adam steps per pass: G_STEPS, from pass G_LATE on G_STEPS_LATE
# (passes that train no step write no records)
G_BOUNCES = 0                              # guided path vertices (0 = all)
# the BSDF fraction falls linearly from 1 to G_ALPHA over the first G_WARM passes, while the
# untrained guide would only add variance


def rmse_estimate(a) -> float:
    """RMSE of the per-pixel mean over pixels and channels, from the per-pixel sample
    variance (both tracers are unbiased, so the MSE of a pixel is its variance of the mean)."""
    import numpy as np
    n = a[..., 3:4].astype(np.float64)
    m = a[..., 0:3] / n
    var = np.maximum(a[..., 4:7] / n - m * m, 0.0) / np.maximum(n - 1.0, 1.0)
    return float(np.sqrt(np.mean(var)))


def time_to_rmse(ctx, scene, guide, max_passes, seed0, target=G_TARGET, keep=False) -> dict:
    """Renders passes of G_PASS_GRID^2 spp into one accumulation buffer until the RMSE
    estimate reaches `target` (keep: also return the accumulation buffer as "image"). With a guide, every pass also writes training records and
    runs G_STEPS Adam steps on G_BATCH of them; every pass's samples count in the image.
    Returns the seconds of render + training, spp, rmse, training seconds, guided network rows,
    and the summed launch times (ms) of trace, shade and inference with the BVH queries. The
    RMSE checks are outside the clock."""
    import numpy as np
    from banshee_amd import hip
    acc = ctx.alloc((G_RES, G_RES, 8), np.float32)
    cap = G_RES * G_RES * G_PASS_GRID ** 2 * 4 if guide is not None else 0   # up to 4 records per path
    recs = ctx.alloc((cap, hip.RECORD_FLOATS), np.float32) if cap else None
    ctx.sync()
    clock = train = 0.0
    e = float("inf")
    k = 0
    tot = {"guide_rows": 0, "guide_ms": 0.0, "trace_ms": 0.0, "shade_ms": 0.0, "rays": 0}
    while k < max_passes:
        t = time.perf_counter()
        steps = (G_STEPS if k < G_LATE else G_STEPS_LATE) if guide is not None else 0
        if guide is not None:
            guide.set_bsdf_fraction(max(G_ALPHA, 1.0 - (1.0 - G_ALPHA) * k / G_WARM))
        n = scene.render_guided(G_RES, G_RES, G_PASS_GRID, G_DEPTH, seed0 + k, guide=guide,
                                records=recs if steps > 0 else None, accum=acc, accumulate=k > 0, download=False)
        st = scene.stats()
        for key in ("guide_rows", "guide_ms", "trace_ms", "shade_ms"):
            tot[key] += st[key]
        tot["rays"] += st["rays_closest"] + st["rays_any"]
        if steps > 0 and n:
            t2 = time.perf_counter()
            guide.train(recs, min(n, cap), steps, G_BATCH)
            ctx.sync()
            train += time.perf_counter() - t2
        ctx.sync()
        clock += time.perf_counter() - t
        k += 1
        if k % G_CHECK == 0:
            e = rmse_estimate(acc.numpy())
            if e <= target:
                break
    spp = k * G_PASS_GRID ** 2
    return {"s": clock, "spp": spp, "rmse": e, "train_s": train, "rows": tot["guide_rows"],
            "image": acc.numpy() if keep else None,
            "infer_ms": tot["guide_ms"], "ms_per_spp": clock * 1e3 / spp,
            "trace_ms_per_spp": tot["trace_ms"] / spp, "shade_ms_per_spp": tot["shade_ms"] / spp,
            "rays_per_sample": tot["rays"] / (spp * G_RES * G_RES)}


def make_context(args):
    from banshee_amd import hip
    ctx = hip.Context(0)
    if args.refill_lanes or args.paths_per_chunk or args.trace_blocks:
        ctx.set_tuning(paths_per_chunk=args.paths_per_chunk, trace_blocks_per_cu=args.trace_blocks,
                       refill_lanes=args.refill_lanes)
    return ctx


def training_rate(ctx, rows=1 << 18, steps=30) -> dict:
    """Adam steps of the guide on `rows` synthetic records (uniform positions and roughness,
    isotropic directions, weights in [0.1, 2)), launched back to back and timed by the wall
    clock around them (the launches run ahead of the device, so the clock sees device time)."""
    import numpy as np
    from banshee_amd import hip
    rng = np.random.default_rng(5)
    rec = np.zeros((rows, 12), np.float32)
    rec[:, 0:7] = rng.uniform(0.0, 1.0, (rows, 7))
    d = rng.normal(size=(rows, 3))
    rec[:, 8:11] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, 11] = rng.uniform(0.1, 2.0, rows)
    dr = ctx.to_device(rec)
    guide = hip.Guide(ctx, seed=3)
    for _ in range(3):
        guide.train_step(dr, want_loss=False)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(steps):
        guide.train_step(dr, want_loss=False)
    ctx.sync()
    s = (time.perf_counter() - t) / steps
    tflops = rows * TRAIN_FLOP_PER_ROW / s / 1e12
    return {"bound": "mfma", "kernel": "k_guide_train_fused (one Adam step: forward, loss, backward, weight "
                                       "gradients; FP16 MFMA, FP32 sums)",
            "achieved": round(tflops, 1), "peak": MFMA_FP16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / MFMA_FP16_PEAK_TFLOPS, 4), "rows": rows, "flop_per_row": TRAIN_FLOP_PER_ROW,
            "ms_per_step": round(s * 1e3, 4)}


def qflag(v: int):
    """--bvh-quantized -> hip.Scene(quantized=...): 0 the default (None), 1 True, -1 False."""
    return None if v == 0 else v > 0


def guided_comparison(ctx, bvh_width, quantized=None) -> dict:
    from banshee_amd import hip
    from banshee_amd.scenes import occluded_room
    sc = occluded_room(width=G_RES, height=G_RES)
    scene = hip.Scene(ctx, sc, width=bvh_width, quantized=quantized)
    scene.render_guided(G_RES, G_RES, 1, G_DEPTH, 1, download=False)   # warm the wavefront buffers
    ctx.sync()
    p = time_to_rmse(ctx, scene, None, 4096, 100000)
    guide = hip.Guide(ctx, seed=1, bsdf_fraction=1.0, lr=G_LR, beta2=0.99, guided_bounces=G_BOUNCES)
    g = time_to_rmse(ctx, scene, guide, 4096, 200000)
    tp, sp, ep = p["s"], p["spp"], p["rmse"]
    tg, sg, eg, tt, rows, infer_ms = g["s"], g["spp"], g["rmse"], g["train_s"], g["rows"], g["infer_ms"]
    tflops = rows * MLP_FLOP_PER_ROW / (infer_ms / 1e3) / 1e12 if infer_ms > 0 else 0.0
    keys = ("ms_per_spp", "trace_ms_per_spp", "shade_ms_per_spp", "rays_per_sample")
    return {"workload": f"occluded room ({sc.ntris} tris, the light behind a wall with 4 openings totalling 5% of its "
                        f"area, the camera's side lit only through them), "
                        f"{G_RES}x{G_RES}, depth {G_DEPTH}, quirk-free estimator",
            "target_rmse": G_TARGET, "plain_s": round(tp, 3), "plain_spp": sp, "plain_rmse": round(ep, 5),
            "guided_s": round(tg, 3), "guided_spp": sg, "guided_rmse": round(eg, 5),
            "guided_train_s": round(tt, 3), "time_ratio": round(tg / tp, 3),
            "per_spp": {"plain": {k: round(p[k], 4) for k in keys}, "guided": {k: round(g[k], 4) for k in keys}},
            "mlp_roofline": {"bound": "mfma", "kernel": "k_guide_forward (inference, FP16 MFMA)",
                             "achieved": round(tflops, 1), "peak": MFMA_FP16_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": round(tflops / MFMA_FP16_PEAK_TFLOPS, 4), "rows": int(rows),
                             "flop_per_row": MLP_FLOP_PER_ROW, "inference_ms": round(infer_ms, 2),
                             "training": training_rate(ctx)},
            "guide": {"bsdf_fraction": G_ALPHA, "warmup_passes": G_WARM, "lr": G_LR, "guided_bounces": G_BOUNCES,
                      "adam_steps_per_pass": [G_STEPS, G_STEPS_LATE, G_LATE], "batch": G_BATCH,
                      "spp_per_pass": G_PASS_GRID ** 2}}


def rmse_check(ctx, bvh_width, ref_spp=4096, quantized=None) -> dict:
    """The guided-vs-plain comparison at the headline RMSE target and at half of it, each run's
    final image also scored against a converged plain render of ref_spp samples per pixel:
    RMSE_true = sqrt(mean (m - m_ref)^2 - mean var(m_ref)), the reference's own variance of the
    mean taken out. Not part of the default line (about a minute of GPU time)."""
    import numpy as np
    from banshee_amd import hip
    from banshee_amd.scenes import occluded_room
    sc = occluded_room(width=G_RES, height=G_RES)
    scene = hip.Scene(ctx, sc, width=bvh_width, quantized=quantized)
    ref = ctx.alloc((G_RES, G_RES, 8), np.float32)
    for k in range(ref_spp // 64):
        scene.render_guided(G_RES, G_RES, 8, G_DEPTH, 900000 + k, accum=ref, accumulate=k > 0, download=False)
    ctx.sync()
    r = ref.numpy().astype(np.float64)
    nr = r[..., 3:4]
    mr = r[..., 0:3] / nr
    vr = np.maximum(r[..., 4:7] / nr - mr * mr, 0.0) / np.maximum(nr - 1.0, 1.0)

    def score(res):
        a = res.pop("image").astype(np.float64)
        n = a[..., 3:4]
        m = a[..., 0:3] / n
        res["rmse_true"] = float(np.sqrt(max(np.mean((m - mr) ** 2) - np.mean(vr), 0.0)))
        return {k: (round(v, 5) if isinstance(v, float) else v) for k, v in res.items()
                if k in ("s", "spp", "rmse", "rmse_true", "train_s")}

    out = {"workload": f"occluded room ({sc.ntris} tris), {G_RES}x{G_RES}, depth {G_DEPTH}, quirk-free estimator",
           "reference_spp": ref_spp, "reference_rmse_estimate": round(float(np.sqrt(np.mean(vr))), 6), "targets": {}}
    for i, target in enumerate((G_TARGET, G_TARGET / 2)):
        p = score(time_to_rmse(ctx, scene, None, 4096, 100000 + 10000 * i, target, keep=True))
        guide = hip.Guide(ctx, seed=1, bsdf_fraction=1.0, lr=G_LR, beta2=0.99, guided_bounces=G_BOUNCES)
        g = score(time_to_rmse(ctx, scene, guide, 4096, 200000 + 10000 * i, target, keep=True))
        guide.close()
        out["targets"][str(target)] = {"plain": p, "guided": g, "time_ratio": round(g["s"] / p["s"], 3),
                                       "rmse_true_ratio_at_end": round(g["rmse_true"] / p["rmse_true"], 3)}
    return out


def measured_traffic(layout: str):
    """HBM bytes per trace launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    of this workload (calibrated against a copy kernel), or None."""
    import re
    path = os.path.join(ROOT, "profiles", "k_trace_refill_terrain3M_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(rf"traffic_bytes_per_launch\[width={layout}\]: (\d+)", open(path).read())
    return int(m.group(1)) if m else None


def batch_workload(args) -> dict:
    """The batched workload: 64 seeded terrain scenes (1k-2M triangles, log-uniform),
    128^2-1024^2, 4-64 spp, depth 4-12, rendered in one persistent wavefront captured as one
    hipGraph with autotuned chunk size and trace grid; compared with rendering each job alone."""
    from banshee_amd import hip
    from banshee_amd.scenes import batch_jobs, terrain_for_tris
    ctx = make_context(args)
    t0 = time.time()
    cfgs = batch_jobs(n=64, seed=100)
    scenes, jobs = [], []
    for c in cfgs:
        sc = terrain_for_tris(c["tris"], seed=c["seed"], width=c["res"], height=c["res"])
        s = hip.Scene(ctx, sc, width=args.bvh_width, quantized=qflag(args.bvh_quantized), max_leaf=args.max_leaf)
        scenes.append(s)
        jobs.append(hip.BatchJob(s, "render", width=c["res"], height=c["res"], img_grid=c["grid"],
                                 max_depth=c["depth"], seed=c["seed"]))
    setup_s = time.time() - t0
    tris = sum(s.info()["tris"] for s in scenes)
    t1 = time.time()
    best, best_ms = ctx.autotune(jobs, paths=[1 << 21, 1 << 22, 1 << 23], blocks=[2, 3, 4, 5])
    tune_s = time.time() - t1
    tuning = dict(best, use_graph=1)
    for _ in range(args.warmup):
        ctx.batch_run(jobs, tuning=tuning)
    ctx.sync()
    rays = samples = 0
    t2 = time.perf_counter()
    for _ in range(args.steps):
        ctx.batch_run(jobs, tuning=tuning)
        st = ctx.stats()
        rays += st["rays_closest"] + st["rays_any"]
        samples += st["camera_samples"]
    ctx.sync()
    dt = time.perf_counter() - t2
    # the same jobs one by one (bsh_render each, same tuning, no graph)
    ctx.set_tuning(best["paths_per_chunk"], best["trace_blocks_per_cu"], 0, best["refill_lanes"], best["claim_rays"])
    t3 = time.perf_counter()
    for j in jobs:
        j.scene.render(j.params.width, j.params.height, j.params.img_grid, j.params.max_depth, j.params.seed,
                       accum=j.accum, download=False)
    ctx.sync()
    solo_s = time.perf_counter() - t3
    return {"metric": "batched GI Mrays/s (64 scenes, one hipGraph launch)", "value": round(rays / dt / 1e6, 2),
            "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "64 seeded terrain scenes, tris log-uniform 1k-2M, 128^2-1024^2, 4-64 spp, "
                                   "depth 4-12, one persistent wavefront as one hipGraph",
                       "tris_total": int(tris), "samples_per_step": samples // args.steps,
                       "samples_per_s": round(samples / dt, 1), "tuning": best,
                       "autotune_s": round(tune_s, 2), "setup_s": round(setup_s, 2),
                       "one_by_one_s": round(solo_s, 3), "batched_s": round(dt / args.steps, 3),
                       "batch_speedup": round(solo_s / (dt / args.steps), 3)}}


def warm_bakes(ctx, scene, probes, coeffs, n, min_s=0.5):
    """Untimed bakes before a timed probe run: n of them, and more until min_s seconds have
    passed. A bake takes about 10 ms, so a few of them alone leave the timed bakes to run
    before the GPU's clocks have settled (11.3 ms per bake after one warm-up bake in a fresh
    process, against 9.5 ms after the GI frames)."""
    t0 = time.perf_counter()
    k = 0
    while k < n or time.perf_counter() - t0 < min_s:
        scene.bake_sh(probes, 1024, 2, 6, 1, coeffs=coeffs, download=False)
        ctx.sync()
        k += 1


def probe_workload(args) -> dict:
    """SH probe baking: 4096 probes on a 16^3 grid in a 2M-triangle terrain scene, 1024 rays
    per probe, depth 6, order-3 SH (9 coefficients x RGB), all probes in one wavefront."""
    from banshee_amd import hip
    from banshee_amd.scenes import probe_grid, terrain_for_tris
    ctx = make_context(args)
    sc = terrain_for_tris(2_000_000, seed=7)
    scene = hip.Scene(ctx, sc, width=args.bvh_width, quantized=qflag(args.bvh_quantized), max_leaf=args.max_leaf)
    probes = probe_grid(sc, 16)
    coeffs = ctx.alloc((len(probes), 9, 3), "float32")
    warm_bakes(ctx, scene, probes, coeffs, args.warmup)
    rays = paths = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        scene.bake_sh(probes, 1024, 2, 6, 1 + k, coeffs=coeffs, download=False)
        st = scene.stats()
        rays += st["rays_closest"] + st["rays_any"]
        paths += st["camera_samples"]
    ctx.sync()
    dt = time.perf_counter() - t0
    guided = None if args.no_guided else probe_guided_comparison(ctx, scene, probes)
    return {"metric": "SH probe baking Mrays/s", "value": round(rays / dt / 1e6, 2), "unit": "Mrays/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "4096 probes (16^3 grid) in a 2M-triangle terrain scene, 1024 rays/probe, "
                                   "depth 6, order-3 SH", "tris": int(sc.ntris),
                       "probe_paths_per_s": round(paths / dt, 1), "probes_per_s": round(len(probes) * args.steps / dt, 1)},
            "guided": guided}


P_RAYS, P_DEPTH, P_TRAIN_PASSES, P_STEPS = 1024, 6, 12, 8


def probe_guided_comparison(ctx, scene, probes) -> dict:
    """Plain vs guided proposal for the probe bake, both with the quirk-free estimator; the
    guided bake draws the probe rays themselves from the uniform/guide mixture and guides the
    continuation. The guide trains on the probe paths' own records over P_TRAIN_PASSES bakes
    (uniform fraction falling from 1 to 0.5), then two independent bakes of each kind give the
    variance of a coefficient as mean((c_1 - c_2)^2) / 2 over all probes and coefficients; each
    bake is timed alone. efficiency = (var_plain t_plain) / (var_guided t_guided), training
    excluded (reported on its own). dc_mean_z: the mean over probes of the DC difference (the
    three channels averaged first, since they come from the same paths) over its standard error."""
    import numpy as np
    from banshee_amd import hip
    n = len(probes)
    dp = ctx.to_device(np.ascontiguousarray(probes, np.float32))
    c = [ctx.alloc((n, 9, 3), "float32") for _ in range(2)]

    def bake_pair(guide, seed0):
        times, out = [], []
        for k in range(2):
            ctx.sync()
            t = time.perf_counter()
            scene.bake_sh_guided(dp, P_RAYS, 2, P_DEPTH, seed0 + k, guide=guide, coeffs=c[k], download=False)
            ctx.sync()
            times.append(time.perf_counter() - t)
            out.append(c[k].numpy().astype(np.float64))
        d2 = (out[0] - out[1]) ** 2 / 2
        dc = [o[:, 0, :].mean(1) for o in out]          # channels share their paths: average them first
        return float(np.mean(d2)), float(np.mean(times)), dc[0], float(np.mean((dc[0] - dc[1]) ** 2 / 2))

    bake_pair(None, 50)   # warm
    var_p, t_p, cp, vdc_p = bake_pair(None, 300)
    g = hip.Guide(ctx, seed=1, bsdf_fraction=1.0, lr=G_LR, beta2=0.99)
    cap = n * P_RAYS * 4
    recs = ctx.alloc((cap, hip.RECORD_FLOATS), np.float32)
    ctx.sync()
    t = time.perf_counter()
    for k in range(P_TRAIN_PASSES):
        g.set_bsdf_fraction(max(0.5, 1.0 - 0.5 * k / (P_TRAIN_PASSES // 2)))
        nr = scene.bake_sh_guided(dp, P_RAYS, 2, P_DEPTH, 1000 + k, guide=g, records=recs, download=False)
        if nr:
            g.train(recs, min(nr, cap), P_STEPS, G_BATCH)
    ctx.sync()
    t_train = time.perf_counter() - t
    var_g, t_g, cg, vdc_g = bake_pair(g, 400)
    # bias check: z of the mean over probes of the channel-averaged DC difference
    dc_z = float(np.mean(cg - cp) / np.sqrt((vdc_p + vdc_g) / n))
    return {"estimator": "quirk-free; probe rays from the uniform/guide mixture weighted by 1/(4 pi p_mix), "
                         "guided continuation at surface vertices",
            "rays_per_probe": P_RAYS, "depth": P_DEPTH, "train_passes": P_TRAIN_PASSES,
            "plain_bake_ms": round(t_p * 1e3, 2), "guided_bake_ms": round(t_g * 1e3, 2),
            "train_s": round(t_train, 3), "coeff_var_plain": var_p, "coeff_var_guided": var_g,
            "var_ratio": round(var_g / var_p, 4), "time_ratio": round(t_g / t_p, 4),
            "efficiency": round(var_p * t_p / (var_g * t_g), 3), "dc_mean_z": round(dc_z, 2)}


def measured_ceilings(ctx) -> dict:
    """Streaming-read ceilings of this GPU in GB/s (bsh_diag_stream_read): a 4 GiB buffer read
    once per pass (HBM: 16x the Infinity Cache), a 128 MiB buffer (inside the 256 MB Infinity
    Cache) and a 2 MiB buffer read whole by every XCD (inside each XCD's 4 MB L2)."""
    import numpy as np
    out = {}
    for key, nbytes, reps, per_xcd in (("hbm_4GiB", 4 << 30, 3, False), ("infinity_cache_128MiB", 128 << 20, 20, False),
                                       ("l2_2MiB_per_xcd", 2 << 20, 100, True)):
        buf = ctx.alloc(nbytes // 4, np.float32)
        # the best of cached and non-temporal loads at 8 and 32 workgroups per CU (non-temporal
        # loads on the full grid are the faster form from HBM, cached loads on 8 per CU from L2)
        out[key] = round(max(ctx.stream_read_gbs(buf, reps, per_xcd, nt, small)
                             for nt in (False, True) for small in (False, True)), 1)
        buf.free()
    return out


def ray_classes(ls, s_node, s_tri) -> dict:
    """Per ray class, from the per-launch statistics of a counting frame: camera rays (launch 0),
    extension and MIS rays (closest hit, launches >= 1) and shadow rays (any hit, launches >= 1),
    with node visits and triangle tests per ray and B_ray = 32 (ray) + 16 (hit) + S_node V_node +
    S_tri V_tri."""
    def cls(rays, nodes, tris):
        v_n, v_t = nodes / max(rays, 1), tris / max(rays, 1)
        return {"rays": int(rays), "node_visits_per_ray": round(v_n, 3), "tri_tests_per_ray": round(v_t, 3),
                "bytes_per_ray": round(48 + s_node * v_n + s_tri * v_t, 1)}
    sec = ls[1:]
    return {"camera": cls(ls[0]["rays_closest"], ls[0]["nodes_closest"], ls[0]["tris_closest"]),
            "extension_mis": cls(sum(l["rays_closest"] for l in sec), sum(l["nodes_closest"] for l in sec),
                                 sum(l["tris_closest"] for l in sec)),
            "shadow": cls(sum(l["rays_any"] for l in sec), sum(l["nodes_any"] for l in sec),
                          sum(l["tris_any"] for l in sec))}


def isolated_launches(ctx, scene, classes, seed) -> dict:
    """One untimed frame of the workload on one wavefront lane at the two-lane default's per-lane
    grid (3 trace workgroups per CU), so no trace launch shares the GPU with another lane: per
    launch class, the summed HIP-event durations on the launch stream and the algorithmic bytes
    of the rays it traced (rays x that class's B_ray)."""
    old = ctx.tuning()
    try:
        ctx.set_tuning(0, 3, 0, 0, 0, 1)
        scene.render(W, H, GRID, DEPTH, seed, download=False)
        ctx.sync()
        ls = ctx.launch_stats()
        st = scene.stats()
    finally:
        ctx.set_tuning(**old)
    b = {k: classes[k]["bytes_per_ray"] for k in classes}
    cam_bytes = ls[0]["rays_closest"] * b["camera"]
    sec_bytes = sum(l["rays_closest"] * b["extension_mis"] + l["rays_any"] * b["shadow"] for l in ls[1:])
    cam_ms = ls[0]["trace_ms"]
    sec_ms = sum(l["trace_ms"] for l in ls[1:])
    return {"launches": st["trace_launches"], "trace_ms": cam_ms + sec_ms, "bytes": cam_bytes + sec_bytes,
            "camera": {"ms": cam_ms, "bytes": cam_bytes, "rays": ls[0]["rays_closest"]},
            "secondary": {"ms": sec_ms, "bytes": sec_bytes,
                          "rays": sum(l["rays_closest"] + l["rays_any"] for l in ls[1:])}}


def batch_line(ctx, args) -> dict:
    """The batched workload (64 seeded terrain scenes, 1k-2M triangles, 128^2-1024^2, 4-64 spp,
    depth 4-12) with the library's default launch shape captured as one hipGraph: Mrays/s over
    `steps` runs, against the same jobs rendered one bsh_render call each."""
    from banshee_amd import hip
    from banshee_amd.scenes import batch_jobs, terrain_for_tris
    cfgs = batch_jobs(n=64, seed=100)
    scenes, jobs = [], []
    for c in cfgs:
        s = hip.Scene(ctx, terrain_for_tris(c["tris"], seed=c["seed"], width=c["res"], height=c["res"]),
                      width=args.bvh_width, quantized=qflag(args.bvh_quantized))
        scenes.append(s)
        jobs.append(hip.BatchJob(s, "render", width=c["res"], height=c["res"], img_grid=c["grid"],
                                 max_depth=c["depth"], seed=c["seed"]))
    tuning = {"use_graph": 1}
    ctx.batch_run(jobs, tuning=tuning)
    ctx.sync()
    rays = 0
    t0 = time.perf_counter()
    steps = max(1, min(args.steps, 5))
    for _ in range(steps):
        ctx.batch_run(jobs, tuning=tuning)
        st = ctx.stats()
        rays += st["rays_closest"] + st["rays_any"]
    ctx.sync()
    dt = (time.perf_counter() - t0) / steps
    t1 = time.perf_counter()
    for j, c in zip(jobs, cfgs):
        j.scene.render(c["res"], c["res"], c["grid"], c["depth"], c["seed"], accum=j.accum, download=False)
    ctx.sync()
    solo = time.perf_counter() - t1
    return {"value": round(rays / steps / dt / 1e6, 1), "unit": "Mrays/s", "s_per_batch": round(dt, 4),
            "one_by_one_s": round(solo, 4), "batch_speedup": round(solo / dt, 3), "steps": steps,
            "tris_total": int(sum(s.info()["tris"] for s in scenes)), "tuning": "library default, one hipGraph",
            "workload": "64 seeded terrain scenes, tris log-uniform 1k-2M, 128^2-1024^2, 4-64 spp, depth 4-12"}


def probes_line(ctx, args) -> dict:
    """SH probe baking: 4096 probes (16^3 grid) in a 2M-triangle terrain scene, 1024 rays per
    probe, depth 6, order-3 SH, in one wavefront: Mrays/s and probes/s."""
    from banshee_amd import hip
    from banshee_amd.scenes import probe_grid, terrain_for_tris
    sc = terrain_for_tris(2_000_000, seed=7)
    scene = hip.Scene(ctx, sc, width=args.bvh_width, quantized=qflag(args.bvh_quantized), max_leaf=args.max_leaf)
    probes = probe_grid(sc, 16)
    coeffs = ctx.alloc((len(probes), 9, 3), "float32")
    warm_bakes(ctx, scene, probes, coeffs, 1)
    rays = 0
    steps = max(1, min(args.steps, 10))
    t0 = time.perf_counter()
    for k in range(steps):
        scene.bake_sh(probes, 1024, 2, 6, 2 + k, coeffs=coeffs, download=False)
        st = scene.stats()
        rays += st["rays_closest"] + st["rays_any"]
    ctx.sync()
    dt = (time.perf_counter() - t0) / steps
    return {"value": round(rays / steps / dt / 1e6, 1), "unit": "Mrays/s", "ms_per_bake": round(dt * 1e3, 3),
            "probes_per_s": round(len(probes) / dt, 1), "steps": steps, "tris": int(sc.ntris),
            "workload": "4096 probes (16^3 grid) in a 2M-triangle terrain, 1024 rays/probe, depth 6, order-3 SH"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-guided", action="store_true", help="skip the guided-vs-plain time-to-RMSE comparison")
    ap.add_argument("--no-extra", action="store_true", help="skip the batch and probe-baking figures")
    ap.add_argument("--bvh-width", type=int, default=4, choices=(2, 4), help="BVH node width")
    ap.add_argument("--max-leaf", type=int, default=4, help="triangles per BVH leaf, 1..8")
    ap.add_argument("--bvh-quantized", type=int, default=0, choices=(-1, 0, 1),
                    help="width 4: 0 (default) or 1 64-B nodes with 8-bit child boxes, -1 128-B float "
                         "nodes")
    ap.add_argument("--paths-per-chunk", type=int, default=0,
                    help="paths in flight per wavefront chunk (0 = library default, 2^23)")
    ap.add_argument("--refill-lanes", type=int, default=0,
                    help="idle lanes per wave that trigger a ray refill in the 4-wide trace kernel "
                         "(0 = library default, -1 = whole 64-ray batches per wave)")
    ap.add_argument("--trace-blocks", type=int, default=0,
                    help="persistent trace workgroups per CU, 1..5 (0 = library default, 5)")
    ap.add_argument("--rmse-check", action="store_true",
                    help="run the guided comparison at the RMSE target and at half of it, scoring each final "
                         "image against a converged render, and print its JSON")
    ap.add_argument("--guided-only", action="store_true",
                    help="run only the guided-vs-plain comparison and print its JSON (for profiling)")
    ap.add_argument("--workload", default="gi", choices=("gi", "batch", "probes"),
                    help="gi: the headline line (default); batch: 64 scenes in one launch; probes: SH baking")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the path would shard by pixels, samples or probes with one final reduction)",
              file=sys.stderr)
        sys.exit(2)

    from banshee_amd import hip
    from banshee_amd.scenes import terrain_scene

    if args.rmse_check:
        print(json.dumps(rmse_check(make_context(args), args.bvh_width, quantized=qflag(args.bvh_quantized))), flush=True)
        return
    if args.guided_only:
        print(json.dumps(guided_comparison(make_context(args), args.bvh_width, qflag(args.bvh_quantized))), flush=True)
        return
    if args.workload == "batch":
        print(json.dumps(batch_workload(args)), flush=True)
        return
    if args.workload == "probes":
        print(json.dumps(probe_workload(args)), flush=True)
        return

    t0 = time.time()
    sc = terrain_scene(width=W, height=H)
    ctx = make_context(args)
    scene = hip.Scene(ctx, sc, width=args.bvh_width, quantized=qflag(args.bvh_quantized), max_leaf=args.max_leaf)
    info = scene.info()
    setup_s = time.time() - t0
    accum = ctx.alloc((H, W, 8), "float32")

    # counting frames (untimed): node visits and triangle tests per ray class, with this layout and,
    # for the layout-independent figure, with the binary layout of the same triangles
    s_node = int(info["node_bytes"])
    layout = f"{info['width']}q" if info["quantized"] else str(info["width"])
    scene.render(W, H, 2, DEPTH, SEED, accum=accum, download=False, count=True)
    ctx.sync()
    classes = ray_classes(ctx.launch_stats(), s_node, 48)
    bin_scene = hip.Scene(ctx, sc, width=2)
    bin_scene.render(W, H, 2, DEPTH, SEED, accum=accum, download=False, count=True)
    ctx.sync()
    classes_bin = ray_classes(ctx.launch_stats(), 64, 36)
    del bin_scene
    ceilings = measured_ceilings(ctx)
    iso = isolated_launches(ctx, scene, classes, SEED + 1000)

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
    # the dominant kernel's own launch time: the isolated frame's HIP events (no second lane)
    avg_launch_s = iso["trace_ms"] / 1e3 / iso["launches"]
    bytes_per_launch = iso["bytes"] / iso["launches"]
    achieved = bytes_per_launch / avg_launch_s / 1e9
    # the same bytes anchored to the timed steps' wall clock (includes every other kernel of the frame)
    b_mix = iso["bytes"] / (iso["camera"]["rays"] + iso["secondary"]["rays"])
    anchored = rays / dt * b_mix / 1e9
    hbm, l2 = ceilings["hbm_4GiB"], ceilings["l2_2MiB_per_xcd"]
    cam_gbs = iso["camera"]["bytes"] / (iso["camera"]["ms"] * 1e-3) / 1e9
    sec_gbs = iso["secondary"]["bytes"] / (iso["secondary"]["ms"] * 1e-3) / 1e9
    cb = classes_bin
    b_ref = (iso["camera"]["rays"] * cb["camera"]["bytes_per_ray"] +
             iso["secondary"]["rays"] * (cb["extension_mis"]["bytes_per_ray"] * cb["extension_mis"]["rays"] +
                                         cb["shadow"]["bytes_per_ray"] * cb["shadow"]["rays"]) /
             max(cb["extension_mis"]["rays"] + cb["shadow"]["rays"], 1))
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
                   "bvh_width": int(info["width"]), "bvh_node_bytes": int(info["node_bytes"]),
                   "samples_per_s": round(samples / dt, 1), "rays_per_step": rays // args.steps,
                   "setup_s": round(setup_s, 2), "bvh_build_ms": round(info["build_ms"], 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(layout),
                     "traffic_source": "profiles/k_trace_refill_terrain3M_pmc.md (FETCH_SIZE x2.0 + WRITE_SIZE x1.0)",
                     "kernel": "k_trace_refill" if int(info["width"]) == 4 and args.refill_lanes != -1 else "k_trace_queue",
                     "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                     "launch_timing": "HIP events on the launch stream, in an untimed frame of the workload on one "
                                      "lane at the default per-lane grid (3 workgroups/CU), so no launch overlaps "
                                      "another lane's kernels",
                     "avg_launch_ms_two_lanes": round(trace_ms / trace_launches, 4),
                     "bytes_per_ray": round(b_mix, 1), "anchored_to_steps_gbs": round(anchored, 1),
                     "anchored_frac": round(anchored / HBM_PEAK_GBS, 4),
                     "measured_ceilings_gbs": ceilings,
                     "frac_of_measured": {"hbm": round(achieved / hbm, 4), "l2": round(achieved / l2, 4)},
                     "incoherent_bounces": {"rays": int(iso["secondary"]["rays"]), "achieved": round(sec_gbs, 1),
                                            "frac_of_measured_hbm": round(sec_gbs / hbm, 4),
                                            "frac_of_measured_l2": round(sec_gbs / l2, 4)},
                     "camera_rays": {"rays": int(iso["camera"]["rays"]), "achieved": round(cam_gbs, 1),
                                     "frac_of_measured_hbm": round(cam_gbs / hbm, 4)},
                     "ray_classes": classes,
                     "layout_independent": {"s_node": 64, "s_tri": 36, "ray_classes": classes_bin,
                                            "achieved": round(b_ref / (iso["trace_ms"] * 1e-3) / 1e9, 1),
                                            "frac_of_measured_hbm": round(b_ref / (iso["trace_ms"] * 1e-3) / 1e9 / hbm,
                                                                          4)}},
        "cpu_baseline": None,
    }
    if not args.no_guided:
        out["guided"] = guided_comparison(ctx, args.bvh_width, qflag(args.bvh_quantized))
    if not args.no_extra:
        out["batch"] = batch_line(ctx, args)
        out["probes"] = probes_line(ctx, args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(None)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
