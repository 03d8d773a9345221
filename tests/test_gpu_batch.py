"""Batching: many scenes and probe sets in one wavefront (optionally one hipGraph) give
bit-identical results to running each job alone; the autotuner only changes speed."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import batch_jobs, cornell_box, terrain_for_tris
    cfgs = batch_jobs(n=5, seed=300, min_tris=1_000, max_tris=40_000, min_res=32, max_res=96, min_grid=2,
                      max_grid=4, min_depth=3, max_depth=8)
    scenes, jobs = [], []
    for k, c in enumerate(cfgs):
        sc = terrain_for_tris(c["tris"], seed=c["seed"], width=c["res"], height=c["res"])
        # chunks mix the node layouts: the float and quantised 4-wide ones (traced together by
        # the queue kernel) and the binary one
        layout = ["4q", 4, "4q", 2, 4][k]
        s = hip.Scene(ctx, sc, width=4 if layout == "4q" else layout, quantized=layout == "4q")
        scenes.append(s)
        jobs.append(("render", s, dict(width=c["res"], height=c["res"], img_grid=c["grid"], max_depth=c["depth"],
                                       seed=c["seed"])))
    cb = hip.Scene(ctx, cornell_box(width=48, height=48))
    scenes.append(cb)
    probes = np.array([[1, 1, 1], [0.5, 1.5, 1.5], [1.5, 0.3, 1.2]], np.float32)
    jobs.append(("bake", cb, dict(probes=probes, rays_per_probe=256, lmax=2, max_depth=5, seed=9)))
    jobs.append(("render", cb, dict(width=48, height=48, img_grid=2, max_depth=5, seed=4)))
    return scenes, jobs


def _make(spec):
    from banshee_amd import hip
    kind, s, kw = spec
    return hip.BatchJob(s, "render" if kind == "render" else "bake", **kw)


@pytest.mark.parametrize("use_graph", [0, 1])
def test_batch_equals_solo(ctx, use_graph):
    scenes, specs = _jobs(ctx)
    solo = []
    for spec in specs:
        j = _make(spec)
        ctx.batch_run([j])
        solo.append(j.result())
    # a small chunk size forces chunks that mix jobs and split jobs across chunks; repeated
    # runs with eager work in between (graph mode: the first run captures the graph, the later
    # ones replay it with new job tables and accumulation buffers), then a run with another
    # chunk size, which must capture a new graph
    for rep, chunk in enumerate((20000, 20000, 20000, 30000)):
        batch = [_make(spec) for spec in specs]
        ctx.batch_run(batch, tuning={"paths_per_chunk": chunk, "trace_blocks_per_cu": 3, "use_graph": use_graph})
        replayed = ctx.stats()["graph_replayed"]
        if use_graph and rep in (1, 2):
            assert replayed == 1, rep
        if not use_graph or rep == 3:
            assert replayed == 0, rep
        for k, (a, j) in enumerate(zip(solo, batch)):
            b = j.result()
            for x, y in zip(a, b):
                if x is not None:
                    np.testing.assert_array_equal(x, y, err_msg=f"rep {rep} job {k}")
        eager = _make(specs[-1])
        ctx.batch_run([eager])
        assert ctx.stats()["graph_replayed"] == 0
    st = ctx.stats()
    assert st["rays_closest"] > 0


def test_batch_render_equals_bsh_render(ctx, cornell):
    from banshee_amd import hip
    s = hip.Scene(ctx, cornell)
    S, Q, N, _ = s.render(64, 64, 2, 5, 3)
    j = hip.BatchJob(s, "render", width=64, height=64, img_grid=2, max_depth=5, seed=3)
    ctx.batch_run([j], tuning={"paths_per_chunk": 5000, "trace_blocks_per_cu": 2, "use_graph": 1})
    S2, Q2, N2, _ = j.result()
    np.testing.assert_array_equal(S, S2)
    np.testing.assert_array_equal(Q, Q2)


def test_autotune_picks_a_candidate(ctx):
    scenes, specs = _jobs(ctx)
    jobs = [_make(spec) for spec in specs]
    old = ctx.tuning()
    try:
        best, ms = ctx.autotune(jobs, paths=[1 << 16, 1 << 20], blocks=[2, 5])
        assert best["paths_per_chunk"] in (1 << 16, 1 << 20) and best["trace_blocks_per_cu"] in (2, 5)
        assert ms > 0
        assert ctx.tuning()["paths_per_chunk"] == best["paths_per_chunk"]
    finally:
        ctx.set_tuning(**old)


@pytest.mark.parametrize("bad", [{"claim_rays": 32}, {"claim_rays": 5000}, {"trace_blocks_per_cu": 6},
                                 {"refill_lanes": 65}, {"paths_per_chunk": -1}, {"use_graph": 2}])
def test_batch_run_rejects_out_of_range_tuning(ctx, cornell, bad):
    """A tuning passed to bsh_batch_run is checked like bsh_context_set_tuning's, before any
    kernel is launched: the job's output buffer is left as it was."""
    from banshee_amd import hip
    s = hip.Scene(ctx, cornell)
    j = hip.BatchJob(s, "render", width=16, height=16, img_grid=1, max_depth=2, seed=1)
    j.accum.upload(np.full((16, 16, 8), 7.0, np.float32))
    with pytest.raises(hip.BshError, match=list(bad)[0]):
        ctx.batch_run([j], tuning=bad)
    assert np.all(j.accum.numpy() == 7.0)


@pytest.mark.parametrize("claim", [64, 200, 4096])
def test_claim_size_changes_no_result(ctx, claim):
    """Trace waves claiming claim_rays queue entries per atomic (a reserve kept per wave) give
    the same batch outputs as the automatic claim; out-of-range sizes are rejected."""
    from banshee_amd import hip
    scenes, specs = _jobs(ctx)
    ref = []
    for spec in specs:
        j = _make(spec)
        ctx.batch_run([j], tuning={"paths_per_chunk": 1 << 14, "trace_blocks_per_cu": 2})
        ref.append(j.result())
    jobs = [_make(spec) for spec in specs]
    ctx.batch_run(jobs, tuning={"paths_per_chunk": 1 << 14, "trace_blocks_per_cu": 2, "claim_rays": claim})
    for j, r in zip(jobs, ref):
        for a, b in zip(j.result(), r):
            if a is not None:
                np.testing.assert_array_equal(a, b)
    old = ctx.tuning()
    with pytest.raises(hip.BshError):
        ctx.set_tuning(0, 0, 0, 0, 32)
    assert ctx.tuning() == old


def test_batch_jobs_match_the_reference(ctx, cornell, terrain100k):
    """A mixed batch, captured as one hipGraph with chunks that mix jobs: the golden Cornell and
    terrain crops (per camera sample, against the compiled reference's fixtures), terrain jobs
    of several sizes, and a Cornell probe set (per probe ray and per coefficient, against the
    CPU restatement's bake)."""
    from banshee_amd import hip
    from banshee_amd.scenes import batch_jobs, terrain_for_tris
    from conftest import golden
    import oracle_lib
    from test_gpu_render import _per_sample_agreement
    cs, ts = hip.Scene(ctx, cornell, width=4), hip.Scene(ctx, terrain100k, width=2)
    jobs, fixtures = [], []
    for name, s in (("cornell", cs), ("terrain", ts)):
        fx = golden(f"gi_{name}")
        W, H, grid, depth, seed = (int(fx[k]) for k in ("W", "H", "grid", "depth", "seed"))
        jobs.append(hip.BatchJob(s, "render", width=W, height=H, img_grid=grid, max_depth=depth, seed=seed,
                                 crop=tuple(int(v) for v in fx["crop"]), per_sample=True))
        fixtures.append(fx)
    others = []
    for c in batch_jobs(n=4, seed=500, min_tris=2_000, max_tris=200_000, min_res=32, max_res=128, max_grid=4):
        s = hip.Scene(ctx, terrain_for_tris(c["tris"], seed=c["seed"], width=c["res"], height=c["res"]))
        others.append(s)
        jobs.append(hip.BatchJob(s, "render", width=c["res"], height=c["res"], img_grid=c["grid"],
                                 max_depth=c["depth"], seed=c["seed"]))
    probes = np.array([[1, 1, 1], [0.5, 1.5, 1.5], [1.5, 0.3, 1.2], [0.2, 0.2, 1.9]], np.float32)
    jobs.append(hip.BatchJob(cs, "bake", probes=probes, rays_per_probe=1024, lmax=2, max_depth=5, seed=13,
                             radiance=True))
    ctx.batch_run(jobs, tuning={"paths_per_chunk": 30000, "trace_blocks_per_cu": 3, "use_graph": 1})
    for name, j, fx in zip(("cornell", "terrain"), jobs, fixtures):
        S, Q, N, PS = j.result()
        frac, _ = _per_sample_agreement(PS, fx["crop_samples"])
        print(f"batched {name} crop: {frac:.5f} of the camera samples agree with the reference")
        assert frac >= 0.97
        np.testing.assert_allclose(S, PS.astype(np.float64).sum(axis=2), rtol=1e-5, atol=1e-5)
    c, rad = jobs[-1].result()
    if oracle_lib.available("cpu"):
        cr, rr = oracle_lib.Oracle("cpu").scene(cornell).bake_sh(probes, 1024, 2, 5, 13)
        close = np.all(np.abs(rad - rr) <= 1e-4 + 1e-3 * np.abs(rr), -1)
        print(f"batched probe rays agreeing with the restatement: {close.mean():.5f}")
        assert close.mean() >= 0.97
        from test_gpu_sh import _sh_fp64, probe_dirs
        Y = np.abs(_sh_fp64(probe_dirs(13, len(probes), 1024)))
        dL = np.abs(rad.astype(np.float64) - rr)
        bound = 4 * np.pi / 1024 * (np.einsum("pnk,pnc->pkc", Y, dL) +
                                    1e-5 * np.einsum("pnk,pnc->pkc", Y, np.abs(rr.astype(np.float64)))) + 1e-7
        assert np.all(np.abs(c - cr) <= bound)


def test_full_batch_shape_in_one_graph(ctx):
    """The batched workload at its full shape: 64 seeded terrain scenes, 1k-2M triangles
    (log-uniform), 128^2-1024^2, 4-64 spp, depth 4-12, in one persistent wavefront captured as
    one hipGraph with the library's default launch shape. Every job's sums equal its solo
    bsh_render's bit for bit, and the batch issues exactly the solo runs' rays."""
    from banshee_amd import hip
    from banshee_amd.scenes import batch_jobs, terrain_for_tris
    cfgs = batch_jobs(n=64, seed=100)
    jobs, scenes = [], []
    for c in cfgs:
        s = hip.Scene(ctx, terrain_for_tris(c["tris"], seed=c["seed"], width=c["res"], height=c["res"]))
        scenes.append(s)
        jobs.append(hip.BatchJob(s, "render", width=c["res"], height=c["res"], img_grid=c["grid"],
                                 max_depth=c["depth"], seed=c["seed"]))
    tris = sum(s.info()["tris"] for s in scenes)
    old = ctx.tuning()
    try:
        ctx.set_tuning(0, 0, 0, 0, 0)
        solo, rays = [], [0, 0]
        for j, c in zip(jobs, cfgs):
            S, Q, N, _ = j.scene.render(c["res"], c["res"], c["grid"], c["depth"], c["seed"])
            st = j.scene.stats()
            rays[0] += st["rays_closest"]
            rays[1] += st["rays_any"]
            solo.append((S, Q, N))
        ctx.batch_run(jobs, tuning={"use_graph": 1})
        st = ctx.stats()
    finally:
        ctx.set_tuning(**old)
    print(f"64 jobs, {tris} triangles, {st['camera_samples']} camera samples, "
          f"{st['rays_closest'] + st['rays_any']} rays in one graph: {st['total_ms']:.1f} ms")
    assert 1_000 <= min(c["tris"] for c in cfgs) and max(c["tris"] for c in cfgs) <= 2_000_000
    assert (st["rays_closest"], st["rays_any"]) == tuple(rays)
    for k, (j, (S, Q, N)) in enumerate(zip(jobs, solo)):
        S2, Q2, N2, _ = j.result()
        assert np.array_equal(S, S2) and np.array_equal(Q, Q2) and np.array_equal(N, N2), f"job {k}"
