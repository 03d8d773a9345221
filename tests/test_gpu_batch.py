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
        s = hip.Scene(ctx, sc, width=4 if k % 2 else 2)   # one chunk mixes both node layouts
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
    # runs (graph mode: one captured graph per run) with eager work in between
    for rep in range(3):
        batch = [_make(spec) for spec in specs]
        ctx.batch_run(batch, tuning={"paths_per_chunk": 20000, "trace_blocks_per_cu": 3, "use_graph": use_graph})
        for k, (a, j) in enumerate(zip(solo, batch)):
            b = j.result()
            for x, y in zip(a, b):
                if x is not None:
                    np.testing.assert_array_equal(x, y, err_msg=f"rep {rep} job {k}")
        eager = _make(specs[-1])
        ctx.batch_run([eager])
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
    best, ms = ctx.autotune(jobs, paths=[1 << 16, 1 << 20], blocks=[2, 5])
    assert best["paths_per_chunk"] in (1 << 16, 1 << 20) and best["trace_blocks_per_cu"] in (2, 5)
    assert ms > 0
    assert ctx.tuning()["paths_per_chunk"] == best["paths_per_chunk"]
    ctx.set_tuning(1 << 21, 5, 0)


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
