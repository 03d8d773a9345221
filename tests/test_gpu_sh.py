"""Spherical harmonics on the GPU: basis, env-map projection, irradiance and probe baking."""
import numpy as np
import pytest

import oracle_lib
from conftest import golden

pytestmark = pytest.mark.gpu


def sh_close(got, ref, rtol=1e-5, floor=1e-2):
    """|got - ref| <= rtol * max(|ref|, floor * |c00|) per channel. Near-zero coefficients
    carry only FP32 summation noise (about 1e-7 absolute for sums of ~1000 terms of size
    ~0.01), so their tolerance is floored at 1e-2 of the DC term."""
    c00 = np.abs(ref[..., 0:1, :])
    scale = np.maximum(np.abs(ref), floor * c00) + 1e-12
    return np.abs(got - ref) <= rtol * scale


def env_projection_fp64(env, lmax=2):
    """ShProjectEnvironmentMap (shproject.cpp:73-130) in float64 with the closed-form basis."""
    H, W = env.shape[:2]
    th = np.pi / H / 2 + np.arange(H) * np.pi / H
    ph = 2 * np.pi / W / 2 + np.arange(W) * 2 * np.pi / W
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)
    Y = _sh_fp64(w)
    wt = (np.sin(T) * (np.pi / H) * (2 * np.pi / W))[..., None]
    return np.einsum("hwk,hwc->kc", Y * wt, env.astype(np.float64))


def test_basis_bit_exact(ctx):
    fx = golden("sh")
    np.testing.assert_array_equal(ctx.diag_sh_eval(fx["dirs"], 4), fx["ylm4"])


def test_env_projection(ctx):
    fx = golden("sh")
    got = ctx.sh_project_env(fx["env"], 2)
    # against an FP64 projection of the same map: 1e-5 relative (floored at 1e-2 |c00|)
    assert np.all(sh_close(got[None], env_projection_fp64(fx["env"])[None]))
    # against the reference's own FP32 sequential sum: within that sum's rounding (1e-4)
    assert np.all(sh_close(got[None], fx["proj"][None], rtol=1e-4))
    const = ctx.sh_project_env(np.ones((64, 128, 3), np.float32), 2)
    assert np.all(sh_close(const[None], env_projection_fp64(np.ones((64, 128, 3)))[None]))
    assert np.all(sh_close(const[None], fx["const_proj"][None], rtol=1e-4))
    assert abs(const[0, 0] - np.sqrt(4 * np.pi)) < 2e-3   # Riemann sum of Y00 over the sphere


def test_irradiance(ctx):
    fx = golden("sh")
    got = ctx.sh_irradiance(2, fx["cin"], fx["normals"])
    np.testing.assert_allclose(got, fx["irr"], rtol=2e-6, atol=1e-7)


PROBES = np.array([[1.0, 1.0, 1.0], [0.5, 1.5, 1.5], [1.5, 0.3, 1.2], [0.2, 0.2, 1.9], [1.0, 1.7, 0.4]],
                  np.float32)


def _sh_fp64(d):
    """Closed forms of the reference basis for lmax = 2 (math/sh.cpp:97-153: z pole,
    Condon-Shortley phase, index l^2 + l + m, m < 0 from sin(|m| phi))."""
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    return np.stack([0.28209479177387814 * np.ones_like(x), -0.4886025119029199 * y, 0.4886025119029199 * z,
                     -0.4886025119029199 * x, 1.0925484305920792 * x * y, -1.0925484305920792 * y * z,
                     0.31539156525252005 * (3 * z * z - 1), -1.0925484305920792 * x * z,
                     0.5462742152960396 * (x * x - y * y)], -1)


def probe_dirs(seed, P, N):
    """The stratified probe-ray directions [P, N, 3] regenerated in float64 from the keyed
    stream (stream 4, counters 0 and 1 give the jitter of cell j of probe p, sample p*N + j)."""
    from test_gpu_shading import _keyed_np
    g = int(round(np.sqrt(N)))
    j = np.arange(N)
    cs = (np.arange(P)[:, None] * N + j[None]).astype(np.uint32).ravel()
    st = np.full(cs.shape, 4, np.uint32)
    uy = _keyed_np(seed, cs, st, np.zeros_like(cs)).astype(np.float64) * 2.328306e-10
    ux = _keyed_np(seed, cs, st, np.ones_like(cs)).astype(np.float64) * 2.328306e-10
    u = ((j % g)[None].repeat(P, 0).ravel() + ux) / g
    v = ((j // g)[None].repeat(P, 0).ravel() + uy) / g
    z = 1 - 2 * u
    r = np.sqrt(np.maximum(0, 1 - z * z))
    return np.stack([r * np.cos(2 * np.pi * v), r * np.sin(2 * np.pi * v), z], -1).reshape(P, N, 3)


def test_probe_projection_of_a_radiance_tape(ctx, cornell):
    """Given the GPU's own per-ray radiance, its FP32 projection matches an FP64 projection."""
    from banshee_amd import hip
    s = hip.Scene(ctx, cornell)
    N, seed = 1024, 7
    c, rad = s.bake_sh(PROBES, N, 2, 5, seed, radiance=True)
    Y = _sh_fp64(probe_dirs(seed, PROBES.shape[0], N))
    ref = 4 * np.pi / N * np.einsum("pnk,pnc->pkc", Y, rad.astype(np.float64))
    ok = sh_close(c, ref, rtol=1e-5)
    assert ok.all(), np.abs(c - ref).max()


@pytest.mark.skipif(not oracle_lib.available("cpu"), reason="CPU restatement not built")
def test_probe_bake_matches_restatement(ctx, cornell):
    from banshee_amd import hip
    s = hip.Scene(ctx, cornell)
    N, seed, depth = 1024, 11, 5
    c, rad = s.bake_sh(PROBES, N, 2, depth, seed, radiance=True)
    cpu = oracle_lib.Oracle("cpu").scene(cornell)
    cr, rr = cpu.bake_sh(PROBES, N, 2, depth, seed)
    close = np.all(np.abs(rad - rr) <= 1e-4 + 1e-3 * np.abs(rr), -1)
    print(f"probe rays agreeing with the restatement: {close.mean():.5f}")
    assert close.mean() >= 0.97
    # The coefficients are fixed by the rays: both sides project with the same directions, so
    # |c - cr| <= (4 pi / N) sum_j |L_j - L'_j| |Y_i(w_j)| plus the FP32 rounding of the GPU's
    # sum (1e-5 of sum_j |L_j Y_i(w_j)|, over 16 sequential terms per lane and a 64-lane tree).
    Y = _sh_fp64(probe_dirs(seed, PROBES.shape[0], N))
    dL = np.abs(rad.astype(np.float64) - rr)
    bound = 4 * np.pi / N * (np.einsum("pnk,pnc->pkc", np.abs(Y), dL) +
                             1e-5 * np.einsum("pnk,pnc->pkc", np.abs(Y), np.abs(rr.astype(np.float64)))) + 1e-7
    err = np.abs(c - cr)
    print(f"coefficient error / ray-implied bound: max {np.max(err / bound):.3f}")
    assert np.all(err <= bound)


@pytest.mark.parametrize("quirk_free", [False, True])
def test_all_miss_bake_closed_form(ctx, quirk_free):
    """Probe rays that all miss see the background bg alone: c_00 = sqrt(4 pi) bg and every
    l >= 1 coefficient within 5 sqrt(4 pi / N) |bg| of 0 (test_probe_pins.py derives both),
    with the reference estimator and the quirk-free one; and the coefficients equal the CPU
    restatement's for the same directions."""
    from banshee_amd import hip
    from banshee_amd.scenes import far_speck
    from test_probe_pins import BG, LMAX, N, PROBES as P3, SEED, check_all_miss_coefficients
    sc = far_speck()
    s = hip.Scene(ctx, sc)
    c, rad = s.bake_sh(P3, N, LMAX, 5, SEED, radiance=True, quirk_free=quirk_free, bg=BG)
    assert np.array_equal(rad, np.broadcast_to(BG, rad.shape)), "a probe ray hit the speck"
    check_all_miss_coefficients(c, N, BG, rtol=1e-5)
    if oracle_lib.available("cpu"):
        cr, _ = oracle_lib.Oracle("cpu").scene(sc).bake_sh(P3, N, LMAX, 5, SEED, bg=BG)
        assert np.all(sh_close(c, cr, rtol=1e-5))
