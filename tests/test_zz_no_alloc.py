"""bsh_sh_project_env_noalloc (no reference counterpart: for callers that cannot allocate device
memory) against bsh_sh_project_env on the same maps, through the C ABI. Both run the same two
kernels over the same host tables, so the coefficients agree bit for bit. Two calls in a row, the
second on a smaller map and a lower order, check that nothing the first call left in the shared
scratch array leaks into the second."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_env_projection_matches_the_allocating_entry_point(ctx):
    rng = np.random.default_rng(11)
    for H, W, lmax in ((1024, 2048, 4), (96, 160, 2)):
        env = (rng.random((H, W, 3)) * 4).astype(np.float32)
        print(f"noalloc env projection: {W}x{H}, lmax {lmax}", flush=True)
        ref = ctx.sh_project_env(env, lmax)
        got = ctx.sh_project_env_noalloc(env, lmax)
        assert got.shape == ref.shape == ((lmax + 1) ** 2, 3)
        np.testing.assert_array_equal(got, ref)
    # the same scratch serves a constant map afterwards: c00 is the Riemann sum of Y00
    const = ctx.sh_project_env_noalloc(np.ones((64, 128, 3), np.float32), 2)
    np.testing.assert_array_equal(const, ctx.sh_project_env(np.ones((64, 128, 3), np.float32), 2))
    assert abs(const[0, 0] - np.sqrt(4 * np.pi)) < 2e-3


def test_noalloc_env_projection_rejects_maps_beyond_its_scratch(ctx):
    from banshee_amd import hip
    H, W = 2048, 4096   # at lmax 4 this needs 9.9 MB of scratch, more than the 4 MiB array
    big = ctx.alloc((H, W, 3), np.float32)   # a real buffer of that size, so nothing could read past it
    with pytest.raises(hip.BshError, match="scratch"):
        _call_raw(ctx, big, W, H, 4)
    # at lmax 2 the same map fits (3.6 MB)
    assert _call_raw(ctx, big, W, H, 2).shape == (9, 3)


def _call_raw(ctx, env_dev, W, H, lmax):
    import ctypes
    from banshee_amd import hip
    dc = ctx.alloc(((lmax + 1) ** 2, 3), np.float32)
    hip._check(hip.lib().bsh_sh_project_env_noalloc(ctx.h, env_dev.ptr, ctypes.c_int(W), ctypes.c_int(H),
                                                    ctypes.c_int(lmax), dc.ptr, ctypes.c_void_p(None)),
               "bsh_sh_project_env_noalloc")
    ctx.sync()
    return dc.numpy()
