"""Closed-form pins of the probe bake's direction map and normalisation (CPU restatement).

Probe rays that all miss return the background bg, so the bake's coefficients are
c_i = (4 pi / N) sum_j bg Y_i(w_j), whatever the estimator (reference or quirk-free):

* c_00 = (4 pi / N) N bg Y_00 = sqrt(4 pi) bg exactly (Y_00 = 1 / (2 sqrt(pi)), sh.cpp:97-153),
  which pins the 4 pi / N factor;
* every l >= 1 coefficient is (4 pi) bg times the sample mean of Y_lm over the N directions.
  For uniform directions on the sphere E[Y_lm] = 0 and Var[Y_lm] = 1 / (4 pi), so the mean's
  standard error is at most sqrt(4 pi / N) |bg| (stratification only lowers it). A direction
  map that is not uniform fails: z = 1 - u gives E[z] = 1/2 (c_10 = 3.07 bg), a map uniform in
  theta gives E[3 z^2 - 1] = 1/2 (c_20 = 1.98 bg).
"""
import numpy as np
import pytest

import oracle_lib
from test_gpu_sh import _sh_fp64, probe_dirs

BG = np.array([0.3, 0.7, 1.1], np.float32)
PROBES = np.array([[0, 0, 0], [1.5, -2, 0.5], [-3, 1, 2]], np.float32)
N, LMAX, SEED = 4096, 2, 5


def check_all_miss_coefficients(c, N, bg, k=5.0, rtol=1e-6):
    """c: [P, 9, 3]. c_00 = sqrt(4 pi) bg to rtol (1e-6 for the FP64 sums of the restatement, 1e-5
    for FP32 sums); |c_lm| <= k sqrt(4 pi / N) |bg| for l >= 1."""
    c = np.asarray(c, np.float64)
    dc = np.sqrt(4 * np.pi) * bg.astype(np.float64)
    np.testing.assert_allclose(c[:, 0, :], np.broadcast_to(dc, c[:, 0, :].shape), rtol=rtol)
    bound = k * np.sqrt(4 * np.pi / N) * np.abs(bg)
    assert np.all(np.abs(c[:, 1:, :]) <= bound), np.abs(c[:, 1:, :]).max(axis=(0, 1)) / bound


@pytest.mark.skipif(not oracle_lib.available("cpu"), reason="CPU restatement not built")
def test_all_miss_bake_fixes_dc_and_uniformity():
    from banshee_amd.scenes import far_speck
    cpu = oracle_lib.Oracle("cpu").scene(far_speck())
    c, rad = cpu.bake_sh(PROBES, N, LMAX, 5, SEED, bg=BG)
    assert np.array_equal(rad, np.broadcast_to(BG, rad.shape)), "a probe ray hit the speck"
    check_all_miss_coefficients(c, N, BG)


def test_the_uniformity_check_rejects_wrong_maps():
    """The statistic above, evaluated on the bake's own stratified cells with a wrong z map, fails:
    the check has the power to see a non-uniform direction map."""
    d = probe_dirs(SEED, 1, N)[0]
    good = 4 * np.pi / N * _sh_fp64(d).sum(0)[:, None] * BG[None].astype(np.float64)
    check_all_miss_coefficients(good[None], N, BG)
    g = int(np.sqrt(N))
    j = np.arange(N)
    u = ((j % g) + 0.5) / g
    v = ((j // g) + 0.5) / g
    for z in (1 - u, np.cos(np.pi * u)):      # half-range z; uniform in theta
        r = np.sqrt(np.maximum(0, 1 - z * z))
        w = np.stack([r * np.cos(2 * np.pi * v), r * np.sin(2 * np.pi * v), z], -1)
        bad = 4 * np.pi / N * _sh_fp64(w).sum(0)[:, None] * BG[None].astype(np.float64)
        with pytest.raises(AssertionError):
            check_all_miss_coefficients(bad[None], N, BG)
    # a wrong normalisation (2 pi / N) fails the DC pin
    with pytest.raises(AssertionError):
        check_all_miss_coefficients(good[None] / 2, N, BG)
