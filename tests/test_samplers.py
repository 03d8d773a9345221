"""The reference's samplers (sampler/*) on the keyed stream, pinned by what their mathematics
fixes, on the fixture the compiled reference wrote (tests/golden/samplers.npz) and, where the
reference oracle is built, on live calls into it:

* RegularSampler(g): the cell centres ((i // g + 1/2) / g, (i mod g + 1/2) / g), x outer
  (regular_sampler.h:46-58); RegularSampler(1) always (1/2, 1/2).
* StratifiedSampler(g): sample s lies in cell (s mod g, s // g) (stratified_sampler.cpp:9-23).
* CmjSampler(g) (A. Kensler, Pixar memo 13-01): every pixel's g^2 samples are multi-jittered --
  one per coarse cell, and one per column and per row of the fine g^2 x g^2 grid.
* SobolSampler(n): a pixel's first 2^k samples of the XOR-scrambled (0,2)-sequence form a
  (0,k,2)-net: exactly one point in every elementary interval of area 2^-k.
* RandomSampler(1) and StratifiedSampler(1) draw the same two numbers in the same order, so
  their keyed samples coincide (both build float2(NextFloat(), NextFloat())).

A plausible slip (a swapped coordinate, a lost scramble, a wrong permutation bound) breaks
one of these."""
import numpy as np
import pytest

import oracle_lib
from conftest import golden

FX = golden("samplers")


def _img(kind, arg):
    return FX[f"img_{kind}_{arg}"].astype(np.float64)


def test_regular_is_the_cell_centres():
    g = 3
    a = _img("regular", g)
    i = np.arange(g * g)
    want = np.stack([(i // g + 0.5) / g, (i % g + 0.5) / g], -1)
    np.testing.assert_allclose(a, np.broadcast_to(want, a.shape), atol=1e-7)
    np.testing.assert_array_equal(FX["nee_regular"], 0.5)


def test_stratified_samples_stay_in_their_cells():
    g = 4
    a = _img("stratified", g)
    s = np.arange(g * g)
    assert np.all(np.floor(a[..., 0] * g) == s % g) and np.all(np.floor(a[..., 1] * g) == s // g)


@pytest.mark.parametrize("g", [3, 4])
def test_cmj_is_multi_jittered(g):
    a = _img("cmj", g)
    n = g * g
    for px in a:
        assert sorted(np.floor(px[:, 0] * g).astype(int) * g + np.floor(px[:, 1] * g).astype(int)) == list(range(n))
        assert sorted(np.floor(px[:, 0] * n).astype(int)) == list(range(n))
        assert sorted(np.floor(px[:, 1] * n).astype(int)) == list(range(n))
    # the pattern differs from pixel to pixel (a new pattern draw per pixel)
    assert len({tuple(np.round(px[0], 6)) for px in a}) == len(a)


@pytest.mark.parametrize("arg,spp", [(5, 8), (16, 16)])
def test_sobol_pixels_are_02_nets(arg, spp):
    a = _img("sobol", arg)
    assert a.shape[1] == spp   # num_samples = upper_power_of_two(n)
    k = int(np.log2(spp))
    for px in a:
        for bx in range(k + 1):
            by = k - bx
            cells = np.floor(px[:, 0] * 2 ** bx).astype(int) * 2 ** by + np.floor(px[:, 1] * 2 ** by).astype(int)
            assert sorted(cells) == list(range(spp)), (bx, by)
    # scrambles are drawn per pixel (Reset after every pixel)
    assert len({tuple(px[0]) for px in a}) == len(a)


def test_random_one_equals_stratified_one():
    np.testing.assert_array_equal(FX["nee_random"], FX["nee_stratified"])
    r = _img("random", 5)
    assert np.all((r >= 0) & (r < 1))
    c = FX["nee_cmj"]
    assert np.all((c >= 0) & (c < 1)) and len(np.unique(c[:, 0])) == len(c)


@pytest.mark.skipif(not oracle_lib.available("ref"), reason="reference oracle not built")
def test_fixture_matches_live_reference():
    o = oracle_lib.Oracle("ref")
    seed, npix = int(FX["seed"]), int(FX["pixels"])
    for key in FX.files:
        if key.startswith("img_"):
            kind, arg = key[4:].rsplit("_", 1)
            np.testing.assert_array_equal(o.image_samples(kind, int(arg), seed, npix), FX[key])
        elif key.startswith("nee_"):
            np.testing.assert_array_equal(o.nee_samples(key[4:], seed, 64), FX[key])
