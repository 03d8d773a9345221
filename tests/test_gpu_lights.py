"""The light table: the quirk-free estimator picks one light per vertex in proportion to its power
(area x luminance) once a scene has more lights than the bounce-0 loop covers, and weighs it by 1 /
its probability; the reference estimator keeps GiTracer's loop over every light at bounce 0
(tracer/gitracer.cpp:72-84) and refuses scenes beyond its 32 lights."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_power_picked_lights_match_the_closed_form(ctx):
    """40 emissive triangles of different radiance over a Lambert floor, depth 1, quirk-free: the
    image mean equals (albedo / pi) sum_j e_j integral_j V cos_l^2 cos_p / r^2 dA, the emitters'
    radiance being GetLe = e cos_l (material/emissive.h:84-95) and V the visibility past the other
    emitters (the lights hang at different heights, and lower ones shade higher ones by a few
    percent: without V the sum is 4-9 % high)."""
    from analytic import camera_footprint, emitter_irradiance
    from banshee_amd import hip
    from banshee_amd.scenes import lit_floor_many
    sc = lit_floor_many(40)
    S = hip.Scene(ctx, sc)
    s, q, n, _ = S.render_guided(64, 64, 16, 1, 13)
    s, q = s.astype(np.float64), q.astype(np.float64)
    nn = n.astype(np.float64)[..., None]
    mean = s / nn
    var = np.maximum(q / nn - mean * mean, 0) * nn / (nn - 1)
    m = mean.mean(axis=(0, 1))
    se = np.sqrt((var / nn).sum(axis=(0, 1))) / (64 * 64)
    pts = camera_footprint(sc.camera)
    E = np.zeros(3)
    lights = sc.meta["lights"]
    for j, (tri, e) in enumerate(zip(lights, sc.meta["radiance"])):
        others = np.delete(lights, j, axis=0)
        E += e.astype(np.float64) * np.mean([emitter_irradiance(p, [0, 1, 0], tri, [0, -1, 0], cos_power=2, m=80,
                                                                occluders=others) for p in pts])
    expect = sc.meta["albedo"] / np.pi * E
    z = (m - expect) / se
    print(f"40 lights, one picked by power per sample: mean {m}, closed form {expect}, z {z}")
    assert np.all(np.abs(z) < 4.5)
    assert np.all(np.abs(m / expect - 1) < 0.01)


def test_reference_estimator_covers_at_most_32_lights(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import lit_floor_many
    S = hip.Scene(ctx, lit_floor_many(40))
    with pytest.raises(hip.BshError, match="at most 32"):
        S.render(16, 16, 1, 1, 1)
    S32 = hip.Scene(ctx, lit_floor_many(32))
    _, _, n, _ = S32.render(16, 16, 1, 2, 1)
    assert np.all(n == 1)
