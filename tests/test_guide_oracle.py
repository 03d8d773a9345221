"""Pins of the FP64 guide restatement (oracle/guide_ref.py) from outside itself (CPU).

* the vMF density integrates to 1 (quadrature on a fine lat-long grid);
* Jakob's sampler reproduces the mixture density (chi-square over direction bins);
* the analytic backward equals central finite differences of the loss, per parameter block;
* Adam on records drawn from a known vMF lobe drives the loss to that lobe's entropy.
"""
import numpy as np
import pytest

from oracle import guide_ref as G


def sphere_grid(n_theta=800, n_phi=1600):
    th = (np.arange(n_theta) + 0.5) * np.pi / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (np.pi / n_theta) * (2 * np.pi / n_phi)).ravel()
    return w, dA


def grid_pdf(o_row, w, chunk=1 << 15):
    """The mixture pdf of one output row at every grid direction, evaluated one slice of the
    grid at a time (the row broadcast, not repeated): the whole grid at once builds
    temporaries of several GB."""
    return np.concatenate([G.mixture_pdf(o_row[None], wc[None])[0] for wc in np.split(w, range(chunk, len(w), chunk))])


def random_outputs(rng, n, lk_range=(-1.5, 5.0)):
    o = np.zeros((n, 48))
    o[:, :8] = rng.normal(0, 1.0, (n, 8))
    o[:, 8:32] = rng.normal(0, 1.0, (n, 24))
    o[:, 32:40] = rng.uniform(*lk_range, (n, 8))
    return o


@pytest.mark.parametrize("kappa", [0.135, 1.0, 7.0, 150.0, 403.0])
def test_vmf_normalises(kappa):
    w, dA = sphere_grid(1600 if kappa > 100 else 800, 3200 if kappa > 100 else 1600)
    mu = np.array([0.3, -0.5, 0.81])
    mu /= np.linalg.norm(mu)
    p = G.vmf_pdf(mu[None], np.array([kappa]), w)
    assert abs(np.sum(p * dA) - 1) < 2e-4


def test_mixture_normalises():
    rng = np.random.default_rng(3)
    w, dA = sphere_grid()
    o = random_outputs(rng, 4, (-1.5, 3.5))
    for r in range(4):
        q = grid_pdf(o[r], w)
        assert abs(np.sum(q * dA) - 1) < 1e-4


def test_sampler_matches_density():
    """Chi-square of 40k Jakob samples in 48 equal-area bins against the integrated pdf."""
    rng = np.random.default_rng(5)
    o = random_outputs(rng, 1, (0.0, 3.0))[0]
    n = 40000
    u = rng.random((n, 3))
    d = np.array([G.sample(o, *u[i]) for i in range(n)])
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-9)

    def bin_of(w):   # 6 bands equal in z times 8 sectors in phi: equal areas
        zb = np.clip(((w[:, 2] + 1) / 2 * 6).astype(int), 0, 5)
        pb = np.clip(((np.arctan2(w[:, 1], w[:, 0]) + np.pi) / (2 * np.pi) * 8).astype(int), 0, 7)
        return zb * 8 + pb

    counts = np.bincount(bin_of(d), minlength=48)
    gw, dA = sphere_grid()
    q = grid_pdf(o, gw)
    expect = np.bincount(bin_of(gw), weights=q * dA, minlength=48) * n
    chi2 = np.sum((counts - expect) ** 2 / expect)
    assert chi2 < 90   # 47 dof: p(chi2 > 90) < 1e-4


def records(rng, n, lobe=None):
    rec = np.zeros((n, 12))
    rec[:, 0:3] = rng.random((n, 3))
    rec[:, 3] = rng.uniform(0.1, 1.0, n)
    nn = rng.normal(size=(n, 3))
    rec[:, 4:7] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    rec[:, 7] = 1.0
    if lobe is None:
        d = rng.normal(size=(n, 3))
        rec[:, 8:11] = d / np.linalg.norm(d, axis=1, keepdims=True)
    else:
        rec[:, 8:11] = lobe
    rec[:, 11] = rng.exponential(1.0, n)
    return rec


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(11)
    p = G.he_init(2).astype(np.float64)
    rec = records(rng, 96)
    L, g = G.loss_and_grad(p, rec)
    assert np.isclose(L, G.loss(p, rec))
    offs, _ = G.offsets()
    for (k, n), (wo, bo) in zip(G.SIZES, offs):
        idx = list(rng.integers(wo, wo + k * n, 6)) + list(rng.integers(bo, bo + n, 3))
        for i in idx:
            h = 1e-6 * max(1.0, abs(p[i]))
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (G.loss(pp, rec) - G.loss(pm, rec)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_output_gradient_matches_finite_differences():
    """Every output channel, including the clamped-log-kappa edge, against FD of -log q."""
    rng = np.random.default_rng(12)
    o = random_outputs(rng, 8)
    o[0, 32] = 6.5   # clamped: zero gradient
    d = rng.normal(size=(8, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = rng.random(8)
    g = G.dloss_dout(o, d, c)

    def f(oo):
        return float(np.sum(c * -np.log(G.mixture_pdf(oo, d))))
    for r in range(8):
        for j in range(40):
            h = 1e-6
            op, om = o.copy(), o.copy()
            op[r, j] += h
            om[r, j] -= h
            fd = (f(op) - f(om)) / (2 * h)
            assert abs(fd - g[r, j]) <= 1e-7 + 1e-5 * abs(fd), (r, j, fd, g[r, j])
    assert g[0, 32] == 0


def adam_train(p, rec_fn, steps, lr=5e-3, b1=0.9, b2=0.99, eps=1e-8, fp16=False):
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    losses = []
    for t in range(1, steps + 1):
        L, g = G.loss_and_grad(p, rec_fn(t), fp16)
        losses.append(L)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, losses


def vmf_entropy(kappa):
    return -np.log(kappa / (2 * np.pi * (1 - np.exp(-2 * kappa)))) - kappa * (1 / np.tanh(kappa) - 1 / kappa - 1)


def draw_vmf(rng, mu, kappa, n):
    """Independent numpy draws (Jakob's inversion written in vector form)."""
    u1, u2 = rng.random(n), rng.random(n)
    w = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * kappa)) / kappa
    s = np.sqrt(np.maximum(0, 1 - w * w))
    t1 = G.orthovector(mu)
    t2 = np.cross(mu, t1)
    phi = 2 * np.pi * u2
    return mu * w[:, None] + np.outer(s * np.cos(phi), t1) + np.outer(s * np.sin(phi), t2)


def test_training_recovers_a_lobe():
    rng = np.random.default_rng(21)
    mu = np.array([0.2, 0.9, -0.3])
    mu /= np.linalg.norm(mu)
    kappa = 20.0

    def batch(t):
        rec = records(rng, 256)
        rec[:, 8:11] = draw_vmf(rng, mu, kappa, 256)
        rec[:, 11] = 1.0
        return rec
    p, losses = adam_train(G.he_init(4).astype(np.float64), batch, 300)
    H = vmf_entropy(kappa)
    assert losses[0] > H + 1.0
    assert np.mean(losses[-20:]) < H + 0.15


def test_emitter_quadrature_matches_lambert_formula():
    """The quadrature behind the direct-light closed-form test, pinned by Lambert's
    polygon formula: for uniform radiance the area-measure integrand is cos_l cos_p / r^2."""
    from analytic import emitter_irradiance, lambert_polygon_irradiance
    tri = np.array([[-0.5, 1.5, -0.5], [0.5, 1.5, -0.5], [0.1, 1.5, 0.6]])
    for p in ([0.3, 0.0, 0.2], [2.0, 0.0, -1.0], [0.0, 0.0, 0.0]):
        q = emitter_irradiance(p, [0, 1, 0], tri, [0, -1, 0], cos_power=1)
        e = lambert_polygon_irradiance(p, [0, 1, 0], tri)
        assert abs(q - e) < 2e-5 * e, (p, q, e)


def test_feature_order():
    """Block b of the encoding is band b: (sin, cos) of 2^b pi v_d at 2d, 2d+1, then the raw
    slots v_2b, v_2b+1; exactly five features are zero padding."""
    v = np.array([[0.1, 0.2, 0.3, 0.4, -0.5, 0.6, 0.7]])
    F = G.features(v)[0]
    assert np.isclose(F[16 * 3 + 2 * 4], np.sin(8 * np.pi * -0.5))
    assert np.isclose(F[16 * 5 + 2 * 6 + 1], np.cos(32 * np.pi * 0.7))
    assert F[16 * 0 + 14] == 0.1 and F[16 * 2 + 15] == 0.6 and F[16 * 3 + 14] == 0.7
    assert G.PAD_FEATURES == [63, 78, 79, 94, 95]
    assert np.all(F[G.PAD_FEATURES] == 0)
