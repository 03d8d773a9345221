"""Guide network kernels on the device against the FP64 restatement (oracle/guide_ref.py).

Tolerances: the device stores weights, encoded inputs, hidden activations and the
backpropagated gradients in FP16 and accumulates in FP32. Forward outputs are compared with
the FP64 network whose weights and activations are rounded to FP16 at the same points
(relative L2 error <= 5e-3); gradients with the FP64 analytic backward of that network
(relative L2 error <= 1e-2 per parameter block).
"""
import numpy as np
import pytest

from oracle import guide_ref as G

pytestmark = pytest.mark.gpu


def raw_inputs(rng, n):
    x = np.zeros((n, 8), np.float32)
    x[:, 0:3] = rng.random((n, 3))
    x[:, 3] = rng.uniform(0.05, 1.0, n)
    nn = rng.normal(size=(n, 3))
    x[:, 4:7] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    x[:, 7] = 1.0
    return x


def records(rng, n, dirs=None):
    rec = np.zeros((n, 12), np.float32)
    x = raw_inputs(rng, n)
    rec[:, 0:8] = x
    if dirs is None:
        d = rng.normal(size=(n, 3))
        dirs = d / np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, 8:11] = dirs
    rec[:, 11] = rng.exponential(1.0, n)
    return rec


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_mfma_tile_layout(ctx):
    """v_mfma_f32_16x16x16_f16 operand/result layout as the guide kernels assume it:
    small integers make every product and sum exact."""
    rng = np.random.default_rng(0)
    a = rng.integers(-8, 9, (16, 16)).astype(np.float16)
    b = rng.integers(-8, 9, (16, 16)).astype(np.float16)
    d = ctx.diag_mfma(a, b)
    assert np.array_equal(d, a.astype(np.float32) @ b.astype(np.float32))


def test_forward_matches_fp64(ctx):
    from banshee_amd import hip
    g = hip.Guide(ctx, seed=3)
    p = g.get_params()
    assert p.size == G.N_PARAMS
    rng = np.random.default_rng(1)
    for n in (1, 15, 1000, 4099):
        x = raw_inputs(rng, n)
        out = g.forward(x)
        ref16, _, _ = G.forward(p, x[:, :7], fp16=True)
        ref64, _, _ = G.forward(p, x[:, :7], fp16=False)
        assert np.all(np.isfinite(out))
        assert rel(out, ref16) <= 5e-3, (n, rel(out, ref16))
        assert rel(out, ref64) <= 3e-2, (n, rel(out, ref64))


def test_set_params_round_trip(ctx):
    from banshee_amd import hip
    g = hip.Guide(ctx, seed=5)
    p = G.he_init(9)
    g.set_params(p)
    assert np.array_equal(g.get_params(), p)
    x = raw_inputs(np.random.default_rng(2), 512)
    ref16, _, _ = G.forward(p, x[:, :7], fp16=True)
    assert rel(g.forward(x), ref16) <= 5e-3


def test_guide_distribution_matches_numpy(ctx):
    """Device lobe decoding, mixture pdf and Jakob sampling on random outputs."""
    rng = np.random.default_rng(4)
    n = 4096
    o = np.zeros((n, 48), np.float32)
    o[:, :8] = rng.normal(0, 1.5, (n, 8))
    o[:, 8:32] = rng.normal(0, 1.0, (n, 24))
    o[:, 32:40] = rng.uniform(-2.5, 6.5, (n, 8))
    u = rng.random((n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r = ctx.diag_guide(o, u, d)
    assert np.allclose(r[:, 5], 1.0, atol=1e-5)
    q = G.mixture_pdf(o.astype(np.float64), d.astype(np.float64))
    assert np.allclose(r[:, 4], q, rtol=3e-4, atol=1e-7)
    # samples: skip the rows whose lobe choice sits within float rounding of a CDF step
    pi = G.lobes(o.astype(np.float64))[0]
    cdf = np.cumsum(pi, axis=1)
    near = np.min(np.abs(cdf - u[:, :1].astype(np.float64)), axis=1) < 1e-5
    s = np.array([G.sample(o[i], *u[i].astype(np.float64)) for i in range(n)])
    ok = ~near
    assert ok.sum() > n - 20
    err = np.linalg.norm(r[ok, :3] - s[ok], axis=1)
    assert np.max(err) < 2e-3 and np.mean(err) < 1e-5, (np.max(err), np.mean(err))
    qs = G.mixture_pdf(o.astype(np.float64), r[:, :3].astype(np.float64))
    assert np.allclose(r[:, 3], qs, rtol=3e-4, atol=1e-7)


@pytest.mark.parametrize("n", [16, 1000, 5000])
def test_gradient_matches_fp64(ctx, n):
    from banshee_amd import hip
    g = hip.Guide(ctx, seed=7)
    p = g.get_params()
    rec = records(np.random.default_rng(n), n)
    grad, loss = g.gradient(rec)
    L, ref = G.loss_and_grad(p, rec, fp16=True)
    assert abs(loss - L) <= 1e-3 * abs(L) + 1e-4, (loss, L)
    offs, _ = G.offsets()
    for (k, m), (wo, bo) in zip(G.SIZES, offs):
        assert rel(grad[wo:wo + k * m], ref[wo:wo + k * m]) <= 1e-2, ("W", wo, rel(grad[wo:wo + k * m], ref[wo:wo + k * m]))
        assert rel(grad[bo:bo + m], ref[bo:bo + m]) <= 1e-2, ("b", bo, rel(grad[bo:bo + m], ref[bo:bo + m]))
    for j in G.PAD_FEATURES:                      # padding features carry no gradient
        assert np.all(grad[j * 64:(j + 1) * 64] == 0)


def test_fp32_mode_forward_matches_fp64(ctx):
    """FP32 mode (bsh_guide_forward_fp32): the network from the FP32 master parameters
    without FP16 rounding agrees with the FP64 restatement to relative L2 <= 1e-5, and the
    FP16 kernels' outputs lie within their own bound of it."""
    from banshee_amd import hip
    g = hip.Guide(ctx, seed=3)
    p = g.get_params()
    rng = np.random.default_rng(11)
    for n in (1, 15, 1000, 4099):
        x = raw_inputs(rng, n)
        out32 = g.forward(x, fp32=True)
        ref64, _, _ = G.forward(p, x[:, :7], fp16=False)
        assert rel(out32, ref64) <= 1e-5, (n, rel(out32, ref64))
        assert rel(g.forward(x), out32) <= 3e-2


def fp32_gradient_check(g, rec, tol):
    p = g.get_params()
    grad, loss = g.gradient(rec, fp32=True)
    L, ref = G.loss_and_grad(p, rec, fp16=False)
    assert abs(loss - L) <= tol * abs(L) + 1e-7, (loss, L)
    offs, _ = G.offsets()
    worst = 0.0
    for (k, m), (wo, bo) in zip(G.SIZES, offs):
        for a, b in ((wo, wo + k * m), (bo, bo + m)):
            e = rel(grad[a:b], ref[a:b])
            worst = max(worst, e)
            assert e <= tol, (a, e)
    for j in G.PAD_FEATURES:
        assert np.all(grad[j * 64:(j + 1) * 64] == 0)
    return worst


@pytest.mark.parametrize("n", [16, 1000, 5000])
def test_fp32_mode_gradient_matches_fp64(ctx, n):
    """FP32 mode (bsh_guide_gradient_fp32): loss and every parameter block's gradient agree
    with the FP64 analytic backward (itself pinned by finite differences) to relative
    L2 <= 1e-5."""
    from banshee_amd import hip
    g = hip.Guide(ctx, seed=7)
    worst = fp32_gradient_check(g, records(np.random.default_rng(n), n), 1e-5)
    print(f"n={n}: worst block relative error {worst:.2e}")


def test_fp32_mode_gradient_of_a_trained_guide(ctx):
    """The same check after training on records from one sharp lobe, where the network's
    concentrations are large (kappa up to e^6) and its mixture weights uneven."""
    from banshee_amd import hip
    g = hip.Guide(ctx, seed=12, lr=5e-3, beta2=0.99)
    rng = np.random.default_rng(5)
    mu = np.array([0.3, -0.5, 0.81])
    mu /= np.linalg.norm(mu)
    d = rng.normal(size=(8192, 3)) * 0.1 + mu
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rec = records(rng, 8192, d)
    dr = ctx.to_device(rec)
    for _ in range(150):
        g.train_step(dr, want_loss=False)
    out = g.forward(rec[:, :8], fp32=True)
    lk = float(np.max(out[:, 32:40]))
    assert lk > 1.0, lk                            # concentrated lobes
    worst = fp32_gradient_check(g, records(np.random.default_rng(6), 2000, d[:2000]), 1e-5)
    print(f"trained (max log kappa {lk:.2f}): worst block relative error {worst:.2e}")


def test_adam_step_and_fp16_copies(ctx):
    """One Adam step moves every parameter with a clear gradient by lr * sign(g) (the first
    Adam step), and the next forward uses the refreshed FP16 copies."""
    from banshee_amd import hip
    lr = 1e-3
    g = hip.Guide(ctx, seed=8, lr=lr, beta1=0.9, beta2=0.99, eps=1e-12)
    p0 = g.get_params().astype(np.float64)
    rec = records(np.random.default_rng(3), 2048)
    _, ref = G.loss_and_grad(p0, rec, fp16=True)
    g.train_step(rec)
    p1 = g.get_params().astype(np.float64)
    big = np.abs(ref) > 0.05 * np.max(np.abs(ref))
    assert big.sum() > 100
    assert np.allclose((p1 - p0)[big], -lr * np.sign(ref[big]), rtol=1e-3, atol=1e-7)
    x = raw_inputs(np.random.default_rng(6), 777)
    ref16, _, _ = G.forward(p1, x[:, :7], fp16=True)
    assert rel(g.forward(x), ref16) <= 5e-3


def vmf_entropy(kappa):
    return -np.log(kappa / (2 * np.pi * (1 - np.exp(-2 * kappa)))) - kappa * (1 / np.tanh(kappa) - 1 / kappa - 1)


def test_training_on_device_recovers_a_lobe(ctx):
    """Adam on the device, records drawn from one vMF lobe (kappa 20): the weighted
    cross-entropy falls from log(4 pi) to within 0.15 nats of the lobe's entropy."""
    from banshee_amd import hip
    rng = np.random.default_rng(21)
    mu = np.array([0.2, 0.9, -0.3])
    mu /= np.linalg.norm(mu)
    kappa = 20.0
    g = hip.Guide(ctx, seed=4, lr=5e-3, beta1=0.9, beta2=0.99)
    t1 = G.orthovector(mu)
    t2 = np.cross(mu, t1)
    losses = []
    for _ in range(300):
        u1, u2 = rng.random(4096), rng.random(4096)
        w = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * kappa)) / kappa
        s = np.sqrt(np.maximum(0, 1 - w * w))
        d = mu * w[:, None] + np.outer(s * np.cos(2 * np.pi * u2), t1) + np.outer(s * np.sin(2 * np.pi * u2), t2)
        rec = records(rng, 4096, d)
        rec[:, 11] = 1.0
        losses.append(g.train_step(rec))
    H = vmf_entropy(kappa)
    assert losses[0] > H + 1.0
    assert np.mean(losses[-20:]) < H + 0.15, (np.mean(losses[-20:]), H)
    # the trained mixture puts its mass on the lobe
    x = raw_inputs(rng, 64)
    q = G.mixture_pdf(g.forward(x).astype(np.float64), np.repeat(mu[None], 64, 0))
    assert np.all(q > 0.5 * kappa / (2 * np.pi))


def test_multi_step_train_equals_single_steps(ctx):
    """bsh_guide_train(steps, batch) is the same Adam steps as train_step on the strided batches
    it documents, and reproduces bit for bit."""
    from banshee_amd import hip
    rec = records(np.random.default_rng(12), 3000)
    d = ctx.to_device(rec)
    g1 = hip.Guide(ctx, seed=6)
    g2 = hip.Guide(ctx, seed=6)
    g3 = hip.Guide(ctx, seed=6)
    g1.train(d, 3000, 3, 1024)
    g3.train(d, 3000, 3, 1024)
    stride = 3000 // 1024
    for s in range(3):
        off = s % stride
        g2.train_step(rec[off::stride][:1024])
    assert np.array_equal(g1.get_params(), g2.get_params())
    assert np.array_equal(g1.get_params(), g3.get_params())


def test_checkpoint_resume_is_bit_exact(ctx):
    """Training interrupted after a 5-step bsh_guide_train call and resumed on a new guide from
    bsh_guide_get_state ends, after a 7-step call, exactly where the same two calls on one guide
    end (each call picks its batches by its own step index): parameters, both moments, step count
    and the FP16 network's outputs. Resuming from bsh_guide_set_params instead (Adam's moments
    reset) does not, which is what the state adds."""
    from banshee_amd import hip
    rec = records(np.random.default_rng(21), 4096)
    d = ctx.to_device(rec)
    full = hip.Guide(ctx, seed=9)
    full.train(d, 4096, 5, 1024)
    full.train(d, 4096, 7, 1024)
    first = hip.Guide(ctx, seed=9)
    first.train(d, 4096, 5, 1024)
    st = first.get_state()
    assert st["step"] == 5
    resumed = hip.Guide(ctx, seed=3)          # other initial weights: the state overrides them
    resumed.set_state(st)
    resumed.train(d, 4096, 7, 1024)
    a, b = full.get_state(), resumed.get_state()
    assert a["step"] == b["step"] == 12
    for k in ("params", "m1", "m2"):
        assert np.array_equal(a[k], b[k]), k
    x = raw_inputs(np.random.default_rng(2), 512)
    assert np.array_equal(full.forward(x), resumed.forward(x))
    naive = hip.Guide(ctx, seed=9)
    naive.set_params(st["params"])
    naive.train(d, 4096, 7, 1024)
    assert not np.array_equal(naive.get_params(), a["params"])
