"""TEST INFRASTRUCTURE ONLY -- FP64 numpy restatement of the path-guiding network.

Only tests/ may import this module, and only as the checker of the device kernels in
banshee_amd/csrc/guide.hip and guide_dev.h. The learned guide has no counterpart in the
reference, so there is no reference file to follow. The restatement below follows the
network's definition, and the tests pin it from outside itself:

* vmf_pdf integrates to 1 over the sphere (quadrature), and the Jakob sampler's
  histogram matches the density (chi-square).
* loss_and_grad's analytic gradient matches central finite differences of loss() in
  FP64, for every parameter block.
* Training with it on records drawn from a known vMF lobe recovers that lobe.

Network (parameters in one flat FP32 vector, weights input-major W[k][n], then bias):
  features(7 raw): 6 blocks of 16, block b = band b; position p < 14 is sin (p even) or cos
  (p odd) of 2^b pi raw_{p/2}, positions 14, 15 the raw values raw_{2b}, raw_{2b+1} (zero
  past raw_6): 91 features and 5 zeros
  -> 4 x (64, ReLU) -> 48 outputs: logits[0:8], mu raw[8:32], log kappa[32:40] in [-2, 6].
Loss over records (position, roughness, normal, direction, weight w):
  L = sum_r (c_r / n) * (-log q(dir_r)),  q = sum_k pi_k vMF(dir; mu_k, kappa_k),
  c_r = min(n w_r / sum w, 256)  (weights normalised to mean 1, then capped).
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT, LOBES, RAW, BANDS = 96, 64, 48, 8, 7, 6
MAX_RECORD_WEIGHT = 256.0
LK_LO, LK_HI = -2.0, 6.0
SIZES = [(N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT)]


def offsets():
    """(weight offset, bias offset) per layer in the flat parameter vector."""
    out, o = [], 0
    for k, n in SIZES:
        out.append((o, o + k * n))
        o += k * n + n
    return out, o


N_PARAMS = offsets()[1]


def unpack(p):
    offs, _ = offsets()
    Ws, bs = [], []
    for (k, n), (wo, bo) in zip(SIZES, offs):
        Ws.append(np.asarray(p[wo:wo + k * n], np.float64).reshape(k, n))
        bs.append(np.asarray(p[bo:bo + n], np.float64))
    return Ws, bs


def pack(Ws, bs):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in zip(Ws, bs)])


PAD_FEATURES = [16 * b + p for b in range(BANDS) for p in (14, 15) if 2 * b + p - 14 >= RAW]


def features(raw):
    """raw: n x 7 -> n x 96 in the block-per-band order."""
    raw = np.asarray(raw, np.float64)
    n = raw.shape[0]
    F = np.zeros((n, N_IN))
    for band in range(BANDS):
        for d in range(RAW):
            a = (2.0 ** band) * np.pi * raw[:, d]
            F[:, 16 * band + 2 * d] = np.sin(a)
            F[:, 16 * band + 2 * d + 1] = np.cos(a)
        for e in (2 * band, 2 * band + 1):
            if e < RAW:
                F[:, 16 * band + 14 + (e - 2 * band)] = raw[:, e]
    return F


def forward(p, raw, fp16=False):
    """Network outputs (n x 48) and the cached activations. fp16=True rounds weights,
    encoded inputs and hidden activations to binary16 where the device kernels store them
    (the accumulation stays exact here; the device accumulates in FP32)."""
    Ws, bs = unpack(p)
    r16 = (lambda x: x.astype(np.float16).astype(np.float64)) if fp16 else (lambda x: x)
    X = r16(features(raw))
    acts, pre = [X], []
    h = X
    for li in range(4):
        z = h @ r16(Ws[li]) + bs[li]
        pre.append(z)
        h = r16(np.maximum(z, 0.0))
        acts.append(h)
    out = h @ r16(Ws[4]) + bs[4]
    return out, acts, pre


def lobes(o):
    """o: n x 48 -> (pi n x 8, mu n x 8 x 3, kappa n x 8, |r| n x 8, clamped log kappa mask)."""
    o = np.asarray(o, np.float64)
    lg = o[:, :LOBES]
    e = np.exp(lg - lg.max(axis=1, keepdims=True))
    pi = e / e.sum(axis=1, keepdims=True)
    r = o[:, 8:32].reshape(-1, LOBES, 3)
    rn = np.linalg.norm(r, axis=2)
    safe = np.where(rn > 0, rn, 1.0)
    mu = np.where(rn[..., None] > 0, r / safe[..., None], np.array([0.0, 0.0, 1.0]))
    lk = o[:, 32:40]
    kappa = np.exp(np.clip(lk, LK_LO, LK_HI))
    inside = (lk > LK_LO) & (lk < LK_HI)
    return pi, mu, kappa, rn, inside


def vmf_pdf(mu, kappa, w):
    """vMF density on S^2 in the stable form kappa / (2 pi (1 - e^{-2k})) exp(k (mu.w - 1))."""
    c = np.sum(mu * w, axis=-1)
    return kappa / (2 * np.pi * (1 - np.exp(-2 * kappa))) * np.exp(kappa * (c - 1))


def mixture_pdf(o, w):
    """o: n x 48, w: n x 3 (or n x m x 3) -> q."""
    pi, mu, kappa, _, _ = lobes(o)
    w = np.asarray(w, np.float64)
    if w.ndim == 2:
        return np.sum(pi * vmf_pdf(mu, kappa, w[:, None, :]), axis=1)
    return np.sum(pi[:, None, :] * vmf_pdf(mu[:, None], kappa[:, None], w[:, :, None, :]), axis=2)


def orthovector(n):
    """A unit vector orthogonal to n (same branch rule as the device helper)."""
    n = np.asarray(n, np.float64)
    if abs(n[2]) > np.sqrt(0.5):
        v = np.array([0.0, -n[2], n[1]])
    else:
        v = np.array([-n[1], n[0], 0.0])
    return v / np.linalg.norm(v)


def sample(o_row, u0, u1, u2):
    """One draw from the mixture of one output row: lobe by the weight CDF, then
    W. Jakob (2012), 'Numerically stable sampling of the von Mises-Fisher distribution on S^2'."""
    pi, mu, kappa, _, _ = lobes(np.asarray(o_row, np.float64)[None])
    pi, mu, kappa = pi[0], mu[0], kappa[0]
    k = LOBES - 1
    c = 0.0
    for j in range(LOBES):
        c += pi[j]
        if u0 < c:
            k = j
            break
    kap = kappa[k]
    w = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * kap)) / kap
    w = min(max(w, -1.0), 1.0)
    s = np.sqrt(max(0.0, 1 - w * w))
    t1 = orthovector(mu[k])
    t2 = np.cross(mu[k], t1)
    phi = 2 * np.pi * u2
    d = mu[k] * w + t1 * (s * np.cos(phi)) + t2 * (s * np.sin(phi))
    return d / np.linalg.norm(d)


def split_records(rec):
    rec = np.asarray(rec, np.float64).reshape(-1, 12)
    raw = np.concatenate([rec[:, 0:4], rec[:, 4:7]], axis=1)   # x, y, z, roughness, normal
    return raw, rec[:, 8:11], rec[:, 11]


def record_coeffs(wt):
    n = len(wt)
    return np.minimum(wt * n / wt.sum(), MAX_RECORD_WEIGHT) / n


def loss(p, rec, fp16=False):
    raw, d, wt = split_records(rec)
    o, _, _ = forward(p, raw, fp16)
    q = mixture_pdf(o, d)
    c = record_coeffs(wt)
    ok = (q > 0) & (c > 0)
    return float(np.sum(c[ok] * -np.log(q[ok])))


def dloss_dout(o, d, c):
    """Gradient of sum_r c_r (-log q_r) with respect to the 48 outputs (analytic)."""
    pi, mu, kappa, rn, inside = lobes(o)
    v = vmf_pdf(mu, kappa, d[:, None, :])
    q = np.sum(pi * v, axis=1)
    ok = (q > 0) & (c > 0)
    qs = np.where(ok, q, 1.0)
    gam = pi * v / qs[:, None]
    g = np.zeros_like(o, dtype=np.float64)
    cc = np.where(ok, c, 0.0)[:, None]
    g[:, :LOBES] = cc * (pi - gam)
    gm = -(cc * gam * kappa)[..., None] * d[:, None, :]                  # d/d mu_k
    proj = gm - mu * np.sum(mu * gm, axis=2, keepdims=True)
    gr = np.where(rn[..., None] > 0, proj / np.where(rn > 0, rn, 1.0)[..., None], 0.0)
    g[:, 8:32] = gr.reshape(-1, 24)
    e = np.exp(-2 * kappa)
    dlogv = 1 / kappa - 2 * e / (1 - e) + (np.sum(mu * d[:, None, :], axis=2) - 1)
    g[:, 32:40] = np.where(inside, -cc * gam * kappa * dlogv, 0.0)
    return g


def loss_and_grad(p, rec, fp16=False):
    """(loss, gradient in the flat parameter layout), by backpropagation in FP64."""
    raw, d, wt = split_records(rec)
    o, acts, pre = forward(p, raw, fp16)
    c = record_coeffs(wt)
    q = mixture_pdf(o, d)
    ok = (q > 0) & (c > 0)
    L = float(np.sum(c[ok] * -np.log(q[ok])))
    Ws, bs = unpack(p)
    r16 = (lambda x: x.astype(np.float16).astype(np.float64)) if fp16 else (lambda x: x)
    gW, gb = [None] * 5, [None] * 5
    dz = dloss_dout(o, d, c)
    for li in range(4, -1, -1):
        gW[li] = acts[li].T @ dz
        gb[li] = dz.sum(axis=0)
        if li > 0:
            dh = dz @ r16(Ws[li]).T
            dz = dh * (pre[li - 1] > 0)
    return L, pack(gW, gb)


def he_init(seed=0):
    """A random parameter vector of the device initialiser's shape and scale (not its stream)."""
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    for li, (k, n) in enumerate(SIZES):
        lim = np.sqrt(6.0 / (91 if li == 0 else k))
        W = rng.uniform(-lim, lim, (k, n))
        if li == 0:
            W[PAD_FEATURES] = 0
        Ws.append(W)
        bs.append(rng.uniform(-0.1, 0.1, n))
    return pack(Ws, bs).astype(np.float32)
