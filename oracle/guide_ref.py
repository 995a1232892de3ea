"""ORACLE (test infrastructure only): CPU float64 restatement of the learned
path guide, the checker for the GPU kernels of hastyrenderer_amd/csrc/guide*.

The guide has no counterpart in the reference renderer.  Network: input
78 = 12-frequency sin/cos encoding of the position normalised to the scene
box (72) + shading normal (3) + outgoing direction (3), padded to 80; four
hidden layers of 64 with ReLU; 41 outputs = 8 von Mises-Fisher lobes x
(mean direction 3, log concentration 1, logit 1) and a selection logit z,
padded to 48.  Weights are stored transposed (W^T[out][in], padded), biases
separately.

  mu_k = r_k / |r_k|,  kappa_k = exp(clamp(c_k, C_LO, C_HI)),  w = softmax(a)
  vMF(w; mu, kappa) = kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1))
  q(w) = sum_k w_k vMF_k(w)
  loss(sample) = -F log q(w_i)        (F = the sample's weight)
  with a learned selection in [lo, hi], the BSDF density pB > 0 of w_i and
  the density pS > 0 w_i was drawn with:
    c = lo + (hi - lo) sigmoid(z),  loss += F^2 pS / (c q + (1 - c) pB)
    (the second moment of the one-sample estimate; q held fixed in this
    term, and its gradient trains the output row of z only, not the hidden
    layers)

`bf16=True` rounds weights and every layer input to bfloat16 (round to
nearest even) exactly where the GPU's MFMA operands are rounded, so the GPU
agrees with it to FP32-accumulation rounding.  This model is pinned by
central finite differences of its own loss and by closed forms of the vMF
(normalisation, sampling), see tests/test_guide_cpu.py.
"""
from __future__ import annotations

import math

import numpy as np

N_FREQ = 12
N_IN = 78
IN_PAD = 80
HIDDEN = 64
N_LOBES = 8
N_OUT = 40       # mixture outputs
SEL_OUT = 40     # the selection logit
N_OUT_USED = 41
OUT_PAD = 48
DIMS_IN = (IN_PAD, HIDDEN, HIDDEN, HIDDEN, HIDDEN)
DIMS_OUT = (HIDDEN, HIDDEN, HIDDEN, HIDDEN, OUT_PAD)
C_LO, C_HI = -4.0, 9.0  # log-concentration clamp: kappa in [0.018, 8103]


def param_layout():
    """Offsets of W^T[l] (out x in) and b[l] in the flat parameter vector."""
    offs = []
    o = 0
    for din, dout in zip(DIMS_IN, DIMS_OUT):
        w = o
        o += din * dout
        b = o
        o += dout
        offs.append((w, b, din, dout))
    return offs, o


LAYOUT, N_PARAMS = param_layout()


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    r = np.where(np.isnan(a), u, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed=0):
    """He-uniform hidden layers, a small output layer, zero biases; padding
    entries zero (and they stay zero under training)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    for li, (w, b, din, dout) in enumerate(LAYOUT):
        n_in = N_IN if li == 0 else HIDDEN
        n_out = N_OUT_USED if li == len(LAYOUT) - 1 else HIDDEN
        lim = math.sqrt(6.0 / n_in) if li < len(LAYOUT) - 1 else 0.1 * math.sqrt(3.0 / n_in)
        W = np.zeros((dout, din), np.float32)
        W[:n_out, :n_in] = rng.uniform(-lim, lim, (n_out, n_in))
        p[w:w + din * dout] = W.ravel()
    return p


def encode(x, n, wo):
    """Input features (B, 80): per axis a and frequency f, sin(2^f pi x_a),
    cos(2^f pi x_a) at 24a + 2f, 24a + 2f + 1; then n, wo, two zeros.  x is
    the position normalised to [0, 1]^3, in float32 like the GPU; 2^f x_a is
    exact in float32 and pi (2^f x_a) is formed in float64, as the GPU's
    sinpi / cospi of 2^f x_a do."""
    x = np.asarray(x, np.float32)
    B = x.shape[0]
    out = np.zeros((B, IN_PAD), np.float64)
    for a in range(3):
        for f in range(N_FREQ):
            t = (np.float32(2.0 ** f) * x[:, a]).astype(np.float64)
            out[:, 24 * a + 2 * f] = np.sin(math.pi * t)
            out[:, 24 * a + 2 * f + 1] = np.cos(math.pi * t)
    out[:, 72:75] = n
    out[:, 75:78] = wo
    return out


def forward(params, feats, bf16_ops=False):
    """Returns raw outputs (B, 48) and the per-layer (input, pre-activation)."""
    h = feats.astype(np.float64)
    cache = []
    for li, (w, b, din, dout) in enumerate(LAYOUT):
        Wt = params[w:w + din * dout].reshape(dout, din).astype(np.float64)
        bias = params[b:b + dout].astype(np.float64)
        hin = bf16(h) if bf16_ops else h
        Wq = bf16(Wt) if bf16_ops else Wt
        z = hin @ Wq.T + bias
        cache.append((hin, z))
        h = np.maximum(z, 0.0) if li < len(LAYOUT) - 1 else z
    return h, cache


def mixture(raw):
    """raw (B, >=40) -> mu (B,8,3), kappa (B,8), w (B,8), |r| (B,8), c (B,8)."""
    r = raw[:, :N_OUT].reshape(-1, N_LOBES, 5)
    vec = r[:, :, 0:3]
    nr = np.sqrt(np.sum(vec * vec, axis=2))
    mu = vec / np.maximum(nr, 1e-12)[:, :, None]
    c = r[:, :, 3]
    kappa = np.exp(np.clip(c, C_LO, C_HI))
    a = r[:, :, 4]
    a = a - a.max(axis=1, keepdims=True)
    w = np.exp(a)
    w /= w.sum(axis=1, keepdims=True)
    return mu, kappa, w, nr, c


def vmf_pdf(mu, kappa, d):
    """vMF density of directions d (B,3) for lobes mu (B,K,3), kappa (B,K)."""
    cos = np.einsum("bkj,bj->bk", mu, d)
    norm = kappa / (2.0 * math.pi * -np.expm1(-2.0 * kappa))
    return norm * np.exp(kappa * (cos - 1.0))


def pdf(raw, d):
    mu, kappa, w, _, _ = mixture(raw)
    return np.sum(w * vmf_pdf(mu, kappa, d), axis=1)


def selection(raw, lo, hi):
    """Probability of drawing from the guide: lo + (hi - lo) sigmoid(z)."""
    return lo + (hi - lo) / (1.0 + np.exp(-raw[:, SEL_OUT]))


def loss_and_grad_raw(raw, d, F, pB=None, sel=None, pS=None):
    """Mean loss and its gradient with respect to the raw outputs (B, 48)
    (already divided by the batch size): -F log q, plus the selection term
    for samples with pB > 0 and pS > 0 when sel = (lo, hi) with lo < hi."""
    B = raw.shape[0]
    mu, kappa, w, nr, c = mixture(raw)
    v = vmf_pdf(mu, kappa, d)
    q = np.maximum(np.sum(w * v, axis=1), 1e-30)  # all lobes pointing away: keep log finite
    per = -F * np.log(q)
    g_sel = np.zeros(B)
    if sel is not None and pB is not None and pS is not None and sel[1] > sel[0]:
        lo, hi = sel
        use = (pB > 0.0) & (pS > 0.0)
        sg = 1.0 / (1.0 + np.exp(-raw[:, SEL_OUT]))
        cs = lo + (hi - lo) * sg
        pm = np.maximum(cs * q + (1.0 - cs) * pB, 1e-30)
        m2 = F * F * pS / pm
        per = per + np.where(use, m2, 0.0)
        g_sel = np.where(use, -m2 * (q - pB) / pm * (hi - lo) * sg * (1.0 - sg), 0.0)  # d m2 / dz
    loss = np.mean(per)
    s = (w * v) / q[:, None]                    # posterior of each lobe
    g_a = s - w                                  # d log q / d a
    e = np.exp(-2.0 * kappa)
    dlogv_dk = 1.0 / kappa - 2.0 * e / (-np.expm1(-2.0 * kappa)) + (np.einsum("bkj,bj->bk", mu, d) - 1.0)
    inside = (c > C_LO) & (c < C_HI)
    g_c = s * dlogv_dk * kappa * inside
    proj = d[:, None, :] - np.einsum("bkj,bj->bk", mu, d)[:, :, None] * mu
    g_r = (s * kappa / np.maximum(nr, 1e-12))[:, :, None] * proj
    g = np.zeros_like(raw)
    gl = g[:, :N_OUT].reshape(B, N_LOBES, 5)
    gl[:, :, 0:3] = g_r
    gl[:, :, 3] = g_c
    gl[:, :, 4] = g_a
    g[:, :N_OUT] = gl.reshape(B, N_OUT)
    out = -(F[:, None] * g) / B
    out[:, SEL_OUT] = g_sel / B
    return loss, out


def backward(params, cache, g_raw, bf16_ops=False):
    """Parameter gradient (flat, like params) from the raw-output gradient."""
    grad = np.zeros(N_PARAMS, np.float64)
    g = g_raw
    for li in range(len(LAYOUT) - 1, -1, -1):
        w, b, din, dout = LAYOUT[li]
        hin, z = cache[li]
        if li < len(LAYOUT) - 1:
            g = g * (z > 0.0)
        gq = bf16(g) if bf16_ops else g
        grad[w:w + din * dout] = (gq.T @ hin).ravel()
        grad[b:b + dout] = gq.sum(axis=0)
        Wt = params[w:w + din * dout].reshape(dout, din).astype(np.float64)
        Wq = bf16(Wt) if bf16_ops else Wt
        if li == len(LAYOUT) - 1:
            gq = gq.copy()
            gq[:, SEL_OUT] = 0.0  # the selection output does not train the hidden layers
        g = gq @ Wq
    return grad


def adam_step(params, m, v, grad, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1), float32 state like the GPU."""
    g = grad.astype(np.float32)
    m = (np.float32(b1) * m + np.float32(1 - b1) * g).astype(np.float32)
    v = (np.float32(b2) * v + np.float32(1 - b2) * g * g).astype(np.float32)
    mh = m / np.float32(1 - b1 ** t)
    vh = v / np.float32(1 - b2 ** t)
    p = (params - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))).astype(np.float32)
    return p, m, v


def sample_vmf(mu, kappa, xi1, xi2):
    """Direction from one lobe (Jakob's stable inversion): w = 1 +
    log(xi1 + (1 - xi1) e^{-2 kappa}) / kappa about mu.  mu (..., 3);
    kappa, xi1, xi2 (...)."""
    wz = 1.0 + np.log(xi1 + (1.0 - xi1) * np.exp(-2.0 * kappa)) / kappa
    wz = np.clip(wz, -1.0, 1.0)
    s = np.sqrt(np.maximum(0.0, 1.0 - wz * wz))
    phi = 2.0 * math.pi * xi2
    local = np.stack([s * np.cos(phi), s * np.sin(phi), wz], -1)
    # frame around mu
    t = np.where(np.abs(mu[..., 0:1]) > 0.9, np.array([0.0, 1.0, 0.0]), np.array([1.0, 0.0, 0.0]))
    b1 = np.cross(mu, t)
    b1 /= np.linalg.norm(b1, axis=-1, keepdims=True)
    b2 = np.cross(mu, b1)
    return local[..., 0:1] * b1 + local[..., 1:2] * b2 + local[..., 2:3] * mu
