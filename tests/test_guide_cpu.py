"""Pins the CPU model of the learned guide (oracle/guide_ref.py) without a GPU:
central finite differences of its loss against its analytic backward pass,
closed forms of the von Mises-Fisher lobe (normalisation, mean direction
A(kappa) = coth(kappa) - 1/kappa), and Adam's first step."""
import math

import numpy as np
import pytest

from oracle import guide_ref as gr


def _batch(rng, B):
    x = rng.uniform(0, 1, (B, 3))
    n = rng.normal(size=(B, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    wo = rng.normal(size=(B, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wi = rng.normal(size=(B, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    F = rng.uniform(0.1, 2.0, B)
    return x, n, wo, wi, F


def _loss(params, feats, wi, F):
    raw, _ = gr.forward(params, feats)
    loss, _ = gr.loss_and_grad_raw(raw, wi, F)
    return loss


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = gr.init_params(1).astype(np.float64)
    # make the output layer non-trivial so every output gradient matters
    w4, b4, din, dout = gr.LAYOUT[4]
    params[w4:w4 + din * dout] *= 10.0
    x, n, wo, wi, F = _batch(rng, 6)
    feats = gr.encode(x, n, wo)
    raw, cache = gr.forward(params, feats)
    _, g_raw = gr.loss_and_grad_raw(raw, wi, F)
    grad = gr.backward(params, cache, g_raw)
    idx = []
    for (w, b, din, dout), nin, nout in zip(gr.LAYOUT, (78, 64, 64, 64, 64), (64, 64, 64, 64, 40)):
        for _ in range(6):
            idx.append(w + rng.integers(nout) * din + rng.integers(nin))
        idx.append(b + rng.integers(nout))
    h = 1e-6
    for i in idx:
        p = params.copy()
        p[i] += h
        lp = _loss(p, feats, wi, F)
        p[i] -= 2 * h
        lm = _loss(p, feats, wi, F)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_raw_gradient_matches_finite_differences():
    """Covers the vMF/softmax/normalisation derivatives directly (all 40 raw
    outputs, lobes with small and large concentration)."""
    rng = np.random.default_rng(5)
    raw = rng.normal(size=(4, 48))
    raw[:, 3:40:5] = rng.uniform(-3, 8, (4, 8))
    _, n, wo, wi, F = _batch(rng, 4)
    _, g = gr.loss_and_grad_raw(raw, wi, F)
    h = 1e-6
    for b in range(4):
        for i in range(40):
            r = raw.copy()
            r[b, i] += h
            lp, _ = gr.loss_and_grad_raw(r, wi, F)
            r[b, i] -= 2 * h
            lm, _ = gr.loss_and_grad_raw(r, wi, F)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[b, i]) <= 1e-5 * max(1.0, abs(fd)), (b, i)


def test_selection_gradient():
    """The learned selection c = lo + (hi - lo) sigmoid(z) and its loss, the
    second moment F^2 pS / (c q + (1 - c) pB) of the one-sample estimate:
    the gradient reaches z only (q held fixed: the mixture's outputs get the
    same gradient as without the term), matches central differences in z,
    vanishes where q = pB or pB = 0, and points towards more guide samples
    where the guide's density exceeds the BSDF's."""
    rng = np.random.default_rng(9)
    B = 6
    raw = rng.normal(size=(B, 48))
    raw[:, 3:40:5] = rng.uniform(-1, 4, (B, 8))
    _, _, _, wi, F = _batch(rng, B)
    q = gr.pdf(raw, wi)
    pB = q * np.array([0.2, 5.0, 1.0, 0.0, 0.7, 3.0])  # sample 2: pB = q; sample 3: no BSDF density
    pS = 0.5 * q + 0.5 * pB + 0.1
    sel = (0.1, 0.9)
    _, g0 = gr.loss_and_grad_raw(raw, wi, F)
    _, g = gr.loss_and_grad_raw(raw, wi, F, pB=pB, sel=sel, pS=pS)
    np.testing.assert_array_equal(g[:, :gr.N_OUT], g0[:, :gr.N_OUT])
    assert not g0[:, gr.SEL_OUT].any()
    h = 1e-6
    for b in range(B):
        r = raw.copy()
        r[b, gr.SEL_OUT] += h
        lp, _ = gr.loss_and_grad_raw(r, wi, F, pB=pB, sel=sel, pS=pS)
        r[b, gr.SEL_OUT] -= 2 * h
        lm, _ = gr.loss_and_grad_raw(r, wi, F, pB=pB, sel=sel, pS=pS)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[b, gr.SEL_OUT]) <= 1e-7 * max(1.0, abs(fd)), (b, fd, g[b, gr.SEL_OUT])
    assert abs(g[2, gr.SEL_OUT]) < 1e-12 and g[3, gr.SEL_OUT] == 0.0
    assert g[0, gr.SEL_OUT] < 0.0 and g[4, gr.SEL_OUT] < 0.0  # q > pB: raising z lowers the loss
    assert g[1, gr.SEL_OUT] > 0.0 and g[5, gr.SEL_OUT] > 0.0  # q < pB
    # the selection stays inside [lo, hi] whatever the logit
    c = gr.selection(np.array([[0.0] * 40 + [-80.0] + [0.0] * 7, [0.0] * 40 + [80.0] + [0.0] * 7]), *sel)
    assert c[0] == pytest.approx(0.1) and c[1] == pytest.approx(0.9)


def test_selection_minimises_estimator_variance():
    """One vertex whose integrand g = a q + (1 - a) h mixes the guide's lobe
    q with light from the lower hemisphere h that only the BSDF (uniform
    here) reaches: descending the model's selection gradient on samples drawn
    at c = 0.5 reaches the c at which the variance of the one-sample estimate
    g / (c q + (1 - c) pB), computed by quadrature, is smallest, and not the
    KL-optimal c (which trusts the guide more)."""
    rng = np.random.default_rng(11)
    a, N, kappa = 0.7, 20000, 20.0
    # guide: one lobe about +z; BSDF: uniform over the sphere
    raw = np.zeros((1, 48))
    raw[0, 0:3] = (0.0, 0.0, 1.0)
    raw[0, 3] = math.log(kappa)
    raw[0, 4] = 10.0  # lobe 0 takes all the weight
    for k in range(1, 8):
        raw[0, 5 * k + 4] = -30.0
    pB0 = 1.0 / (4.0 * math.pi)
    from_guide = rng.uniform(size=N) < 0.5
    d = np.where(from_guide[:, None],
                 gr.sample_vmf(np.tile(raw[0, 0:3], (N, 1)), np.full(N, kappa), rng.uniform(size=N),
                               rng.uniform(size=N)),
                 _uniform_sphere(rng, N))
    rawN = np.repeat(raw, N, 0)
    q = gr.pdf(rawN, d)
    pB = np.full(N, pB0)
    pS = 0.5 * q + 0.5 * pB
    F = (a * q + (1 - a) * np.where(d[:, 2] < 0.0, 1.0 / (2.0 * math.pi), 0.0)) / pS
    sel = (0.02, 0.98)
    z = 0.0
    for _ in range(40):
        rawN[:, gr.SEL_OUT] = z
        _, g = gr.loss_and_grad_raw(rawN, d, F, pB=pB, sel=sel, pS=pS)
        z -= g[:, gr.SEL_OUT].sum()
    rawN[:, gr.SEL_OUT] = z
    c_learned = gr.selection(rawN[:1], *sel)[0]
    # quadrature over cos(theta): everything is symmetric about +z
    t = np.linspace(-1.0, 1.0, 200001)
    qt = gr.vmf_pdf(np.array([[[0.0, 0.0, 1.0]]] * len(t)), np.full((len(t), 1), kappa),
                    np.stack([np.sqrt(1 - t * t), np.zeros_like(t), t], -1))[:, 0]
    gt = a * qt + (1 - a) * np.where(t < 0.0, 1.0 / (2.0 * math.pi), 0.0)
    cs = np.linspace(0.02, 0.98, 961)
    m2 = [np.trapezoid(gt * gt / (c * qt + (1 - c) * pB0), t) for c in cs]
    kl = [-np.trapezoid(gt * np.log(c * qt + (1 - c) * pB0), t) for c in cs]
    c_var, c_kl = cs[int(np.argmin(m2))], cs[int(np.argmin(kl))]
    print("learned c %.3f, variance-optimal %.3f, KL-optimal %.3f" % (c_learned, c_var, c_kl))
    assert abs(c_learned - c_var) < 0.015
    assert abs(c_learned - c_kl) > 0.04


def _uniform_sphere(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


@pytest.mark.parametrize("kappa", [0.05, 1.0, 20.0, 300.0])
def test_vmf_normalised_and_mean(kappa):
    # integrate in (cos theta, phi) about mu = z: the density depends on cos only
    t = np.linspace(-1.0, 1.0, 400001)
    mu = np.array([[[0.0, 0.0, 1.0]]])
    d = np.stack([np.sqrt(1 - t * t), np.zeros_like(t), t], -1)
    p = gr.vmf_pdf(np.repeat(mu, len(t), 0), np.full((len(t), 1), kappa), d)[:, 0]
    total = 2 * math.pi * np.trapezoid(p, t)
    assert total == pytest.approx(1.0, rel=1e-6)
    mean_cos = 2 * math.pi * np.trapezoid(p * t, t)
    assert mean_cos == pytest.approx(1.0 / math.tanh(kappa) - 1.0 / kappa, rel=1e-5, abs=1e-9)


@pytest.mark.parametrize("kappa", [0.5, 5.0, 80.0])
def test_vmf_sampling_mean_direction(kappa):
    rng = np.random.default_rng(7)
    mu = rng.normal(size=3)
    mu /= np.linalg.norm(mu)
    N = 200000
    d = gr.sample_vmf(np.tile(mu, (N, 1)), np.full(N, kappa), rng.uniform(size=N), rng.uniform(size=N))
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-9)
    m = d.mean(axis=0)
    A = 1.0 / math.tanh(kappa) - 1.0 / kappa
    assert np.allclose(m, A * mu, atol=5.0 / math.sqrt(N))


def test_adam_first_step_closed_form():
    p = np.array([1.0, -2.0, 0.5], np.float32)
    g = np.array([0.3, -4.0, 1e-3], np.float32)
    m = np.zeros(3, np.float32)
    v = np.zeros(3, np.float32)
    p1, m1, v1 = gr.adam_step(p, m, v, g, t=1, lr=1e-3)
    # bias-corrected first step: m_hat = g, v_hat = g^2 -> lr * g / (|g| + eps)
    assert np.allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-6)


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.00390625, 1.005859375, 3.0e38, -0.0], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: nearest even keeps 1
    np.testing.assert_array_equal(gr.bf16(v)[:3], [1.0, 1.0, 1.0078125])


def test_encoding_octaves():
    """The positional encoding is sin / cos of 2^f pi x: unit pairs, each
    octave the angle double of the one below (sin 2t = 2 sin t cos t,
    cos 2t = 1 - 2 sin^2 t), exact values at dyadic points, and the normal
    and outgoing direction copied after the 72 encoded features."""
    rng = np.random.default_rng(5)
    x, n, wo, _, _ = _batch(rng, 64)
    x[0] = (0.5, 0.25, 0.0)
    f = gr.encode(x, n, wo)
    s = f[:, :72:2].reshape(-1, 3, 12)
    c = f[:, 1:72:2].reshape(-1, 3, 12)
    assert np.allclose(s ** 2 + c ** 2, 1.0, atol=1e-12)
    assert np.allclose(s[:, :, 1:], 2 * s[:, :, :-1] * c[:, :, :-1], atol=1e-9)
    assert np.allclose(c[:, :, 1:], 1 - 2 * s[:, :, :-1] ** 2, atol=1e-9)
    # sin(pi / 2) = 1, cos(pi / 2) = 0 (x = 0.5, f = 0); sin(pi / 4 * 2) at x = 0.25, f = 1
    assert abs(s[0, 0, 0] - 1.0) < 1e-15 and abs(c[0, 0, 0]) < 1e-15
    assert abs(s[0, 1, 1] - 1.0) < 1e-15 and abs(s[0, 2, :]).max() == 0.0
    assert np.array_equal(f[:, 72:75], n) and np.array_equal(f[:, 75:78], wo)
    assert not f[:, 78:].any()


def test_table_form_is_the_mixture_density():
    """The per-vertex table the path tracer keeps (lobe k as v_k = kappa_k
    mu_k and b_k = log(w_k norm_k) - kappa_k, density sum_k exp(v_k . d +
    b_k)) is the mixture density of the raw outputs, and it integrates to one
    over the sphere.  The GPU tests compare the device tables against this
    form; dropping the -kappa_k or the weight from b_k fails here."""
    rng = np.random.default_rng(4)
    raw = rng.normal(size=(6, gr.OUT_PAD)) * 2.0
    mu, kappa, w, _, _ = gr.mixture(raw)
    norm = kappa / (2.0 * math.pi * -np.expm1(-2.0 * kappa))
    v = mu * kappa[:, :, None]
    b = np.log(w * norm) - kappa
    # Fibonacci sphere quadrature
    n = 200000
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    phi = math.pi * (1.0 + 5.0 ** 0.5) * i
    r = np.sqrt(1.0 - z * z)
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    for q in range(raw.shape[0]):
        table = np.exp(d @ v[q].T + b[q]).sum(axis=1)
        np.testing.assert_allclose(table, gr.pdf(raw[q:q + 1].repeat(n, 0), d), rtol=1e-9, atol=1e-300)
        assert abs(table.mean() * 4.0 * math.pi - 1.0) < 2e-3
