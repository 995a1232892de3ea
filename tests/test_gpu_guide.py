"""The learned guide's GPU kernels against the CPU model (oracle/guide_ref.py,
itself pinned by finite differences): the BF16 MFMA lane layout, inference,
the fused forward/loss/backward + weight-gradient GEMM, Adam, and that
training lowers the loss on a fixed batch."""
import math

import numpy as np
import pytest

import hastyrenderer_amd as ha
from oracle import guide_ref as gr

from conftest import load_testkernels

pytestmark = pytest.mark.gpu


def _guide(cornell_job, seed=3, scale_out=10.0):
    sc = ha.Scene(cornell_job)
    g = ha.Guide(sc, seed=seed)
    g.set_box([0.0, 0.0, 0.0], [1.0, 1.0, 1.0])
    p = gr.init_params(seed)
    w4, b4, din, dout = gr.LAYOUT[4]
    p[w4:w4 + din * dout] *= scale_out
    g.set_params(p)
    return sc, g, p


def _queries(rng, B):
    x = rng.uniform(0, 1, (B, 3)).astype(np.float32)
    n = rng.normal(size=(B, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    wo = rng.normal(size=(B, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    return x, n.astype(np.float32), wo.astype(np.float32)


def test_mfma_lane_layout():
    L = load_testkernels()
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 9, (16, 16)).astype(np.float32) / 4  # exact in BF16
    B = rng.integers(-8, 9, (16, 16)).astype(np.float32) / 4
    D = np.zeros((16, 16), np.float32)
    assert L.tk_mfma(A, B, D) == 0
    np.testing.assert_array_equal(D, A @ B)


def test_infer_matches_reference(cornell_job):
    sc, g, p = _guide(cornell_job)
    rng = np.random.default_rng(1)
    x, n, wo = _queries(rng, 1000)
    raw = g.infer(np.concatenate([x, n, wo], 1))
    ref, _ = gr.forward(p, gr.encode(x, n, wo), bf16_ops=True)
    err = np.abs(raw - ref)
    scale = np.sqrt(np.mean(ref[:, :40] ** 2))
    print("infer: median |err| / rms %.2e, max %.2e" % (np.median(err) / scale, err.max() / scale), flush=True)
    assert np.median(err) <= 1e-5 * scale
    assert np.mean(err[:, :41] <= 1e-3 * scale) >= 0.99
    assert np.all(raw[:, 41:] == 0.0)


def _tables_reference(raw):
    """The mixture table the path tracer keeps per guided vertex, from raw
    outputs in float64: lobe k as (kappa_k mu_k, log(w_k norm_k) - kappa_k)
    with norm_k = kappa_k / (2 pi (1 - exp(-2 kappa_k))), then the weights."""
    mu, kappa, w, _, _ = gr.mixture(raw)
    norm = kappa / (2.0 * math.pi * -np.expm1(-2.0 * kappa))
    v = mu * kappa[:, :, None]
    b = np.log(w * norm) - kappa
    return v, b, w


def _inference_tables(L, B, out_scale, seed):
    """Mixture tables of B random queries by the path tracer's inference
    code (test kernel), and the CPU model's raw outputs for them."""
    rng = np.random.default_rng(seed)
    p = gr.init_params(seed + 2)
    w4, b4, din4, dout4 = gr.LAYOUT[4]
    p[w4:w4 + din4 * dout4] *= out_scale
    for li, (w, b, din, dout) in enumerate(gr.LAYOUT):  # biases exercise the LDS bias rows
        used = gr.N_OUT_USED if li == len(gr.LAYOUT) - 1 else dout
        p[b:b + used] = rng.normal(0.0, 0.1, used).astype(np.float32)
    wT = np.concatenate([gr.bf16(p[w:w + din * dout]).astype(np.float32).view(np.uint32) >> 16
                         for (w, b, din, dout) in gr.LAYOUT]).astype(np.uint16)
    lo = np.array([-1.0, 0.5, -2.0], np.float32)
    inv = np.array([0.5, 0.25, 0.125], np.float32)
    xn, n, wo = _queries(rng, B)
    x = (xn / inv + lo).astype(np.float32)
    q9 = np.ascontiguousarray(np.concatenate([x, n, wo], 1), np.float32)
    tables = np.zeros((B, 10, 4), np.float32)
    logits = np.zeros(B, np.float32)
    assert L.tk_guide_tables(wT, wT.size, p, p.size, lo, inv, q9, B, tables, logits) == 0
    xq = ((x - lo) * inv).astype(np.float32)  # the kernel's normalisation, in float32
    raw, _ = gr.forward(p, gr.encode(xq, n, wo), bf16_ops=True)
    return tables, logits, raw


def test_inference_tables_match_reference():
    """The path tracer's inference code (the network staged in LDS in
    lobe-major row order, guide::tileOutputs with the softmax across a
    query's four lanes, guide::lobeEntry) against the CPU model: every lobe's
    kappa mu, log-weight term and weight, and the selection logit.  A
    permuted lobe, a swapped weight or a lane group left out of the softmax
    gives O(1) errors; BF16 rounding flips give rare ~1e-3 ones.  B = 1000
    is not a multiple of 16: the last tile is ragged."""
    L = load_testkernels()
    B = 1000
    tables, logits, raw = _inference_tables(L, B, 10.0, 5)
    v, b, w = _tables_reference(raw)
    dv = np.abs(tables[:, :8, :3] - v) / (1.0 + np.abs(v))
    db = np.abs(tables[:, :8, 3] - b) / (1.0 + np.abs(b))
    dw = np.abs(tables[:, 8:, :].reshape(B, 8) - w)
    e = np.maximum(np.maximum(dv.max(axis=(1, 2)), db.max(axis=1)), dw.max(axis=1))
    scale = np.sqrt(np.mean(raw[:, :40] ** 2))
    dl = np.abs(logits - raw[:, gr.SEL_OUT]) / scale
    print("tables: median %.2e, 99%% %.2e, max %.2e; logit median %.2e" % (
        np.median(e), np.quantile(e, 0.99), e.max(), np.median(dl)), flush=True)
    assert np.median(e) <= 1e-4
    assert np.quantile(e, 0.99) <= 1e-2
    assert np.median(dl) <= 1e-5 and np.mean(dl <= 1e-3) >= 0.99
    assert np.allclose(tables[:, 8:, :].reshape(B, 8).sum(axis=1), 1.0, atol=1e-5)


def test_mixture_table_sampling_and_density():
    """guide::mixtureSampleTable and guide::mixturePdfTable (the guided
    vertex's sampling and density) on inference tables: the density equals
    the CPU model's mixture density, and the first and second moments of the
    sampled directions equal the mixture's within five standard errors.  For
    a vMF lobe with A = coth kappa - 1 / kappa, E[d] = A mu and
    E[d d^T] = (A / kappa) I + (1 - 3 A / kappa) mu mu^T.  A wrong lobe pick,
    a concentration off by 30 % in the sampler alone (checked on a CPU copy
    of the sampler: z 9 against 2 when right) or a wrong normalisation fails
    one of them."""
    L = load_testkernels()
    nq, m = 64, 16384
    tables, _, raw = _inference_tables(L, nq, 4.0, 9)
    rng = np.random.default_rng(11)
    xi = rng.random((nq * m, 3), dtype=np.float32)
    ev = rng.normal(size=(nq * m, 3))
    ev /= np.linalg.norm(ev, axis=1, keepdims=True)
    ev = ev.astype(np.float32)
    dirs = np.zeros((nq * m, 3), np.float32)
    pdfs = np.zeros(nq * m, np.float32)
    assert L.tk_guide_mixture(np.ascontiguousarray(tables), nq, m, xi, dirs, ev, pdfs) == 0
    assert np.all(np.abs(np.linalg.norm(dirs, axis=1) - 1.0) <= 1e-5)
    ref = np.concatenate([gr.pdf(raw[i:i + 1].repeat(m, 0), ev[i * m:(i + 1) * m].astype(np.float64))
                          for i in range(nq)])
    ok = ref > 1e-20
    rel = np.abs(pdfs[ok] - ref[ok]) / ref[ok]
    print("mixture pdf: median rel %.2e, 99%% %.2e" % (np.median(rel), np.quantile(rel, 0.99)), flush=True)
    assert np.median(rel) <= 1e-4 and np.quantile(rel, 0.99) <= 1e-2
    mu, kappa, w, _, _ = gr.mixture(raw)
    A = 1.0 / np.tanh(kappa) - 1.0 / kappa
    want1 = np.einsum("bk,bkj->bj", w * A, mu)
    want2 = (np.einsum("bk,ij->bij", w * A / kappa, np.eye(3)) +
             np.einsum("bk,bki,bkj->bij", w * (1.0 - 3.0 * A / kappa), mu, mu))
    zmax = 0.0
    for i in range(nq):
        d = dirs[i * m:(i + 1) * m].astype(np.float64)
        z1 = np.abs(d.mean(axis=0) - want1[i]) / (d.std(axis=0) / math.sqrt(m) + 1e-4)
        dd = np.einsum("ni,nj->nij", d, d)
        z2 = np.abs(dd.mean(axis=0) - want2[i]) / (dd.std(axis=0) / math.sqrt(m) + 1e-4)
        zmax = max(zmax, float(z1.max()), float(z2.max()))
    print("sample moments: max z %.2f" % zmax, flush=True)
    assert zmax <= 5.0


def _samples(rng, B, target=None):
    x, n, wo = _queries(rng, B)
    if target is None:
        wi = rng.normal(size=(B, 3))
        wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    else:
        wi = target + 0.15 * rng.normal(size=(B, 3))
        wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    F = rng.uniform(0.2, 2.0, B).astype(np.float32)
    return x, n, wo, wi.astype(np.float32), F


def test_gradient_matches_reference(cornell_job):
    sc, g, p = _guide(cornell_job)
    rng = np.random.default_rng(2)
    x, n, wo, wi, F = _samples(rng, 1000)
    grad, loss = g.gradient(np.concatenate([x, n, wo, wi, F[:, None]], 1))
    raw, cache = gr.forward(p, gr.encode(x, n, wo), bf16_ops=True)
    lref, g_raw = gr.loss_and_grad_raw(raw, wi.astype(np.float64), F.astype(np.float64))
    gref = gr.backward(p, cache, g_raw, bf16_ops=True)
    assert loss == pytest.approx(lref, rel=1e-3)
    rel = np.linalg.norm(grad - gref) / np.linalg.norm(gref)
    print("gradient: loss %.6f (ref %.6f), relative L2 error %.2e" % (loss, lref, rel), flush=True)
    assert rel < 1e-2
    for (w, b, din, dout) in gr.LAYOUT:
        a, r = grad[w:w + din * dout], gref[w:w + din * dout]
        cos = float(a @ r / (np.linalg.norm(a) * np.linalg.norm(r) + 1e-30))
        assert cos > 0.999
    # padding stays untouched
    w0, b0, din0, dout0 = gr.LAYOUT[0]
    assert np.all(grad[w0:w0 + din0 * dout0].reshape(dout0, din0)[:, 78:] == 0.0)


def test_selection_gradient_matches_reference(cornell_job):
    """Samples carrying the BSDF density of wi and the density it was drawn
    with (15 floats) under a learned selection range: the selection logit's
    gradient (trained through its own output row only) joins the mixture's,
    as the CPU model computes it; the same samples without a range leave
    that row untouched."""
    sc, g, p = _guide(cornell_job)
    rng = np.random.default_rng(6)
    x, n, wo, wi, F = _samples(rng, 1000)
    raw, cache = gr.forward(p, gr.encode(x, n, wo), bf16_ops=True)
    q = gr.pdf(raw, wi.astype(np.float64))
    pB = (q * rng.uniform(0.1, 4.0, 1000)).astype(np.float32)
    pB[::7] = 0.0
    pS = (0.5 * q + 0.5 * pB).astype(np.float32)
    sm = np.concatenate([x, n, wo, wi, F[:, None], pB[:, None], pS[:, None]], 1)
    w4, b4, din, dout = gr.LAYOUT[4]
    grad0, _ = g.gradient(sm)
    assert not grad0[w4 + gr.SEL_OUT * din:w4 + (gr.SEL_OUT + 1) * din].any() and grad0[b4 + gr.SEL_OUT] == 0.0
    g.set_selection(0.1, 0.9)
    grad, loss = g.gradient(sm)
    lref, g_raw = gr.loss_and_grad_raw(raw, wi.astype(np.float64), F.astype(np.float64),
                                       pB=pB.astype(np.float64), sel=(0.1, 0.9), pS=pS.astype(np.float64))
    gref = gr.backward(p, cache, g_raw, bf16_ops=True)
    # the hidden layers see the mixture's gradient only
    np.testing.assert_allclose(grad[:w4], grad0[:w4], rtol=1e-5, atol=1e-7)
    assert loss == pytest.approx(lref, rel=1e-3)
    rel = np.linalg.norm(grad - gref) / np.linalg.norm(gref)
    row, rrow = grad[w4 + gr.SEL_OUT * din:w4 + (gr.SEL_OUT + 1) * din], gref[w4 + gr.SEL_OUT * din:w4 + (gr.SEL_OUT + 1) * din]
    cos = float(row @ rrow / (np.linalg.norm(row) * np.linalg.norm(rrow) + 1e-30))
    print("selection gradient: loss %.6f (ref %.6f), relative L2 error %.2e, row cos %.6f" % (loss, lref, rel, cos),
          flush=True)
    assert rel < 1e-2 and cos > 0.999
    assert grad[b4 + gr.SEL_OUT] == pytest.approx(gref[b4 + gr.SEL_OUT], rel=1e-2)
    with pytest.raises(ha.HastyError):
        g.set_selection(0.5, 0.5)


def test_adam_step_matches_reference(cornell_job):
    sc, g, p = _guide(cornell_job)
    rng = np.random.default_rng(4)
    x, n, wo, wi, F = _samples(rng, 512)
    sm = np.concatenate([x, n, wo, wi, F[:, None]], 1)
    grad, _ = g.gradient(sm)
    g.train_step(sm, lr=1e-3)
    p1 = g.get_params()
    pref, _, _ = gr.adam_step(p.copy(), np.zeros_like(p), np.zeros_like(p), grad, t=1, lr=1e-3)
    close = np.abs(p1 - pref) <= 1e-6 + 1e-6 * np.abs(pref)
    assert close.mean() >= 0.999
    assert np.all(np.abs(p1 - p) <= 1.01e-3)


def test_training_lowers_loss(cornell_job):
    """Directions clustered about a fixed target: Adam on one fixed batch
    raises the learned density at the target well above uniform (1 / 4 pi)."""
    sc, g, p = _guide(cornell_job, scale_out=1.0)
    rng = np.random.default_rng(5)
    target = np.array([0.3, 0.8, -0.52])
    target /= np.linalg.norm(target)
    x, n, wo, wi, F = _samples(rng, 4096, target)
    sm = np.concatenate([x, n, wo, wi, F[:, None]], 1)
    first = g.train_step(sm, lr=3e-3)
    for _ in range(150):
        last = g.train_step(sm, lr=3e-3)
    print("training: loss %.4f -> %.4f" % (first, last), flush=True)
    assert last < first - 1.0
    raw = g.infer(np.concatenate([x[:64], n[:64], wo[:64]], 1)).astype(np.float64)
    dens = gr.pdf(raw, np.tile(target, (64, 1)))
    assert np.median(dens) > 5.0 / (4 * math.pi)
