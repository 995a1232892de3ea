"""Guided rendering: unbiased against the plain tracer, and the trained guide
lowers the error at equal samples on a room lit indirectly (the light faces
the ceiling, so next-event estimation sees its back from almost everywhere
and the BSDF rarely finds the bright patch)."""
import numpy as np
import pytest

import hastyrenderer_amd as ha
from hastyrenderer_amd import scenes

pytestmark = pytest.mark.gpu


def relmse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


@pytest.fixture(scope="module")
def room_job(tmp_path_factory):
    return scenes.write_cornell(str(tmp_path_factory.mktemp("room")), width=48, height=48, spp=64, light_up=True)


@pytest.fixture(scope="module")
def room_reference(room_job):
    sc = ha.Scene(room_job)
    n = 16384
    c, _, _ = sc.render_pt(seed=11, first_sample=0, num_samples=n)
    return c / n


def test_guided_render_unbiased(room_job, room_reference):
    sc = ha.Scene(room_job)
    g = ha.Guide(sc, seed=0)
    sc.set_guide(g, fraction=0.5, train_lr=1e-3)
    n = 1024
    c, nrm, alb = sc.render_pt(seed=3, first_sample=0, num_samples=n)
    loss, ns = sc.guide_stats()
    print("guided online training: last loss %.4f on %d samples" % (loss, ns), flush=True)
    assert ns > 0 and np.isfinite(loss)
    img = c / n
    assert np.all(np.isfinite(img))
    ratio = img.mean() / room_reference.mean()
    print("guided / plain image mean: %.4f" % ratio, flush=True)
    assert abs(ratio - 1.0) < 0.02
    # depth-0 aux buffers do not depend on the sampling strategy
    plain = ha.Scene(room_job)
    _, pn, pa = plain.render_pt(seed=3, first_sample=0, num_samples=4)
    _, gn, ga = sc.render_pt(seed=3, first_sample=0, num_samples=4)
    np.testing.assert_allclose(gn, pn, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("max_depth,max_throughput", [(2, 3.0e38), (255, 0.5)])
def test_guided_depth_and_throughput_limits_unbiased(room_job, room_reference, max_depth, max_throughput):
    """Guiding limited to shallow vertices or to paths of small throughput:
    the inference kernel and the shading kernel pick the same vertices (the
    stored throughput before the vertex's MIS weight), so every guided vertex
    has its mixture, and the image mean stays the plain tracer's."""
    sc = ha.Scene(room_job)
    g = ha.Guide(sc, seed=0)
    sc.set_guide_depth(max_depth, max_throughput)
    sc.set_guide(g, fraction=0.5, train_lr=1e-3)
    n = 1024
    c, _, _ = sc.render_pt(seed=3, first_sample=0, num_samples=n)
    img = c / n
    assert np.all(np.isfinite(img))
    ratio = img.mean() / room_reference.mean()
    print("limits (%d, %g): guided / plain image mean %.4f" % (max_depth, max_throughput, ratio), flush=True)
    assert abs(ratio - 1.0) < 0.02


def test_learned_selection_unbiased_and_trained(room_job, room_reference):
    """The guide/BSDF selection probability learned per vertex: still
    unbiased, every probability inside its range, and training moves it
    away from where it starts (0.5, the middle of the range)."""
    sc = ha.Scene(room_job)
    g = ha.Guide(sc, seed=0)
    g.set_selection(0.1, 0.9)
    sc.set_guide(g, fraction=0.5, train_lr=1e-3)
    n = 1024
    c, _, _ = sc.render_pt(seed=3, first_sample=0, num_samples=n)
    img = c / n
    assert np.all(np.isfinite(img))
    ratio = img.mean() / room_reference.mean()
    print("learned selection: guided / plain image mean %.4f" % ratio, flush=True)
    assert abs(ratio - 1.0) < 0.02
    # the selection the trained guide assigns to hit points seen from the camera
    rng = np.random.default_rng(0)
    q = np.zeros((2048, 9), np.float32)
    q[:, 0:3] = rng.uniform(0.05, 0.95, (2048, 3))
    q[:, 3:6] = (0.0, 1.0, 0.0)  # floor-facing normal
    d = rng.normal(size=(2048, 3))
    d[:, 1] = np.abs(d[:, 1])
    q[:, 6:9] = d / np.linalg.norm(d, axis=1, keepdims=True)
    raw = g.infer(q)
    sel = 0.1 + 0.8 / (1.0 + np.exp(-raw[:, 40].astype(np.float64)))
    print("learned selection on the floor: mean %.3f, range [%.3f, %.3f]" % (sel.mean(), sel.min(), sel.max()),
          flush=True)
    assert np.all((sel >= 0.1) & (sel <= 0.9))
    assert abs(sel.mean() - 0.5) > 0.02 or sel.std() > 0.02


def test_trained_guide_lowers_error(room_job, room_reference):
    spp = 64
    plain = ha.Scene(room_job)
    pc, _, _ = plain.render_pt(seed=5, first_sample=0, num_samples=spp)
    sc = ha.Scene(room_job)
    g = ha.Guide(sc, seed=0)
    sc.set_guide(g, fraction=0.5, train_lr=1e-3)
    sc.render_pt(seed=7, first_sample=0, num_samples=512)  # online training
    sc.set_guide(g, fraction=0.5, train_lr=0.0)
    gc, _, _ = sc.render_pt(seed=5, first_sample=0, num_samples=spp)
    e_plain = relmse(pc / spp, room_reference)
    e_guided = relmse(gc / spp, room_reference)
    print("relMSE at %d spp: plain %.5f, guided %.5f (x%.2f)" % (spp, e_plain, e_guided, e_plain / e_guided),
          flush=True)
    assert e_guided < e_plain


def _trimmed(img, ref, keep=0.999):
    e = np.sort(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2), axis=-1).ravel())
    return float(np.mean(e[: int(e.size * keep)]))


@pytest.fixture(scope="module")
def ring_job(tmp_path_factory):
    """The metal-ring box (dispersive prism, GGX gold torus, 1 % ceiling
    light) at 128^2, depth 32, with a reduced torus."""
    return scenes.write_caustic_scene(str(tmp_path_factory.mktemp("ring")), width=128, height=128, passes=64,
                                      n_major=192, n_minor=48, max_depth=32)


def _photon_guided(job, train_spp=256, selection=None):
    sc = ha.Scene(job)
    g = ha.Guide(sc, seed=0)
    if selection:
        g.set_selection(*selection)
    sc.set_guide_training(minibatch=16384, weight_cap=100.0)
    sc.set_guide_photons(1 << 21, 0.5)
    sc.set_guide(g, fraction=0.5, train_lr=1e-3)
    sc.render_pt(seed=77, first_sample=0, num_samples=train_spp)  # online training with caustic photons
    loss, ns = sc.guide_stats()
    assert ns > 0 and np.isfinite(loss)
    sc.set_guide(g, fraction=0.5, train_lr=0.0)
    return sc, g


@pytest.fixture(scope="module")
def ring_reference(ring_job):
    plain = ha.Scene(ring_job)
    ref, _, _ = plain.render_pt(seed=101, first_sample=0, num_samples=16384)
    return ref / 16384


@pytest.mark.parametrize("selection", [None, (0.05, 0.95)], ids=["fixed", "learned"])
def test_photon_trained_guide_unbiased_and_gains(ring_job, ring_reference, selection):
    """Caustic light paths as training samples, with the fixed guide
    fraction 0.5 or the learned selection: the guided image agrees with the
    plain tracer's in image mean and 4 x 4 region means (batches give the
    Monte Carlo error), and at equal samples its error outside the worst
    0.1 % of pixels is lower than the plain tracer's and than a guide
    trained without the photons."""
    plain = ha.Scene(ring_job)
    ref = ring_reference
    sc, g = _photon_guided(ring_job, selection=selection)
    batches, spp = 8, 256
    pm, gm = [], []
    for b in range(batches):
        pm.append(plain.render_pt(seed=21, first_sample=b * spp, num_samples=spp)[0] / spp)
        gm.append(sc.render_pt(seed=22, first_sample=b * spp, num_samples=spp)[0] / spp)
    reg = lambda x: np.array([[x[i * 32:(i + 1) * 32, j * 32:(j + 1) * 32].mean() for j in range(4)]
                              for i in range(4)])
    P, G = np.array([reg(x) for x in pm]), np.array([reg(x) for x in gm])
    se = np.sqrt(P.var(0, ddof=1) / batches + G.var(0, ddof=1) / batches)
    z = np.abs(G.mean(0) - P.mean(0)) / np.maximum(se, 1e-12)
    pimg, gimg = np.mean([x.mean() for x in pm]), np.mean([x.mean() for x in gm])
    se_img = np.sqrt(np.var([x.mean() for x in pm], ddof=1) / batches + np.var([x.mean() for x in gm], ddof=1) / batches)
    print("photon-trained guide (%s): image mean plain %.5f guided %.5f (z %.2f), region z max %.2f" % (
        selection, pimg, gimg, abs(gimg - pimg) / se_img, z.max()), flush=True)
    assert abs(gimg - pimg) < 4 * se_img and z.max() < 5.0
    # equal samples: trimmed relMSE, guided with photons < plain and < guided without photons
    pe = _trimmed(np.mean(pm, 0), ref)
    ge = _trimmed(np.mean(gm, 0), ref)
    sc2 = ha.Scene(ring_job)
    g2 = ha.Guide(sc2, seed=0)
    sc2.set_guide_training(minibatch=16384, weight_cap=100.0)
    sc2.set_guide(g2, fraction=0.5, train_lr=1e-3)
    sc2.render_pt(seed=77, first_sample=0, num_samples=256)
    sc2.set_guide(g2, fraction=0.5, train_lr=0.0)
    ne = _trimmed(np.mean([sc2.render_pt(seed=22, first_sample=b * spp, num_samples=spp)[0] / spp
                           for b in range(batches)], 0), ref)
    print("trimmed relMSE at %d spp: plain %.5f, guided %.5f, guided without photons %.5f" % (
        batches * spp, pe, ge, ne), flush=True)
    assert ge < 0.8 * pe and ge < ne
