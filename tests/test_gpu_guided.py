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
