"""hasty_scene_create (a scene from the arrays Scene::Scene holds, read here
by an independent Python reader) against hasty_scene_load of the same files:
identical scene summaries, hit records and films, bit for bit."""
import numpy as np
import pytest

import hastyrenderer_amd as ha
from scene_reader import read_job

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["cornell", "ggx_small", "caustic_small", "cornell_sky"])
def test_create_from_descriptor_equals_load(which, request):
    job = request.getfixturevalue(which + "_job")
    print("[create] %s" % which, flush=True)
    kw, _ = read_job(job)
    a = ha.Scene(job)
    b = ha.Scene.create(**kw)
    ia, ib = a.info(), b.info()
    for k in ("num_triangles", "num_shapes", "num_materials", "num_light_triangles", "has_environment", "bvh_nodes",
              "bvh_depth", "width", "height", "max_depth", "num_samples", "use_mis"):
        assert ia[k] == ib[k], k
    rng = np.random.default_rng(4)
    o = rng.uniform(-1.5, 2.5, (20000, 3))
    d = rng.normal(size=(20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    assert np.array_equal(a.trace_hits(rays).view(np.uint8), b.trace_hits(rays).view(np.uint8))
    spp = 2
    if ia["width"] * ia["height"] > 64 * 64:
        for s in (a, b):
            s.set_settings(width=64, height=48)
    fa = a.render_pt(3, 0, spp)
    fb = b.render_pt(3, 0, spp)
    for x, y in zip(fa, fb):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert np.isfinite(fa[0]).all() and fa[0].sum() > 0

