"""Batching, graph replay and the autotuner: every job of a batch gets exactly
the films it gets when rendered alone (its paths' random streams are keyed by
(pixel, sample), so sharing the path pool with other jobs changes nothing),
graph replay equals eager launches, and any tuned launch configuration gives
the same images."""
import numpy as np
import pytest

import hastyrenderer_amd as ha
from hastyrenderer_amd import scenes
from oracle import pyoracle as po

from conftest import rel_ok

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cuboid_job(tmp_path_factory):
    return scenes.write_cuboid_scene(str(tmp_path_factory.mktemp("cuboids")), 3000, 80, 72, 6, seed=7)


def _alone(scene, seed, first, n):
    info = scene.info()
    f = ha.DeviceFilms(info["width"], info["height"])
    scene.render_pt_device(seed, first, n, *f.ptrs)
    out = f.download()
    f.close()
    return out


def _same(a, b):
    for x, y in zip(a, b):
        assert (x.view(np.uint32) == y.view(np.uint32)).all()


def test_cuboid_scene_parity(cuboid_job):
    """The batching workload's scenes (Lambertian, GGX and glass cuboids with
    media) against the oracle."""
    sc = ha.Scene(cuboid_job)
    c, n, a = sc.render_pt(seed=2, first_sample=0, num_samples=6)
    (rc, rn, ra), _ = po.OracleJob(cuboid_job, seed=2).render(0, 6, threads=16)
    frac = rel_ok(c, rc).mean()
    print("cuboids: pixels within 1e-3: %.6f" % frac, flush=True)
    assert frac >= 0.999


def test_batch_equals_jobs_alone(cornell_job, ggx_small_job, caustic_small_job, cuboid_job):
    specs = [(cornell_job, 5, 3, 6, (64, 48)), (ggx_small_job, 1, 0, 3, None), (caustic_small_job, 3, 1, 2, None),
             (cuboid_job, 9, 0, 5, None), (cornell_job, 6, 0, 4, (40, 40))]
    scs, alone = [], []
    for path, seed, first, n, size in specs:
        sc = ha.Scene(path)
        if size:
            sc.set_settings(width=size[0], height=size[1])
        scs.append(sc)
        alone.append(_alone(sc, seed, first, n))
    films = [ha.DeviceFilms(sc.info()["width"], sc.info()["height"]) for sc in scs]
    ha.render_batch([(sc, sp[1], sp[2], sp[3], f) for sc, sp, f in zip(scs, specs, films)])
    for f, ref in zip(films, alone):
        _same(f.download(), ref)
    # a small pool forces heavy regeneration across job boundaries
    scs[0].set_launch(path_slots=2000)
    for f in films:
        f.zero()
    ha.render_batch([(sc, sp[1], sp[2], sp[3], f) for sc, sp, f in zip(scs, specs, films)])
    for f, ref in zip(films, alone):
        _same(f.download(), ref)


def test_graph_replay_equals_eager(cornell_job):
    sc = ha.Scene(cornell_job)
    sc.set_settings(width=128, height=96)
    sc.set_launch(path_slots=4000)  # many chunks: the graph is replayed many times
    sc.set_launch_ex(use_graph=0)
    eager = _alone(sc, 3, 0, 12)
    sc.set_launch_ex(use_graph=1)
    g1 = _alone(sc, 3, 0, 12)
    g2 = _alone(sc, 3, 0, 12)
    _same(eager, g1)
    _same(eager, g2)
    # odd chunk lengths round up to an even number of bounces
    sc.set_launch_ex(chunk=3, use_graph=1)
    _same(eager, _alone(sc, 3, 0, 12))


def test_autotune_changes_only_time(cornell_job, ggx_small_job):
    sc = ha.Scene(ggx_small_job)
    ref = _alone(sc, 1, 0, 3)
    chosen = sc.autotune(seed=1, trial_samples=1)
    print("autotuned:", chosen, flush=True)
    assert sc.get_launch()["trace_blocks"] == chosen["trace_blocks"]
    _same(ref, _alone(sc, 1, 0, 3))
    # batch autotune
    sc2 = ha.Scene(cornell_job)
    ref2 = _alone(sc2, 0, 0, 2)
    ha.autotune_batch([(sc2, 0, 0, 1, None), (sc, 1, 0, 1, None)], trial_samples=1)
    _same(ref2, _alone(sc2, 0, 0, 2))


@pytest.mark.parametrize("which", ["cornell", "ggx_small", "caustic_small"])
def test_bvh_width_changes_only_time(which, request, cuboid_job):
    """4- and 8-wide BVHs of the same scene: identical hit records, films and
    photon-map passes, bit for bit (conservative box tests and the
    smallest-t / smallest-index tie rule make the closest hit independent of
    the tree); a batch mixing widths equals its jobs alone."""
    job = request.getfixturevalue(which + "_job")
    sc = ha.Scene(job)
    assert sc.info()["bvh_width"] == 4
    rng = np.random.default_rng(3)
    o = rng.uniform(-1.5, 2.5, (20000, 3))
    d = rng.normal(size=(20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    hits4 = sc.trace_hits(rays)
    f4 = _alone(sc, 2, 0, 3)
    sc.set_launch_ex(bvh_width=8)
    info = sc.info()
    print("%s: 8-wide BVH %d nodes, depth %d" % (which, info["bvh_nodes"], info["bvh_depth"]), flush=True)
    assert info["bvh_width"] == 8 and sc.get_launch()["bvh_width"] == 8
    assert np.array_equal(sc.trace_hits(rays).view(np.uint8), hits4.view(np.uint8))
    _same(_alone(sc, 2, 0, 3), f4)
    if which == "caustic_small":
        outs = []
        for w in (4, 8):
            sc.set_launch_ex(bvh_width=w)
            g = ha.Sppm(sc, 5)
            outs.append([g.pass_(k)[0].copy() for k in range(2)])
        for x, y in zip(*outs):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    # a batch whose jobs use different widths
    other = ha.Scene(cuboid_job)
    ref = _alone(other, 4, 0, 2)
    sc.set_launch_ex(bvh_width=8)
    films = [ha.DeviceFilms(s.info()["width"], s.info()["height"]) for s in (sc, other)]
    ha.render_batch([(sc, 2, 0, 3, films[0]), (other, 4, 0, 2, films[1])])
    _same(films[0].download(), f4)
    _same(films[1].download(), ref)
