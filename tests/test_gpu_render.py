"""Closest-hit queries and path-traced films on the GPU vs the CPU oracle.

Parity bar (per the path tracer's contract): triangle ids of closest hits are
identical (the watertight test with smallest-t / smallest-index tie-break is
order independent), and per-pixel radiance at equal spp is within 1e-3
relative on at least 99.9% of pixels.  Both sides use the same counter-based
RNG keys, so paths take the same decisions; only float reassociation of the
accumulated radiance differs.
"""
import os

import numpy as np
import pytest

import hastyrenderer_amd as ha
from hastyrenderer_amd import scenes
from conftest import rel_ok
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _random_rays(n, lo, hi, rng):
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


@pytest.mark.parametrize("which", ["cornell", "ggx_small"])
def test_closest_hit_matches_oracle(which, cornell_job, ggx_small_job):
    job = cornell_job if which == "cornell" else ggx_small_job
    rng = np.random.default_rng(5)
    scene = ha.Scene(job)
    oj = po.OracleJob(job)
    rays = _random_rays(20000, -1.2, 2.2, rng)
    cam = oj.camera_rays(0).reshape(-1, 6)
    rays = np.concatenate([rays, cam[::7]])
    gp, tri, tuv = scene.trace_closest(rays)
    otri, otuv, edge = oj.trace(rays)
    assert (tri == otri).all()
    hit = tri >= 0
    assert (tuv[hit].view(np.uint32) == otuv[hit].view(np.uint32)).all()
    for i in np.nonzero(hit)[0][:50]:
        assert tuple(gp[i]) == oj.tri_ids(tri[i])
    assert (gp[~hit] == 0xFFFFFFFF).all()


@pytest.mark.parametrize("which", ["cornell", "ggx_small"])
def test_trace_hits_records_match_oracle(which, cornell_job, ggx_small_job):
    """hasty_trace_hits: (geomID, primID, t, u, v) bit-equal to the oracle's
    BVH2 hit, Ng = normalize((v1 - v0) x (v2 - v0)) of the hit triangle
    computed in float64 from the oracle's positions (within 2 ulp-ish), and
    the Embree miss convention (0xFFFFFFFF ids, t = +inf, Ng = 0)."""
    job = cornell_job if which == "cornell" else ggx_small_job
    rng = np.random.default_rng(8)
    scene = ha.Scene(job)
    oj = po.OracleJob(job)
    rays = _random_rays(30000, -1.2, 2.2, rng)
    hits = scene.trace_hits(rays)
    otri, otuv, _ = oj.trace(rays)
    hit = otri >= 0
    assert hit.sum() > 1000 and (~hit).sum() > 100
    assert (hits["t"][hit].view(np.uint32) == otuv[hit, 0].view(np.uint32)).all()
    assert (hits["u"][hit].view(np.uint32) == otuv[hit, 1].view(np.uint32)).all()
    assert (hits["v"][hit].view(np.uint32) == otuv[hit, 2].view(np.uint32)).all()
    ids = np.array([oj.tri_ids(t) for t in otri[hit]], np.uint32)
    assert (hits["geom_id"][hit] == ids[:, 0]).all() and (hits["prim_id"][hit] == ids[:, 1]).all()
    for i in np.nonzero(hit)[0][:400]:
        p = oj.tri_pos(otri[i]).astype(np.float64)
        n = np.cross(p[1] - p[0], p[2] - p[0])
        assert np.allclose(hits["ng"][i], n / np.linalg.norm(n), atol=3e-7)
    assert (hits["geom_id"][~hit] == 0xFFFFFFFF).all() and (hits["prim_id"][~hit] == 0xFFFFFFFF).all()
    assert np.isinf(hits["t"][~hit]).all() and (hits["ng"][~hit] == 0).all()
    # the record agrees with the older (tri, t, u, v) query
    _, tri, tuv = scene.trace_closest(rays)
    assert (tri == otri).all()


def test_edge_rays_watertight(cornell_job):
    """Rays aimed exactly at shared vertices/edges of the walls never slip through."""
    scene = ha.Scene(cornell_job)
    oj = po.OracleJob(cornell_job)
    targets = []
    for y in (0.0, 1.0, 2.0):
        for x in (-1.0, -0.5, 0.0, 0.5, 1.0):
            targets.append((x, y, -1.0))  # back wall incl. its diagonal and corners
    targets = np.array(targets, np.float32)
    o = np.array([0.013, 1.0, 3.0], np.float32)
    d = targets - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([np.repeat(o[None], len(d), 0), d], axis=1).astype(np.float32)
    _, tri, _ = scene.trace_closest(rays)
    otri, _, _ = oj.trace(rays)
    btri, _, _ = oj.trace(rays, brute=True)
    assert (tri == otri).all() and (otri == btri).all()
    assert (tri >= 0).all()


def test_cornell_parity_full_config(cornell_job):
    """Procedural Cornell box, 256^2, 64 spp, max depth 8, RR from depth 3, NEE+MIS, seed 0."""
    scene = ha.Scene(cornell_job)
    info = scene.info()
    spp = info["num_samples"]
    assert (info["width"], info["height"], spp, info["max_depth"]) == (256, 256, 64, 8)
    c, n, a = scene.render_pt(seed=0, first_sample=0, num_samples=spp)
    oj = po.OracleJob(cornell_job, seed=0)
    (rc, rn, ra), nan = oj.render(0, spp, threads=16)
    ok = rel_ok(c, rc)
    frac = ok.mean()
    print("cornell: pixels within 1e-3: %.6f, mean gpu %s oracle %s" % (frac, c.mean((0, 1)) / spp,
                                                                        rc.mean((0, 1)) / spp))
    assert frac >= 0.999
    assert rel_ok(a, ra, 1e-5).mean() >= 0.9999
    assert rel_ok(n, rn, 1e-5).mean() >= 0.9999


def test_ggx_grid_parity_reduced(ggx_small_job):
    """GGX grid (env-map NEE, thin lens, conductors + dielectrics) at 96^2 x 4 spp."""
    scene = ha.Scene(ggx_small_job)
    spp = 4
    c, n, a = scene.render_pt(seed=1, first_sample=0, num_samples=spp)
    oj = po.OracleJob(ggx_small_job, seed=1)
    (rc, rn, ra), nan = oj.render(0, spp, threads=16)
    frac = rel_ok(c, rc).mean()
    print("ggx small: pixels within 1e-3: %.6f" % frac)
    assert frac >= 0.999


def test_full_ggx_grid_parity_sampled(ggx_job):
    """The benchmark GGX grid (250,882 triangles, 2048x1024 sky, f/2.8 lens) on a
    160x160 film crop of the same camera, 2 spp."""
    scene = ha.Scene(ggx_job)
    scene.set_settings(width=160, height=160)
    c, n, a = scene.render_pt(seed=1, first_sample=0, num_samples=2)
    oj = po.OracleJob(ggx_job, seed=1)
    oj.set(width=160, height=160)
    (rc, rn, ra), nan = oj.render(0, 2, threads=16)
    frac = rel_ok(c, rc).mean()
    print("ggx full scene: pixels within 1e-3: %.6f" % frac)
    assert frac >= 0.999


def test_cornell_sky_parity(cornell_sky_job):
    scene = ha.Scene(cornell_sky_job)
    c, n, a = scene.render_pt(seed=0, first_sample=0, num_samples=16)
    oj = po.OracleJob(cornell_sky_job, seed=0)
    (rc, rn, ra), nan = oj.render(0, 16, threads=16)
    frac = rel_ok(c, rc).mean()
    print("cornell sky: pixels within 1e-3: %.6f" % frac)
    assert frac >= 0.999


def test_caustic_scene_parity(caustic_small_job):
    """Glass with dispersion and media tracking, metal torus, small area light."""
    scene = ha.Scene(caustic_small_job)
    c, n, a = scene.render_pt(seed=3, first_sample=0, num_samples=16)
    oj = po.OracleJob(caustic_small_job, seed=3)
    (rc, rn, ra), nan = oj.render(0, 16, threads=16)
    frac = rel_ok(c, rc).mean()
    print("caustic: pixels within 1e-3: %.6f" % frac)
    assert frac >= 0.999
    assert scene.info()["num_triangles"] == oj.info()["num_tris"]


def test_deterministic_across_launch_shapes(cornell_job):
    """Bit-identical films whatever the slot pool (heavy regeneration with
    1000 slots) and grid sizes: each path's stream is keyed by (pixel,
    sample), not by the slot that runs it."""
    scene = ha.Scene(cornell_job)
    scene.set_settings(width=128, height=96)
    ref = scene.render_pt(seed=3, first_sample=5, num_samples=6)
    scene.set_launch(path_slots=1000, trace_blocks=64, shade_blocks=32)
    alt = scene.render_pt(seed=3, first_sample=5, num_samples=6)
    scene.set_launch(path_slots=1 << 24, trace_blocks=1024, shade_blocks=4096)
    alt2 = scene.render_pt(seed=3, first_sample=5, num_samples=6)
    for x, y, z in zip(ref, alt, alt2):
        assert (x.view(np.uint32) == y.view(np.uint32)).all()
        assert (x.view(np.uint32) == z.view(np.uint32)).all()


def test_incremental_passes_equal_one_call(cornell_job):
    """Adding samples over several calls equals one call (film sums in sample order)."""
    scene = ha.Scene(cornell_job)
    scene.set_settings(width=64, height=64)
    one = scene.render_pt(seed=0, first_sample=0, num_samples=8)
    films = tuple(np.zeros((64, 64, 3), np.float32) for _ in range(3))
    for k in range(0, 8, 3):
        scene.render_pt(seed=0, first_sample=k, num_samples=min(3, 8 - k), films=films)
    for x, y in zip(one, films):
        assert (x.view(np.uint32) == y.view(np.uint32)).all()


def test_render_thread_mirror(cornell_job):
    """The renderThread host mirror (batches of sample indices, stop flag,
    copyBuffer) yields the same sums as a direct call and computeEstimate
    divides by the sample count."""
    job = ha.RenderJob()
    job.loadJSON(cornell_job)
    job.scene.set_settings(width=48, height=32, num_samples=10)
    job.renderSettings = ha.RenderSettings(job.scene.info())
    bufs = [ha.Image3fAccDoubleBuffer(48, 32) for _ in range(3)]
    td = ha.RenderThreadData(seed=0)
    ha.renderThread(*bufs, job, td, batch=3)
    assert td.stoppedFlag
    assert bufs[0].getReadBuffer().numSamples == 10
    direct = job.scene.render_pt(seed=0, first_sample=0, num_samples=10)
    assert (bufs[0].getReadBuffer().data.view(np.uint32) == direct[0].view(np.uint32)).all()
    est = ha.computeEstimate([bufs[0]])
    assert np.allclose(est, direct[0] / np.float32(1e-7 + 10.0))


def test_stop_flag_stops_between_batches(cornell_job):
    job = ha.RenderJob()
    job.loadJSON(cornell_job)
    job.scene.set_settings(width=16, height=16, num_samples=12)
    job.renderSettings = ha.RenderSettings(job.scene.info())
    job.stopFlag = True
    bufs = [ha.Image3fAccDoubleBuffer(16, 16) for _ in range(3)]
    td = ha.RenderThreadData(seed=0)
    ha.renderThread(*bufs, job, td, batch=4)
    assert td.stoppedFlag and bufs[0].getWriteBuffer().numSamples == 0


def test_headless_cli_matches_library(cornell_job, caustic_small_job, tmp_path):
    """hastyrenderer_amd/hasty_render (C++ host layer with the reference's
    renderThread / renderPhotonThread signatures) writes what the library
    renders: PT estimate = sums / (1e-7 + samples), SPPM = last estimate."""
    import subprocess

    cli = os.path.join(os.path.dirname(ha.__file__), "hasty_render")
    out = tmp_path / "pt"
    out.mkdir()
    r = subprocess.run([cli, cornell_job, str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    img = ha.read_image(str(out / "output.exr"))
    info = ha.Scene(cornell_job).info()
    c, n, a = ha.Scene(cornell_job).render_pt(seed=0, first_sample=0, num_samples=info["num_samples"])
    np.testing.assert_allclose(img, c / np.float32(1e-7 + info["num_samples"]), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(ha.read_image(str(out / "outputAlbedo.exr")),
                               a / np.float32(1e-7 + info["num_samples"]), rtol=1e-6, atol=1e-9)
    out2 = tmp_path / "pm"
    out2.mkdir()
    r = subprocess.run([cli, "--sppm", caustic_small_job, str(out2)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    pm = ha.read_image(str(out2 / "output.exr"))
    job = ha.PMRenderJob()
    job.loadJSON(caustic_small_job)
    st = job.renderSettings
    bufs = [ha.Image3fAccDoubleBuffer(st.width, st.height) for _ in range(3)]
    ha.renderPhotonThread(*bufs, job, ha.RenderThreadData(seed=0))
    np.testing.assert_allclose(pm, bufs[0].getReadBuffer().data / np.float32(1e-7 + 1), rtol=1e-6, atol=1e-9)
    # reference error convention: load failure -> exit code 2
    r = subprocess.run([cli, str(tmp_path / "missing.json"), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2


def _nested_glass_job(folder, shells):
    """`shells` concentric smooth glass cubes (IOR 1.5, slightly tinted)
    around the camera's view axis under a white sky: a camera ray crosses
    every shell twice, entering `shells` nested media on the way in."""
    import json

    os.makedirs(folder, exist_ok=True)
    with open(os.path.join(folder, "m.mtl"), "w") as f:
        f.write("newmtl glass\nKd 0.9 0.95 0.9\nKs 0 0 0\nNs 1000\nNi 1.5\nd 0\n")
    lines = ["mtllib m.mtl"]
    nv = 0
    for k in range(shells):
        h = 0.4 + 0.12 * k
        v, tris = scenes._box((0.0, 0.0, 0.0), (2 * h, 2 * h, 2 * h), 0.0)
        lines += ["o shell%d" % k, "usemtl glass"]
        lines += ["v %r %r %r" % tuple(map(float, p)) for p in v]
        lines += ["f %d %d %d" % (a + 1 + nv, b + 1 + nv, c + 1 + nv) for a, b, c in tris]
        nv += len(v)
    with open(os.path.join(folder, "m.obj"), "w") as f:
        f.write("\n".join(lines) + "\n")
    json.dump({"camera": {"position": [0.05, 0.1, 4.0], "look_at": [0, 0, 0]},
               "world": {"background_color": [1.0, 1.0, 1.0]}, "mesh": "m.obj"},
              open(os.path.join(folder, "scene.json"), "w"))
    # a background colour without light triangles: BSDF sampling only (NEE
    # would need a light triangle, Scene.cpp:508-512)
    json.dump({"settings": {"width": 32, "height": 32, "num_samples": 4, "max_depth": 48,
                            "roulette_start_depth": 40, "use_mis": False}, "scene": "scene.json"},
              open(os.path.join(folder, "job.json"), "w"))
    return os.path.join(folder, "job.json")


def test_media_stack_within_and_beyond_eight_entries(tmp_path):
    """Seven nested glass media: no overflow, films match the oracle (whose
    stack is unbounded like the reference's).  Ten: the stack's ninth entry
    on is counted as an overflow (the pixels through the centre), and the
    render stays finite."""
    job7 = _nested_glass_job(str(tmp_path / "n7"), 7)
    sc = ha.Scene(job7)
    sc.set_profiling(events=False, counters=True)
    c, _, _ = sc.render_pt(seed=3, first_sample=0, num_samples=4)
    assert sc.counters()["media_overflow"] == 0
    (rc, _, _), _ = po.OracleJob(job7, seed=3).render(0, 4, threads=8)
    frac = rel_ok(c, rc).mean()
    print("7 nested media: %.5f of pixels within 1e-3 of the oracle" % frac, flush=True)
    assert frac >= 0.999
    job10 = _nested_glass_job(str(tmp_path / "n10"), 10)
    sc10 = ha.Scene(job10)
    sc10.set_profiling(events=False, counters=True)
    c10, _, _ = sc10.render_pt(seed=3, first_sample=0, num_samples=4)
    ov = sc10.counters()["media_overflow"]
    print("10 nested media: %d paths overflowed the 8-entry stack" % ov, flush=True)
    assert ov > 0 and np.isfinite(c10).all()
