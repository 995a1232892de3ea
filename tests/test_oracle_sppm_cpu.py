"""Pins the photon-mapping oracle without a GPU.

The reference has no photon-mapping tests, so the oracle of
core/src/PhotonMapping.cpp is pinned by:
  * a cross-check against the path-tracing oracle: with no concentrated lobe
    in the scene no photon is stored and a pass's radiance is exactly the
    path tracer's sample of the same index (same random stream);
  * a closed form: a mirror under a small light makes a caustic on a diffuse
    ceiling whose radiance is rho / pi times the irradiance of the light's
    mirror image (numerical integral over the mirrored quad);
  * per-photon closed forms: every stored photon of that scene carries
    emission * area / (1 - q_min) and arrives from below; with a diffuse
    floor nothing is stored (photons are stored only after a bounce);
  * brute force for the cell search (every point within r, in cell order)
    and the minimising property of the cell-size search;
  * the hole filler's documented quirk (the ring's fourth side is never
    searched, PhotonMapping.cpp:653-678) on a hand-computed case.
"""
import math

import numpy as np
import pytest

from hastyrenderer_amd import scenes
from oracle import pyoracle as po


@pytest.fixture(scope="module")
def mirror_job(tmp_path_factory):
    return scenes.write_mirror_caustic(str(tmp_path_factory.mktemp("mirror")), width=32, height=32, passes=64)


@pytest.fixture(scope="module")
def mirror_deep_job(tmp_path_factory):
    return scenes.write_mirror_caustic(str(tmp_path_factory.mktemp("mirror8")), width=16, height=16, passes=1,
                                       max_depth=8)


@pytest.fixture(scope="module")
def diffuse_floor_job(tmp_path_factory):
    return scenes.write_mirror_caustic(str(tmp_path_factory.mktemp("dfloor")), width=16, height=16, passes=1,
                                       max_depth=8, mirror=False)


def ceiling_points(job, w, h):
    rays = po.OracleJob(job).camera_rays(0).reshape(h, w, 6).astype(np.float64)
    o, d = rays[..., :3], rays[..., 3:]
    t = (2.0 - o[..., 1]) / d[..., 1]
    return o + d * t[..., None]


def test_photon_flux_closed_form(mirror_deep_job):
    """Light -> mirror (metallic, albedo 1: reflectance 1) -> ceiling: each
    stored photon has flux emission * A / (1 - q_min), q_min = 1 - 0.1^(1/8),
    and about (1 - q_min) of the emitted photons are stored."""
    sp = po.OracleSppm(mirror_deep_job, seed=3)
    n_emit = 20000
    ph = sp.trace_photons(0, 0, n_emit)
    qmin = 1.0 - 0.1 ** (1.0 / 8.0)
    area = 0.2 * 0.2
    expect = 10.0 * math.pi * area / (1.0 - qmin)
    assert len(ph) > 0
    # stored on the ceiling (y = 2) or on the back of the light (y = 1)
    y = ph[:, 1]
    assert np.all((np.abs(y - 2.0) < 1e-4) | (np.abs(y - 1.0) < 1e-4))
    # wi points back down to the mirror; ngwi = dot(n_g, wi) with n_g = -y
    # on the ceiling and on the light
    wi = ph[:, 3:6]
    assert np.all(wi[:, 1] < 0)
    np.testing.assert_allclose(ph[:, 6], -wi[:, 1], rtol=1e-5)
    np.testing.assert_allclose(ph[:, 7:10], expect, rtol=2e-5)
    # every surviving photon stores once; some miss the 40 x 40 mirror
    p = 1.0 - qmin
    sigma = math.sqrt(n_emit * p * (1 - p))
    assert abs(len(ph) - n_emit * p) < 5 * sigma + 0.01 * n_emit
    assert len(ph) <= n_emit * p + 5 * sigma


def test_no_photon_without_specular_bounce(diffuse_floor_job):
    sp = po.OracleSppm(diffuse_floor_job, seed=3)
    assert len(sp.trace_photons(0, 0, 5000)) == 0


def test_sppm_caustic_closed_form(mirror_job):
    """Radiance of ceiling pixels after 64 passes equals rho / pi * E(x).
    max_depth 2 keeps only the depth-0 gather; pixels near the light's shadow
    (x < 0.3) are left out."""
    w = h = 32
    sp = po.OracleSppm(mirror_job, seed=5, footprint_scale=2.0, store_fraction=16.0, emit_fraction=400.0)
    for k in range(64):
        color, normal, albedo = sp.run_pass(k, w, h)
    info = sp.info()
    assert info["stored"] >= 16 * w * h // 2
    x = ceiling_points(mirror_job, w, h)
    expect = scenes.mirror_caustic_radiance(x.reshape(-1, 3)).reshape(h, w)
    m = x[..., 0] > 0.3
    ratio = color[..., 0][m].mean() / expect[m].mean()
    assert abs(ratio - 1.0) < 0.05, ratio
    # grey scene: the three channels are identical
    np.testing.assert_array_equal(color[..., 0], color[..., 1])
    # the ceiling's normal and albedo at depth 0
    np.testing.assert_allclose(normal[m], np.tile([0.0, -1.0, 0.0], (m.sum(), 1)), atol=1e-6)
    np.testing.assert_allclose(albedo[m], 0.5)


def test_sppm_equals_path_tracer_without_concentrated_lobes(tmp_path):
    """Cornell box (diffuse only): no photon is stored, M = 0, so each pass's
    estimate L is the path tracer's sample of the same index; after n passes
    the film is their running mean (LSum / Nd)."""
    job = scenes.write_cornell(str(tmp_path / "c"), width=48, height=48, spp=3)
    w = h = 48
    sp = po.OracleSppm(job, seed=11)
    pt = po.OracleJob(job, seed=11)
    sums = None
    for k in range(3):
        c, n, a = sp.run_pass(k, w, h)
        assert sp.info()["stored"] == 0
        s, _ = pt.render(k, 1)
        if k == 0:
            np.testing.assert_array_equal(c, s[0])  # bit-exact: same stream, same operations
            np.testing.assert_array_equal(n, s[1])
            np.testing.assert_array_equal(a, s[2])
        sums = s[0].astype(np.float64) if sums is None else sums + s[0]
    np.testing.assert_allclose(c, sums / 3.0, rtol=1e-5, atol=1e-7)
    px = sp.pixels(w, h)
    assert np.all(px[:, 9] == 3) and np.all(px[:, 1] == 0) and np.all(px[:, 3:6] == 0)


def test_radius_neighbors_brute_force():
    rng = np.random.default_rng(1)
    pts = rng.uniform(-1, 1, (3000, 3)).astype(np.float32)
    for cs in (0.05, 0.13, 0.7):
        for _ in range(20):
            c = rng.uniform(-1.2, 1.2, 3).astype(np.float32)
            r = np.float32(rng.uniform(0.01, 0.4))
            got = po.radius_neighbors(pts, cs, c, r)
            d2 = np.sum((pts - c) ** 2, axis=1, dtype=np.float32)
            want = set(np.nonzero(d2 <= r * r)[0].tolist())
            assert set(got.tolist()) == want
            # order: cells in (x, y, z) lexicographic order, then insertion
            cell = np.floor(pts[got].astype(np.float32) / np.float32(cs)).astype(np.int64)
            key = [tuple(v) + (int(i),) for v, i in zip(cell, got)]
            assert key == sorted(key)


def test_cell_size_minimises_cost():
    def cost(r, cs):
        r = r.astype(np.float64)
        return np.sum(5.0 * (r / cs) ** 3 + np.maximum(r, cs) ** 3 * 1e5)

    rng = np.random.default_rng(2)
    for r in (np.full((16, 16), 0.01, np.float32), rng.uniform(0.005, 0.05, (24, 20)).astype(np.float32)):
        cs = po.cell_size(r)
        c0 = cost(r, cs)
        assert c0 <= cost(r, cs * 1.5) * (1 + 1e-6)
        assert c0 <= cost(r, cs * 0.6666) * (1 + 1e-6)
    # constant radius: the optimum of 5 (r/c)^3 + 1e5 c^3 (c > r) is
    # c* = (5e-5 r^3)^(1/6); the search lands within one step of it
    r = 0.01
    cs = po.cell_size(np.full((16, 16), r, np.float32))
    cstar = (5e-5 * r ** 3) ** (1.0 / 6.0)
    assert abs(math.log(cs / cstar)) <= math.log(1.5)


def test_hole_filler_skips_fourth_side():
    img = np.full((3, 3), -1.0, np.float32)
    img[2, 1] = 5.0  # pixel (i=1, j=2)
    # the rings around (1, 0) and (1, 1) reach row 2 only on their bottom
    # side (j2 = j + k), which is never searched: the default 1e-3 comes back
    assert po.get_close_non_negative(img, 1, 0) == pytest.approx(1e-3)
    assert po.get_close_non_negative(img, 1, 1) == pytest.approx(1e-3)
    # found on the right side (i + k), the left side (i - k), the top (j - k)
    assert po.get_close_non_negative(img, 0, 2) == pytest.approx(5.0)
    assert po.get_close_non_negative(img, 2, 2) == pytest.approx(5.0)
    img2 = np.full((3, 3), -1.0, np.float32)
    img2[0, 1] = 7.0  # pixel (1, 0), above (1, 1)
    assert po.get_close_non_negative(img2, 1, 1) == pytest.approx(7.0)
    # the sides are scanned together, step m along each: (i+1, j-1) on the
    # right side (m = 0) wins over (i-1, j) on the left side (m = 1)
    img3 = np.full((3, 3), -1.0, np.float32)
    img3[1, 0], img3[0, 2] = 1.0, 3.0  # pixels (0, 1) and (2, 0)
    assert po.get_close_non_negative(img3, 1, 1) == pytest.approx(3.0)


# ------------------------------------------------------- fixed-point flux ----
def test_fixed_point_gather_agrees_with_fp32(tmp_path):
    """Gathers summed as int64 multiples of 2^-24 agree with the FP32 sums
    of photonMapFluxEstimate (PhotonMapping.cpp:271-296) to float rounding:
    each term is rounded by at most 2^-25 and the FP32 sum's own rounding is
    larger.  Photon counts are identical (the set of photons is the same)."""
    job = scenes.write_caustic_scene(str(tmp_path / "c"), width=48, height=27, passes=2, n_major=48, n_minor=12)
    kw = dict(alpha=2.0 / 3.0, store_fraction=2.0, emit_fraction=100.0, initial_radius=0.064)
    a = po.OracleSppm(job, seed=5, **kw)
    b = po.OracleSppm(job, seed=5, fixed_point_flux=True, **kw)
    for k in range(2):
        ca, _, _ = a.run_pass(k, 48, 27)
        cb, _, _ = b.run_pass(k, 48, 27)
        assert b.info()["flux_overflows"] == 0
        assert np.allclose(cb, ca, rtol=2e-5, atol=1e-7)
    pa, pb = a.pixels(48, 27), b.pixels(48, 27)
    np.testing.assert_array_equal(pa[:, 1], pb[:, 1])  # N: same photon counts M
    np.testing.assert_array_equal(pa[:, 9], pb[:, 9])
    assert np.allclose(pb[:, 3:6], pa[:, 3:6], rtol=2e-5, atol=1e-9)


def test_fixed_point_guard_counts_saturated_gathers(tmp_path):
    """A light of radiance 1e12 makes single photon terms exceed 2^62 units
    of 2^-24: such gathers saturate and are counted, and the default light
    counts none."""
    job = scenes.write_caustic_scene(str(tmp_path / "hot"), width=32, height=18, passes=1, n_major=24, n_minor=8,
                                     light_radiance=1e12)
    o = po.OracleSppm(job, seed=1, store_fraction=2.0, emit_fraction=100.0, initial_radius=0.064,
                      fixed_point_flux=True)
    o.run_pass(0, 32, 18)
    assert o.info()["flux_overflows"] > 0
