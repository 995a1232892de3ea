"""Parity of the GPU photon mapper with the oracle (core/src/PhotonMapping.cpp).

Photons are compared one by one in emission order; the cell size within
float rounding (the GPU folds the cost sums in a tree); films after several
passes per pixel within 1e-3 relative (the GPU accumulates a path's terms
forward, the oracle bottom-up, and gathers with a different cell size may
sum photons in another order)."""
import math

import numpy as np
import pytest

from hastyrenderer_amd import scenes
from oracle import pyoracle as po

from conftest import rel_ok

pytestmark = pytest.mark.gpu


def _gpu_sppm(job, seed, w=0, h=0, **kw):
    import hastyrenderer_amd as ha

    sc = ha.Scene(job)
    if w and h:
        sc.set_settings(width=w, height=h)
    return sc, ha.Sppm(sc, seed, **kw)


def _photon_rows(ph):
    """GPU rows (x, ngwi, wi, flux) -> oracle layout (x, wi, ngwi, flux)."""
    return np.concatenate([ph[:, 0:3], ph[:, 4:7], ph[:, 3:4], ph[:, 7:10]], axis=1)


def test_photons_match_oracle(caustic_small_job):
    w, h = 192, 108
    sc, g = _gpu_sppm(caustic_small_job, seed=9)
    ref = po.OracleSppm(caustic_small_job, seed=9)
    for k in range(2):
        g.pass_(k)
        info = g.info()
        print(f"pass {k}: emitted {info['emitted']} stored {info['stored']} cell {info['cell_size']}", flush=True)
        ph = _photon_rows(g.photons())
        want = ref.trace_photons(k, 0, info["emitted"])
        assert info["emitted"] == min(info["emit_limit"], info["emitted"])
        assert len(ph) == len(want) == info["stored"]
        assert info["stored"] > 1000
        exact = np.all(ph == want[:, :10], axis=1).mean()
        print(f"photons bit-exact: {exact:.5f}", flush=True)
        assert exact > 0.999
        np.testing.assert_allclose(ph, want[:, :10], rtol=1e-4, atol=1e-6)
    # the emission loop's stop rule: stored < store limit before the last photon
    assert info["stored"] >= info["store_limit"] or info["emitted"] == info["emit_limit"]


def test_sppm_passes_match_oracle(caustic_small_job):
    w, h = 192, 108
    sc, g = _gpu_sppm(caustic_small_job, seed=4)
    ref = po.OracleSppm(caustic_small_job, seed=4)
    for k in range(3):
        c, n, a = g.pass_(k)
        rc, rn, ra = ref.run_pass(k, w, h)
        gi, ri = g.info(), ref.info()
        assert gi["emitted"] == ri["emitted"] and gi["stored"] == ri["stored"]
        assert gi["cell_size"] == pytest.approx(ri["cell_size"], rel=1e-4)
        ok = rel_ok(c, rc)
        print(f"pass {k}: {ok.mean():.5f} of pixels within 1e-3", flush=True)
        assert ok.mean() >= 0.999
        assert rel_ok(n, rn, rtol=1e-5).mean() >= 0.9995
        assert rel_ok(a, ra, rtol=1e-5).mean() >= 0.9995
    px = g.pixels()
    rp = ref.pixels(w, h)
    np.testing.assert_array_equal(px["Nd"].ravel(), rp[:, 9].astype(np.int32))
    assert (np.abs(px["N"].ravel() - rp[:, 1]) <= 1e-3 * np.maximum(1.0, rp[:, 1])).mean() >= 0.999
    assert (np.abs(px["radius"].ravel() - rp[:, 0]) <= 1e-4 * rp[:, 0]).mean() >= 0.999
    np.testing.assert_allclose(g.radii().ravel(), ref.radii(w, h).ravel(), rtol=1e-4)


def test_sppm_cornell_matches_oracle(cornell_job):
    w = h = 64
    sc, g = _gpu_sppm(cornell_job, seed=2, w=w, h=h)
    ref = po.OracleSppm(cornell_job, seed=2, width=w, height=h)
    for k in range(2):
        c, n, a = g.pass_(k)
        rc, rn, ra = ref.run_pass(k, w, h)
        assert g.info()["stored"] == 0
        assert rel_ok(c, rc).mean() >= 0.999


def test_sppm_shared_gather_matches_oracle(caustic_small_job):
    """Dense photon maps (16 x W H photons stored, radius 20x the footprint):
    waves with a query of more than 512 candidate photons take the shared
    gather (lanes split the owner's cell lookups and photons); the image and
    the photon counts still follow the oracle's sequential gather."""
    w, h = 96, 54
    kw = dict(alpha=0.4, footprint_scale=20.0, store_fraction=16.0, emit_fraction=4000.0)
    sc, g = _gpu_sppm(caustic_small_job, seed=3, w=w, h=h, **kw)
    ref = po.OracleSppm(caustic_small_job, seed=3, width=w, height=h, **kw)
    for k in range(2):
        c, n, a = g.pass_(k)
        rc, rn, ra = ref.run_pass(k, w, h)
        gi, ri = g.info(), ref.info()
        assert gi["emitted"] == ri["emitted"] and gi["stored"] == ri["stored"]
        ok = rel_ok(c, rc)
        print(f"pass {k}: {gi['stored']} photons, {ok.mean():.5f} of pixels within 1e-3", flush=True)
        assert ok.mean() >= 0.999
        if k == 0:
            found = ref.pixels(w, h)[:, 1] / kw["alpha"]
            print(f"photons found per pixel: mean {found.mean():.0f}, max {found.max():.0f}", flush=True)
            assert (found > 2048).mean() > 0.05  # many waves take the shared gather
            np.testing.assert_allclose(g.pixels()["N"].ravel(), ref.pixels(w, h)[:, 1], rtol=1e-6)


def test_sppm_ragged_image_matches_oracle(caustic_small_job):
    """97 x 53 pixels (not a multiple of 64): the gather kernel's last wave
    has lanes without a pixel, which still serve the shared gathers."""
    w, h = 97, 53
    kw = dict(alpha=0.4, footprint_scale=20.0, store_fraction=16.0, emit_fraction=4000.0)
    sc, g = _gpu_sppm(caustic_small_job, seed=5, w=w, h=h, **kw)
    ref = po.OracleSppm(caustic_small_job, seed=5, width=w, height=h, **kw)
    c, n, a = g.pass_(0)
    rc, rn, ra = ref.run_pass(0, w, h)
    assert g.info()["stored"] == ref.info()["stored"]
    ok = rel_ok(c, rc)
    print(f"{ok.mean():.5f} of pixels within 1e-3", flush=True)
    assert ok.mean() >= 0.999
    np.testing.assert_allclose(g.pixels()["N"].ravel(), ref.pixels(w, h)[:, 1], rtol=1e-6)


def test_sppm_caustic_closed_form(tmp_path):
    """Mirror caustic on a diffuse ceiling: rho / pi * E(x) within 1%."""
    w = h = 64
    job = scenes.write_mirror_caustic(str(tmp_path / "m"), width=w, height=h, passes=32)
    sc, g = _gpu_sppm(job, seed=1, footprint_scale=2.0, store_fraction=64.0, emit_fraction=1600.0)
    for k in range(32):
        c, n, a = g.pass_(k)
    rays = po.OracleJob(job).camera_rays(0).reshape(h, w, 6).astype(np.float64)
    o, d = rays[..., :3], rays[..., 3:]
    x = o + d * ((2.0 - o[..., 1]) / d[..., 1])[..., None]
    expect = scenes.mirror_caustic_radiance(x.reshape(-1, 3)).reshape(h, w)
    m = x[..., 0] > 0.3
    ratio = c[..., 0][m].mean() / expect[m].mean()
    print(f"caustic radiance / closed form = {ratio:.4f} ({g.info()['stored']} photons per pass)", flush=True)
    assert abs(ratio - 1.0) < 0.01


def test_render_photon_thread_mirror(caustic_small_job):
    import hastyrenderer_amd as ha

    job = ha.PMRenderJob()
    job.loadJSON(caustic_small_job)
    st = job.renderSettings
    bufs = [ha.Image3fAccDoubleBuffer(st.width, st.height) for _ in range(3)]
    td = ha.RenderThreadData(seed=4)
    ha.renderPhotonThread(*bufs, job, td)
    assert td.stoppedFlag and job.sppm.info()["passes_done"] == st.numSamples
    assert bufs[0].getReadBuffer().numSamples == 1
    ref = po.OracleSppm(caustic_small_job, seed=4)
    for k in range(st.numSamples):
        rc, _, _ = ref.run_pass(k, st.width, st.height)
    assert rel_ok(bufs[0].getReadBuffer().data, rc).mean() >= 0.999
    est = ha.computeEstimate([bufs[0]])
    assert np.allclose(est, bufs[0].getReadBuffer().data, rtol=1e-6)


def test_sppm_bitexact_run_to_run(caustic_small_job):
    """Two jobs with the same seed give bit-identical films and statistics:
    every pixel sums its photons in a fixed order (cells x -> y -> z, photons
    in emission order), so no atomics or fixed point are needed."""
    outs = []
    for _ in range(2):
        sc, g = _gpu_sppm(caustic_small_job, seed=6)
        for k in range(2):
            c, n, a = g.pass_(k)
        outs.append((c.copy(), g.pixels()["flux"].copy(), g.radii().copy()))
        g.close()
        sc.close()
    for x, y in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(x, y)


# ------------------------------------------------------- fixed-point flux ----
_FIXED_KW = dict(alpha=2.0 / 3.0, store_fraction=4e6 / (1920 * 1080), emit_fraction=100.0, initial_radius=0.064)


def test_fixed_point_flux_matches_oracle(caustic_small_job):
    """int64 fixed-point gathers on the GPU against the oracle's fixed-point
    gathers: photon counts (N) identical, films within 1e-3."""
    w, h = 192, 108
    sc, g = _gpu_sppm(caustic_small_job, seed=8, fixed_point_flux=True, **_FIXED_KW)
    ref = po.OracleSppm(caustic_small_job, seed=8, fixed_point_flux=True, **_FIXED_KW)
    for k in range(2):
        c, _, _ = g.pass_(k)
        rc, _, _ = ref.run_pass(k, w, h)
        gi = g.info()
        assert gi["flux_overflows"] == 0 and ref.info()["flux_overflows"] == 0
        ok = rel_ok(c, rc)
        print("fixed point pass %d: %.5f of pixels within 1e-3" % (k, ok.mean()), flush=True)
        assert ok.mean() >= 0.999
    px, rp = g.pixels(), ref.pixels(w, h)
    same_n = (px["N"].ravel() == rp[:, 1]).mean()
    print("fixed point: N identical on %.5f of pixels" % same_n, flush=True)
    assert same_n >= 0.999


def test_fixed_point_flux_does_not_depend_on_gather_order(caustic_small_job):
    """Three gather forms: each lane alone in the reference's order; the 64
    lanes of a wave sharing every gather (partial sums added by a
    butterfly); the wave reading the union of its queries' cells once, each
    lane summing its own photons.  With fixed-point flux the films and every
    pixel statistic are the same bits for all three (integer sums do not
    depend on the order of the terms); in FP32 the union gather equals the
    lane-alone gather bit for bit (same photons, same order)."""
    forms = {"alone": dict(gather_share_candidates=1 << 30),
             "shared": dict(gather_share_candidates=1),
             "union": dict(gather_share_candidates=1, gather_union_factor=1.0e9)}
    outs = {}
    for fixed in (True, False):
        for name, kw in forms.items():
            sc, g = _gpu_sppm(caustic_small_job, seed=12, fixed_point_flux=fixed, **kw, **_FIXED_KW)
            films = [g.pass_(k)[0].copy() for k in range(3)]
            px = g.pixels()
            outs[(fixed, name)] = (films, px["flux"].copy(), px["radius"].copy(), px["N"].copy())
    same = lambda x, y: np.array_equal(x.view(np.uint32), y.view(np.uint32))
    for other in ("shared", "union"):
        a, b = outs[(True, "alone")], outs[(True, other)]
        assert all(same(x, y) for x, y in zip(a[0], b[0])), other
        assert all(same(x, y) for x, y in zip(a[1:], b[1:])), other
    a, b = outs[(False, "alone")], outs[(False, "union")]
    assert all(same(x, y) for x, y in zip(a[0], b[0]))
    assert all(same(x, y) for x, y in zip(a[1:], b[1:]))
    fa, fb = outs[(False, "alone")][1], outs[(False, "shared")][1]
    print("FP32 flux planes differ between the alone and shared gathers on %.4f of pixels" % (fa != fb).any(-1).mean(),
          flush=True)
    # fixed point agrees with FP32 to float rounding
    assert rel_ok(outs[(True, "alone")][0][-1], outs[(False, "alone")][0][-1]).mean() >= 0.999


def test_fixed_point_guard_counts_saturated_gathers(tmp_path):
    """A 1e12 light saturates single photon terms: the pass completes and
    counts the saturated gathers (no fault, no silent wrap)."""
    job = scenes.write_caustic_scene(str(tmp_path / "hot"), width=32, height=18, passes=1, n_major=24, n_minor=8,
                                     light_radiance=1e12)
    sc, g = _gpu_sppm(job, seed=1, fixed_point_flux=True, store_fraction=2.0, emit_fraction=100.0,
                      initial_radius=0.064)
    g.pass_(0)
    n = g.info()["flux_overflows"]
    print("saturated gathers: %d" % n, flush=True)
    assert n > 0
