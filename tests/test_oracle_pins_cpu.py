"""Pins of the oracle's environment sampling, camera and media against what
the mathematics fixes, each shown to fail on a plausible mistake.

Every check below is computed from quantities the reference defines in
closed form, never by retyping the oracle's own formula:

  * environment map (core/include/Hasty/Environment.h:47-166, Texture.h:119-140,
    VMath.h:671-700): the texel a direction reads (theta from +y, phi from +x
    toward +z, row 0 = top), the sampling density integrating to one over the
    sphere, sampled density == evaluated density, the importance-sampling
    estimator of the sky's integral against a numpy sum over the image, and
    NEE + MIS against BSDF-only sampling on an environment-lit scene;
  * camera (core/src/Camera.cpp:31-81): known-answer rays (centre pixel along
    the view direction, the image edges at half the 45-degree field of view,
    ly scaled by the WIDTH), and the thin lens (all aperture samples of a pixel
    meet in one point, the centre pixel's at the focal distance);
  * media (PathTracing.h:46-59, LightRayInfo.cpp:8-50, BRDF.cpp:540-573): a
    coloured smooth-glass slab at normal incidence under a constant sky has the
    closed-form radiance R + (1-R)^2 tau / (1 - R tau) with tau = T^d.

Each test then plants one mistake in the oracle (oracle::Mutation: a dropped
sin(theta), rows counted from the bottom, ly by the height, the up term's sign,
refocusing along the ray, Beer's law without the distance) and checks that the
same check fails.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po

M = po.mutation


def _write_min_mesh(folder, name="m"):
    """A single tiny diffuse triangle far below the camera (the loader needs a mesh)."""
    with open(os.path.join(folder, name + ".mtl"), "w") as f:
        f.write("newmtl w\nKd 0.5 0.5 0.5\nNs 0\n")
    with open(os.path.join(folder, name + ".obj"), "w") as f:
        f.write("mtllib %s.mtl\no t\nusemtl w\nv 0 -100 0\nv 0.01 -100 0\nv 0 -100 0.01\nf 1 3 2\n" % name)
    return name + ".obj"


def _env_job(folder, env, camera=None, settings=None, mesh=None):
    from hastyrenderer_amd import scenes

    os.makedirs(folder, exist_ok=True)
    scenes.write_pfm(os.path.join(folder, "sky.pfm"), env.astype(np.float32))
    mesh = mesh or _write_min_mesh(folder)
    cam = camera or {"position": [0, 0, 0]}
    json.dump({"camera": cam, "world": {"environment_texture": "sky.pfm"}, "mesh": mesh},
              open(os.path.join(folder, "scene.json"), "w"))
    st = {"width": 8, "height": 8, "num_samples": 1}
    st.update(settings or {})
    json.dump({"settings": st, "scene": "scene.json"}, open(os.path.join(folder, "job.json"), "w"))
    return os.path.join(folder, "job.json")


def _smooth_env(w, h, seed, blob=True):
    """Positive, smooth and asymmetric in every direction, plus one bright blob."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    base = 0.3 + 0.2 * np.sin(2 * np.pi * x / w + 0.7)[..., None] * np.array([1.0, 0.6, 0.3]) \
        + 0.4 * (y / h)[..., None] * np.array([0.2, 0.5, 1.0])
    env = base + 0.05 * rng.random((h, w, 3))
    if blob:
        env[h // 4, (3 * w) // 5] += np.array([40.0, 30.0, 20.0])
    return env.astype(np.float32)


def _dir(theta, phi):
    """SphericalCoordinates::ToDirection (VMath.h:681-690): theta from +y."""
    st = np.sin(theta)
    return np.stack([st * np.cos(phi), np.cos(theta), st * np.sin(phi)], -1)


# ------------------------------------------------------------ environment ----
def _lookup_ok(oj, H, W):
    rows, cols = np.mgrid[0:H, 0:W]
    theta = (rows + 0.5) / H * np.pi
    phi = (cols + 0.5) / W * 2 * np.pi
    got = oj.env_eval(_dir(theta, phi).reshape(-1, 3)).reshape(H, W, 3)
    return bool(np.array_equal(got[..., 0], rows.astype(np.float32)) and
                np.array_equal(got[..., 1], cols.astype(np.float32)))


def test_env_lookup_reads_the_texel_of_a_direction(tmp_path):
    """u = phi / 2 pi, v = 1 - theta / pi, row = (1 - v) H from the top
    (Environment.h:120-128, Texture.h:119-140): the texel centre direction of
    (row, col) reads the texel (row, col).  The image stores its own indices."""
    H, W = 16, 32
    rows, cols = np.mgrid[0:H, 0:W]
    env = np.stack([rows, cols, np.ones_like(rows)], -1).astype(np.float32)
    oj = po.OracleJob(_env_job(str(tmp_path / "lk"), env))
    assert _lookup_ok(oj, H, W)
    # +y is the top row, -y the bottom row; +x is column 0, +z a quarter turn on
    top = oj.env_eval([[0, 1, 0]])[0]
    bottom = oj.env_eval([[0, -1, 0]])[0]
    assert top[0] == 0 and bottom[0] == H - 1
    assert oj.env_eval([[1, 0, 0]])[0][1] == 0
    assert oj.env_eval([[0, 0, 1]])[0][1] == W // 4
    with M(M.ENV_ROW_FROM_BOTTOM):
        assert not _lookup_ok(oj, H, W)


def _pdf_integral(oj, H, W, sub=8):
    nt, nphi = sub * H, sub * W
    theta = (np.arange(nt) + 0.5) / nt * np.pi
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    T, P = np.meshgrid(theta, phi, indexing="ij")
    pdf = oj.env_pdf_n(_dir(T, P).reshape(-1, 3)).reshape(nt, nphi).astype(np.float64)
    return float(np.sum(pdf * np.sin(T)) * (np.pi / nt) * (2 * np.pi / nphi))


def test_env_pdf_integrates_to_one(tmp_path):
    """Midpoint quadrature of the sampling density over the sphere
    (sin(theta) d theta d phi): 1 up to float rounding."""
    H, W = 32, 64
    oj = po.OracleJob(_env_job(str(tmp_path / "pdf"), _smooth_env(W, H, 1)))
    assert _pdf_integral(oj, H, W) == pytest.approx(1.0, abs=2e-3)
    with M(M.ENV_PDF_NO_SIN_THETA):
        assert abs(_pdf_integral(oj, H, W) - 1.0) > 0.05


def _sample_vs_eval(oj, n=20000):
    s = oj.env_sample(5, n)
    assert np.allclose(np.linalg.norm(s[:, :3], axis=1), 1.0, atol=1e-5)
    rel = np.abs(s[:, 3] - s[:, 4]) / s[:, 4]
    return float(np.quantile(rel, 0.999))


def test_env_sampled_pdf_equals_evaluated_pdf(tmp_path):
    """Environment.h:141-153 vs 158-165, in the style of Tests.cpp:93-106."""
    H, W = 32, 64
    oj = po.OracleJob(_env_job(str(tmp_path / "sv"), _smooth_env(W, H, 2)))
    assert _sample_vs_eval(oj) < 1e-3
    with M(M.ENV_PDF_NO_SIN_THETA):
        assert _sample_vs_eval(oj) > 0.05


def _sky_integral(env):
    """Integral of the piecewise-constant lat-long image over the sphere."""
    H, W, _ = env.shape
    ct = np.cos(np.arange(H + 1) / H * np.pi)
    band = (ct[:-1] - ct[1:]) * (2 * np.pi / W)  # solid angle of a texel in row r
    return np.einsum("r,rwc->c", band, env.astype(np.float64))


def _importance_estimate(oj, n):
    s = oj.env_sample(9, n)
    L = oj.env_eval(s[:, :3]).astype(np.float64)
    pdf = oj.env_pdf_n(s[:, :3]).astype(np.float64)
    return np.mean(L / pdf[:, None], axis=0)


def test_env_importance_sampling_estimates_the_sky_integral(tmp_path):
    """E[L(w) / p(w)] over the sampler's directions equals the integral of the
    sky computed from the image by numpy (texel solid angles)."""
    H, W = 32, 64
    env = _smooth_env(W, H, 3)
    oj = po.OracleJob(_env_job(str(tmp_path / "is"), env))
    exact = _sky_integral(env)
    assert np.allclose(_importance_estimate(oj, 200000), exact, rtol=5e-3)
    for m in (M.ENV_ROW_FROM_BOTTOM, M.ENV_PDF_NO_SIN_THETA):
        with M(m):
            assert not np.allclose(_importance_estimate(oj, 200000), exact, rtol=5e-3)


def _env_lit_scene(folder, env):
    """A grey diffuse floor and a raised diffuse block under the sky, camera
    looking down at them; 24x24 film, depth 4, no roulette."""
    from hastyrenderer_amd import scenes

    os.makedirs(folder, exist_ok=True)
    obj = scenes._Obj()
    obj.shape("floor")
    obj.material("grey")
    vb = obj.vertices([[-3, 0, -3], [3, 0, -3], [3, 0, 3], [-3, 0, 3]])
    obj.faces([(0, 2, 1), (0, 3, 2)], vb)
    v, tris = scenes._box((0.2, 0.4, -0.3), (0.8, 0.8, 0.8), 20.0)
    obj.shape("block")
    obj.material("red")
    obj.faces(tris, obj.vertices(v))
    obj.write(os.path.join(folder, "s.obj"), "s.mtl")
    with open(os.path.join(folder, "s.mtl"), "w") as f:
        f.write(scenes._mtl("grey", kd=(0.6, 0.6, 0.6)) + scenes._mtl("red", kd=(0.7, 0.3, 0.2)))
    return _env_job(folder, env, camera={"position": [0, 3, 3], "look_at": [0, 0, 0]},
                    settings={"width": 24, "height": 24, "max_depth": 4, "roulette_q": 0.0}, mesh="s.obj")


def _mean_radiance(job, mis, spp, batches):
    oj = po.OracleJob(job, seed=17)
    oj.set(use_mis=mis)
    vals = []
    for b in range(batches):
        (c, _, _), _ = oj.render(b * spp, spp, threads=8)
        vals.append(c.mean() / spp)
    return float(np.mean(vals)), float(np.std(vals) / math.sqrt(batches))


def test_env_nee_mis_and_bsdf_only_agree(tmp_path):
    """Environment NEE + MIS (Scene.cpp:538-627, PathTracing.cpp:21-95) and
    BSDF-only sampling estimate the same image mean on an environment-lit
    scene; with the density's sin(theta) dropped, NEE is biased."""
    env = _smooth_env(64, 32, 4, blob=False)  # no blob: BSDF-only sampling stays low-variance
    job = _env_lit_scene(str(tmp_path / "lit"), env)
    m1, s1 = _mean_radiance(job, 1, 16, 8)
    m0, s0 = _mean_radiance(job, 0, 16, 8)
    assert abs(m1 - m0) < 4 * math.hypot(s1, s0) + 1e-4
    with M(M.ENV_PDF_NO_SIN_THETA):
        mb, sb = _mean_radiance(job, 1, 16, 8)
    assert abs(mb - m0) > 6 * math.hypot(sb, s0)


# ----------------------------------------------------------------- camera ----
CAM_POS = np.array([0.3, 1.2, 2.0])
CAM_LOOK = np.array([-0.5, 0.7, -1.0])


def _camera_job(folder, W, H, **cam):
    c = {"position": CAM_POS.tolist(), "look_at": CAM_LOOK.tolist()}
    c.update(cam)
    return _env_job(folder, np.ones((4, 8, 3), np.float32), camera=c, settings={"width": W, "height": H})


def _frame():
    """Camera::setLookAt (Camera.cpp:31-39): forward, up orthogonalised from +y, right = forward x up."""
    fwd = (CAM_LOOK - CAM_POS) / np.linalg.norm(CAM_LOOK - CAM_POS)
    up = np.array([0.0, 1.0, 0.0]) - fwd[1] * fwd
    up /= np.linalg.norm(up)
    return fwd, up, np.cross(fwd, up)


def _angle(a, b):
    return math.degrees(math.acos(np.clip(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b), -1, 1)))


def _pinhole_known_answers(oj, W, H):
    """Checks that fail on any one wrong term; returns the failures."""
    fwd, up, right = _frame()
    bad = []
    o, d = oj.camera_ray(W / 2, H / 2)
    if not (np.allclose(o, CAM_POS, atol=1e-6) and np.allclose(d, fwd, atol=2e-6)):
        bad.append("centre")
    for px, sgn in ((W, 1), (0, -1)):  # left / right edges at half the field of view
        _, d = oj.camera_ray(px, H / 2)
        if not (abs(_angle(d, fwd) - 22.5) < 1e-4 and np.sign(np.dot(d, right)) == sgn and abs(np.dot(d, up)) < 1e-6):
            bad.append("edge x=%d" % px)
    # ly is scaled by the WIDTH (Camera.cpp:53-56): H/2 + W/2 is also 22.5 degrees, below centre
    _, d = oj.camera_ray(W / 2, H / 2 + W / 2)
    if not (abs(_angle(d, fwd) - 22.5) < 1e-4 and np.dot(d, up) < 0 and abs(np.dot(d, right)) < 1e-6):
        bad.append("y = H/2 + W/2")
    # top-left corner: tan(angle) = q * sqrt(1 + (H/W)^2), up and left
    _, d = oj.camera_ray(0, 0)
    q = math.tan(math.radians(22.5))
    if not (abs(_angle(d, fwd) - math.degrees(math.atan(q * math.hypot(1, H / W)))) < 1e-4 and
            np.dot(d, up) > 0 and np.dot(d, right) < 0):
        bad.append("corner")
    return bad


def test_pinhole_camera_known_answer_rays(tmp_path):
    W, H = 64, 24  # W != H: ly scaled by H instead of W moves the rows
    oj = po.OracleJob(_camera_job(str(tmp_path / "cam"), W, H))
    assert _pinhole_known_answers(oj, W, H) == []
    with M(M.CAMERA_LY_BY_HEIGHT):
        assert "y = H/2 + W/2" in _pinhole_known_answers(oj, W, H)
    with M(M.CAMERA_UP_SIGN):
        assert "y = H/2 + W/2" in _pinhole_known_answers(oj, W, H)


def _lens_failures(lens, pin, W, H, aperture, focal):
    fwd, up, right = _frame()
    bad = []
    for px, py in ((W / 2, H / 2), (10.3, 7.7), (55.5, 20.25)):
        _, d0 = pin.camera_ray(px, py)
        focal_point = CAM_POS + focal * np.dot(d0, fwd) * d0
        radii = []
        for seed in range(64):
            o, d = lens.camera_ray(px, py, seed=seed)
            off = o - CAM_POS
            radii.append(np.linalg.norm(off))
            to = focal_point - o
            miss = np.linalg.norm(np.cross(to, d)) / np.linalg.norm(to)
            if miss > 2e-6 or np.dot(to, d) <= 0 or abs(np.dot(off, fwd)) > 1e-6:
                bad.append((px, py))
                break
        # origins fill the aperture polygon (6 blades: inradius cos 30 deg)
        if max(radii) > aperture * (1 + 1e-5) or max(radii) < 0.8 * aperture * math.cos(math.pi / 6):
            bad.append(("aperture", px, py))
    # the centre pixel's rays meet at the focal distance along the view direction
    o, d = lens.camera_ray(W / 2, H / 2, seed=3)
    t = np.dot(CAM_POS + focal * fwd - o, d)
    if np.linalg.norm(o + t * d - (CAM_POS + focal * fwd)) > 1e-5:
        bad.append("focal plane")
    return bad


def test_thin_lens_rays_meet_at_the_focal_point(tmp_path):
    W, H, ap, fd = 64, 24, 0.05, 2.5
    lens = po.OracleJob(_camera_job(str(tmp_path / "lens"), W, H, aperture=ap, focal_distance=fd, num_blades=6,
                                    blade_rotation=0.3))
    pin = po.OracleJob(_camera_job(str(tmp_path / "pin"), W, H))
    assert _lens_failures(lens, pin, W, H, ap, fd) == []
    with M(M.LENS_FOCAL_ALONG_RAY):
        assert _lens_failures(lens, pin, W, H, ap, fd) != []


# ------------------------------------------------------------------ media ----
SLAB_T = np.array([0.3, 0.6, 0.9])
SLAB_D = 0.5
SLAB_IOR = 1.5


def _slab_job(folder):
    """A smooth coloured glass slab (d 0, Ns 1000: transmission 1, roughness 0;
    Kd = the medium's transmittance per unit length) of thickness 0.5 in front
    of a camera with a 1-degree field of view, under a constant sky of 1,
    BSDF sampling only, no roulette."""
    from hastyrenderer_amd import scenes

    os.makedirs(folder, exist_ok=True)
    v, tris = scenes._box((0.0, 0.0, 0.0), (4.0, 4.0, SLAB_D), 0.0)
    obj = scenes._Obj()
    obj.shape("slab")
    obj.material("glass")
    obj.faces(tris, obj.vertices(v))
    obj.write(os.path.join(folder, "s.obj"), "s.mtl")
    with open(os.path.join(folder, "s.mtl"), "w") as f:
        f.write(scenes._mtl("glass", kd=tuple(SLAB_T), ns=1000.0, ni=SLAB_IOR, d=0.0))
    json.dump({"camera": {"position": [0, 0, 3], "look_at": [0, 0, 0], "fov": 1.0},
               "world": {"background_color": [1, 1, 1]}, "mesh": "s.obj"},
              open(os.path.join(folder, "scene.json"), "w"))
    json.dump({"settings": {"width": 3, "height": 3, "num_samples": 1, "use_mis": False, "max_depth": 24,
                            "roulette_q": 0.0}, "scene": "scene.json"},
              open(os.path.join(folder, "job.json"), "w"))
    return os.path.join(folder, "job.json")


def _slab_estimate(job, spp=6144):
    oj = po.OracleJob(job, seed=23)
    (c, _, _), nan = oj.render(0, spp, threads=8)
    assert nan == 0
    per_pixel = c.reshape(-1, 3) / spp
    return per_pixel.mean(0), per_pixel.std(0)


def test_glass_slab_transmission_closed_form(tmp_path):
    """Normal incidence: R = ((n-1)/(n+1))^2 at each face, tau = T^d through
    the medium; summing the internal reflections gives
    L = R + (1-R)^2 tau / (1 - R tau) (the (eta_o/eta_t)^2 radiance scalings of
    entering and leaving cancel)."""
    R = ((SLAB_IOR - 1) / (SLAB_IOR + 1)) ** 2
    tau = SLAB_T ** SLAB_D
    expect = R + (1 - R) ** 2 * tau / (1 - R * tau)
    job = _slab_job(str(tmp_path / "slab"))
    mean, _ = _slab_estimate(job)
    # each sample is a single reflect-or-refract path with values in [0, 1]:
    # 9 * 6144 samples give a standard error below 0.006
    assert np.allclose(mean, expect, atol=0.02), (mean, expect)
    with M(M.BEER_NO_DISTANCE):
        bad, _ = _slab_estimate(job)
    assert not np.allclose(bad, expect, atol=0.02)
