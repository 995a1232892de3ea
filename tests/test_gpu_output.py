"""The output step on the GPU (apps/ViewerUtils.cpp:15-152,
PathTracingViewer.cpp:54-166):

  * hasty_estimate_device = computeEstimate (sum of the films from the first
    one with samples, / (1e-7 + total), x 2^exposure), bit for bit against a
    float32 numpy restatement, including the reference's replace-until-
    nonzero quirk;
  * hasty_present_device = preparePresentableImage (gamma 1/2.2, RGBA8);
  * film sets with asynchronous snapshots: every snapshot read back equals a
    render of the samples added so far, and hasty_films_estimate over several
    film sets equals computeEstimate of their snapshots;
  * the headless renderer with several render threads writes the EXR of
    the GPU estimate.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import hastyrenderer_amd as ha

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_estimate_matches_compute_estimate(cornell_job):
    sc = ha.Scene(cornell_job)
    sc.set_settings(width=40, height=24)
    rng = np.random.default_rng(1)
    films = [rng.random((24, 40, 3)).astype(np.float32) * 7 for _ in range(3)]
    dev = ha.DeviceFilms(40, 24)
    dev.upload(films)
    out = ha.DeviceFilms(40, 24)
    for counts, exposure in (([0, 3, 5], 0.0), ([2, 0, 4], 1.5), ([0, 0, 0], -1.0), ([1, 1, 1], 0.0)):
        ha.estimate_device(sc, dev.ptrs, counts, exposure, out.ptrs[0])
        got = out.download()[0]
        start = next((k for k, c in enumerate(counts) if c > 0), len(counts) - 1)
        want = films[start].copy()
        for k in range(start + 1, len(films)):
            want = want + films[k]
        want = want / (np.float32(1e-7) + np.float32(sum(counts)))
        if exposure != 0.0:
            want = want * np.float32(2.0 ** exposure)
        _same(got, want.astype(np.float32))


def test_present_matches_gamma_rgba8(cornell_job):
    sc = ha.Scene(cornell_job)
    sc.set_settings(width=40, height=24)
    rng = np.random.default_rng(2)
    img = (rng.random((24, 40, 3)) * 1.4 - 0.2).astype(np.float32)
    img[0, 0] = (0.0, 1.0, 0.5)
    dev = ha.DeviceFilms(40, 24)
    dev.upload([img, img, img])
    out = ha.DeviceFilms(40, 24)
    ha.present_device(sc, dev.ptrs[0], out.ptrs[0])
    rgba = out.download()[0].view(np.uint32).ravel()[: 40 * 24].reshape(24, 40)
    g = np.power(np.clip(img.astype(np.float64), 0.0, 1.0), 1.0 / 2.2)
    want = np.floor(255.0 * g).astype(np.int64)
    got = np.stack([(rgba >> 24) & 0xFF, (rgba >> 16) & 0xFF, (rgba >> 8) & 0xFF], 2).astype(np.int64)
    assert ((rgba & 0xFF) == 255).all()
    assert np.abs(got - want).max() <= 1 and (got == want).mean() >= 0.999
    assert tuple(got[0, 0]) == (0, 255, int(255 * 0.5 ** (1 / 2.2)))


def test_film_snapshots_equal_renders(cornell_job):
    sc = ha.Scene(cornell_job)
    sc.set_settings(width=64, height=48)
    f = ha.Films(sc)
    assert f.read(wait=False) is None  # nothing rendered yet
    done = 0
    for n in (2, 3, 1):
        f.render(4, done, n)
        f.snapshot()
        done += n
        snap = f.read(wait=True)
        assert snap is not None and snap[3] == done
        ref = sc.render_pt(seed=4, first_sample=0, num_samples=done)
        for x, y in zip(snap[:3], ref):
            _same(x, y)
        assert f.read(wait=True) is None  # each snapshot is handed out once
    (pc, pn, pa), k = f.device()
    assert k == done and pc and pn and pa
    # the estimate over two film sets = computeEstimate of their snapshots
    g = ha.Films(sc)
    g.render(5, 0, 2)
    g.snapshot()
    sg = g.read()
    c, n, a = ha.Films.estimate([f, g], exposure=1.0)
    f.snapshot()
    sf = f.read()
    d = np.float32(1e-7) + np.float32(done + 2)
    _same(c, ((sf[0] + sg[0]) / d) * np.float32(2.0))
    _same(a, (sf[2] + sg[2]) / d)
    f.close()
    g.close()


def test_cli_threads_write_gpu_estimate(cornell_job, tmp_path):
    folder = tmp_path / "job"
    shutil.copytree(os.path.dirname(cornell_job), folder)
    job = json.load(open(folder / "renderJob.json"))
    job["settings"].update({"num_threads": 3, "width": 48, "height": 40, "num_samples": 6})
    json.dump(job, open(folder / "renderJob.json", "w"))
    out = tmp_path / "out"
    out.mkdir()
    cli = os.path.join(os.path.dirname(ha.__file__), "hasty_render")
    r = subprocess.run([cli, str(folder / "renderJob.json"), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    img = ha.read_image(str(out / "output.exr"))
    sc = ha.Scene(str(folder / "renderJob.json"))
    # one thread takes all samples (each fetches all that remain); which one
    # depends on the scheduler, so the image is the render with its seed
    cands = [sc.render_pt(seed=t, first_sample=0, num_samples=6)[0] / (np.float32(1e-7) + np.float32(6))
             for t in range(3)]
    assert any(np.array_equal(img, c) for c in cands)
    for name in ("outputNormal.exr", "outputAlbedo.exr"):
        assert ha.read_image(str(out / name)).shape == (40, 48, 3)
