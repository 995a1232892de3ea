"""Image textures and normal maps on the GPU (Texture.h:105-152,
Scene.cpp:73-114, BRDF.cpp:604-750, LightRayInfo.cpp:14):

  * the textured scene (sRGB checker floor with roughness and metallic maps,
    normal-mapped smooth sphere, albedo-textured glass whose media absorb with
    the entry texel, a box without uv, EXR sky) against the oracle;
  * the same scene from Scene.create with the decoded textures: bit-identical;
  * 1 x 1 textures holding a material's own constants render bit-identically
    to the untextured material (cuboid scene: Lambertian, GGX, glass media);
  * a batch mixing textured and untextured jobs equals its jobs alone;
  * guided rendering runs on textured materials; photon mapping refuses them.
"""
import numpy as np
import pytest

import hastyrenderer_amd as ha
from hastyrenderer_amd import scenes
from oracle import pyoracle as po

from conftest import rel_ok

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def textured_job(tmp_path_factory):
    return scenes.write_textured_scene(str(tmp_path_factory.mktemp("textured")), width=96, height=72)


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_textured_scene_matches_oracle(textured_job):
    sc = ha.Scene(textured_job)
    c, n, a = sc.render_pt(seed=3, first_sample=0, num_samples=8)
    (rc, rn, ra), nan = po.OracleJob(textured_job, seed=3).render(0, 8, threads=16)
    frac = rel_ok(c, rc).mean()
    alb = (a.view(np.uint32) == ra.view(np.uint32)).all(axis=2).mean()
    print("textured scene: colour within 1e-3 on %.5f of pixels, albedo film bit-equal on %.5f" % (frac, alb),
          flush=True)
    assert np.isfinite(c).all() and c.mean() > 0
    assert frac >= 0.999 and alb >= 0.999
    assert np.array_equal(n.view(np.uint32), rn.view(np.uint32)) or rel_ok(n, rn).mean() >= 0.999


def test_textured_scene_create_equals_files(tmp_path):
    """A uv-mapped quad with albedo, roughness, metallic and normal maps, from
    files and from Scene.create with the textures the loader decoded."""
    rng = np.random.default_rng(8)
    folder = tmp_path / "q"
    folder.mkdir()
    scenes.write_png(str(folder / "a.png"), rng.integers(0, 256, (6, 5, 3)))
    scenes.write_pnm(str(folder / "r.pgm"), rng.integers(30, 256, (4, 4)))
    scenes.write_pnm(str(folder / "m.pgm"), rng.integers(0, 256, (3, 2)))
    scenes.write_png(str(folder / "n.png"), rng.integers(90, 256, (5, 7, 3)))
    uv = np.array([(-0.3, -0.2), (1.7, -0.2), (1.7, 1.4), (-0.3, 1.4)], np.float32)
    verts = np.array([(-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1)], np.float32)
    (folder / "m.mtl").write_text("newmtl q\nKd 0.5 0.5 0.5\nKs 0.5 0.5 0.5\nNs 200\n"
                                  "map_Kd a.png\nmap_Ns r.pgm\nmap_Pm m.pgm\nbump n.png\n"
                                  "newmtl light\nKe 20 20 20\n")
    lines = ["mtllib m.mtl", "o quad", "usemtl q"] + ["v %r %r %r" % tuple(map(float, p)) for p in verts]
    lines += ["vt %r %r" % tuple(map(float, t)) for t in uv] + ["f 1/1 2/2 3/3", "f 1/1 3/3 4/4"]
    lines += ["o light", "usemtl light", "v -0.2 1.5 -0.2", "v 0.2 1.5 -0.2", "v 0.2 1.5 0.2", "v -0.2 1.5 0.2",
              "f 5 6 7", "f 5 7 8"]
    (folder / "m.obj").write_text("\n".join(lines) + "\n")
    import json
    cam = {"position": [0.3, 1.6, 2.2], "look_at": [0, 0, 0]}
    settings = {"width": 48, "height": 40, "num_samples": 4, "max_depth": 6, "roulette_start_depth": 2}
    json.dump({"camera": cam, "world": {"background_color": [0.2, 0.3, 0.4]}, "mesh": "m.obj"},
              open(folder / "scene.json", "w"))
    json.dump({"settings": settings, "scene": "scene.json"}, open(folder / "job.json", "w"))
    job = str(folder / "job.json")
    o = po.OracleJob(job)
    mt = np.array([o.material_textures(0), o.material_textures(1)], np.int32)
    textures = [o.texture(t) for t in range(4)]
    p, fwd, right, up = scenes.look_at_frame(cam["position"], cam["look_at"])
    st = dict(scenes._SETTINGS_DEFAULTS)
    st.update(settings)
    desc = dict(
        settings={k: (float(v) if k == "roulette_q" else int(v)) for k, v in st.items()},
        camera={"position": p, "forward": fwd, "right": right, "up": up, "fov_degree": 45.0, "aperture": 0.0,
                "focal_distance": 1.0, "blade_rotation": 0.0, "num_blades": 8, "exposure": 0.0},
        vertices=np.concatenate([verts, np.array([(-0.2, 1.5, -0.2), (0.2, 1.5, -0.2), (0.2, 1.5, 0.2),
                                                  (-0.2, 1.5, 0.2)], np.float32)]),
        vertex_indices=np.array([(0, 1, 2), (0, 2, 3), (4, 5, 6), (4, 6, 7)], np.int32),
        texcoords=uv, texcoord_indices=np.array([(0, 1, 2), (0, 2, 3), (-1, -1, -1), (-1, -1, -1)], np.int32),
        geom_ids=np.array([0, 0, 1, 1], np.uint32), prim_ids=np.array([0, 1, 0, 1], np.uint32),
        material_ids=np.array([0, 0, 1, 1], np.int32),
        materials=[scenes.mtl_material(kd=(0.5, 0.5, 0.5), ks=(0.5, 0.5, 0.5), ns=200.0),
                   scenes.mtl_material(ke=(20, 20, 20), ns=1.0)],
        background=(0.2, 0.3, 0.4), textures=textures, material_textures=mt)
    a = ha.Scene(job).render_pt(seed=1, first_sample=0, num_samples=4)
    b = ha.Scene.create(**desc).render_pt(seed=1, first_sample=0, num_samples=4)
    assert a[0].mean() > 0
    _same(a, b)


def _with_constant_textures(kw):
    """The descriptor with every material's albedo, roughness and metallic
    moved into 1 x 1 textures (and one shared uv for every face)."""
    out = dict(kw)
    ntri = len(kw["vertex_indices"])
    out["texcoords"] = np.array([[0.25, 0.75]], np.float32)
    out["texcoord_indices"] = np.zeros((ntri, 3), np.int32)
    textures, mt = [], []
    for m in kw["materials"]:
        base = len(textures)
        textures += [np.array(m["albedo"], np.float32).reshape(1, 1, 3), np.full((1, 1), m["roughness"], np.float32),
                     np.full((1, 1), m["metallic"], np.float32)]
        mt.append((base, base + 1, base + 2, -1))
    out["textures"] = textures
    out["material_textures"] = np.array(mt, np.int32)
    return out


def test_constant_textures_equal_constant_materials():
    kw = scenes.cuboid_scene_desc(400, 64, 64, 4, seed=3)
    plain = dict(kw)
    plain["texcoords"] = np.array([[0.25, 0.75]], np.float32)
    plain["texcoord_indices"] = np.zeros((len(kw["vertex_indices"]), 3), np.int32)
    a = ha.Scene.create(**plain).render_pt(seed=5, first_sample=0, num_samples=4)
    b = ha.Scene.create(**_with_constant_textures(kw)).render_pt(seed=5, first_sample=0, num_samples=4)
    assert a[0].mean() > 0
    _same(a, b)


def test_textured_batch_equals_alone(textured_job, cornell_job):
    scs = [ha.Scene(textured_job), ha.Scene(cornell_job), ha.Scene(textured_job)]
    scs[1].set_settings(width=64, height=64)
    scs[2].set_settings(width=40, height=30)
    specs = [(2, 0, 3), (4, 1, 2), (6, 2, 3)]
    alone = []
    for sc, (seed, first, n) in zip(scs, specs):
        i = sc.info()
        f = ha.DeviceFilms(i["width"], i["height"])
        sc.render_pt_device(seed, first, n, *f.ptrs)
        alone.append(f.download())
        f.close()
    films = [ha.DeviceFilms(sc.info()["width"], sc.info()["height"]) for sc in scs]
    ha.render_batch([(sc, s, fi, n, f) for sc, (s, fi, n), f in zip(scs, specs, films)])
    for f, ref in zip(films, alone):
        _same(f.download(), ref)


def test_guided_textured_scene_runs(textured_job):
    sc = ha.Scene(textured_job)
    guide = ha.Guide(sc, seed=0)
    sc.set_guide(guide, fraction=0.5, train_lr=1e-3)
    c, _, _ = sc.render_pt(seed=9, first_sample=0, num_samples=8)
    loss, ns = sc.guide_stats()
    assert np.isfinite(c).all() and c.mean() > 0 and ns > 0 and np.isfinite(loss)


def test_photon_mapping_refuses_textures(textured_job):
    with pytest.raises(ha.HastyError, match="texture"):
        ha.Sppm(ha.Scene(textured_job), 1)
