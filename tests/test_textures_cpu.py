"""Image files and image textures on the CPU.

Readers (hastyrenderer_amd.read_image -> hasty_image_read, the product's own
decoders) against an independent decoder and encoder (Pillow) on PNG (grey,
grey + alpha, RGB, RGBA, palette, 1-bit, 16-bit), PPM / PGM and PFM, and
against files written from the OpenEXR layout (scenes.write_exr: NONE, ZIPS,
ZIP; HALF and FLOAT) plus a hand-coded RLE chunk; the EXR writer round-trips.

Texture semantics in the oracle (Texture.h:84-152, Texture.cpp:9-49,
Scene.cpp:73-114, BRDF.cpp:604-628), each from what the mathematics fixes:

  * the texel a camera ray reads: the albedo film of a uv-mapped quad, texel
    by texel, against the ray / plane intersection in float64 (affine uv,
    Repeat wrap of negative and > 1 coordinates, row 0 = top, the sRGB decode
    pow(v / 255, 2.2)); fails with v not flipped;
  * normal maps: a Lambertian plane (albedo rho) under a uniform sky of
    radiance 1 whose normal map tilts the shading normal by theta: BSDF
    samples below the geometric plane return nothing, so the image mean is
    rho (1 + cos theta) / 2 (the sky-view factor of a tilted plane); fails
    without the 2 texel - 1 remap;
  * faces without uv read (1, 0, 0.7) as albedo and the loader's refusals
    (channel counts, rough glass through a roughness texture).
"""
import json
import math
import os
import struct

import numpy as np
import pytest

import hastyrenderer_amd as ha
from hastyrenderer_amd import scenes
from oracle import pyoracle as po

PIL = pytest.importorskip("PIL.Image")


# ------------------------------------------------------------------ readers --
def test_png_matches_pillow(tmp_path):
    rng = np.random.default_rng(0)
    cases = [("L", rng.integers(0, 256, (7, 9), dtype=np.uint8)),
             ("LA", rng.integers(0, 256, (3, 5, 2), dtype=np.uint8)),
             ("RGB", rng.integers(0, 256, (5, 6, 3), dtype=np.uint8)),
             ("RGBA", rng.integers(0, 256, (4, 3, 4), dtype=np.uint8))]
    # a larger smooth image: dynamic Huffman blocks, long matches, several blocks
    yy, xx = np.mgrid[0:200, 0:300]
    big = np.stack([(xx + yy) % 256, (xx * 3) % 256, (yy * 5 + rng.integers(0, 3, yy.shape)) % 256], 2)
    cases.append(("RGB", big.astype(np.uint8)))
    for mode, arr in cases:
        p = str(tmp_path / ("%s_%d.png" % (mode, arr.shape[0])))
        PIL.fromarray(arr, mode).save(p)
        got = ha.read_image(p)
        want = arr.reshape(arr.shape[0], arr.shape[1], -1).astype(np.float32) / 255.0
        assert got.shape == want.shape and np.array_equal(got, want), mode
    a16 = rng.integers(0, 65536, (6, 7)).astype(np.uint16)
    p = str(tmp_path / "g16.png")
    PIL.fromarray(a16).save(p)  # mode I;16
    assert np.array_equal(ha.read_image(p)[..., 0], a16.astype(np.float32) / 65535.0)
    pal = PIL.fromarray(rng.integers(0, 5, (4, 6), dtype=np.uint8), "P")
    pal.putpalette([int(x) for x in rng.integers(0, 256, 15)])
    p = str(tmp_path / "pal.png")
    pal.save(p)
    assert np.array_equal(ha.read_image(p), np.asarray(pal.convert("RGB")).astype(np.float32) / 255.0)
    bits = PIL.fromarray(rng.integers(0, 2, (5, 13)).astype(bool))
    p = str(tmp_path / "bits.png")
    bits.save(p)
    assert np.array_equal(ha.read_image(p)[..., 0], np.asarray(bits).astype(np.float32))


def test_own_png_writer_reads_back_in_pillow(tmp_path):
    """scenes.write_png (used by the textured test scenes) against Pillow's decoder."""
    rng = np.random.default_rng(1)
    rgb = rng.integers(0, 256, (9, 11, 3))
    scenes.write_png(str(tmp_path / "a.png"), rgb)
    assert np.array_equal(np.asarray(PIL.open(str(tmp_path / "a.png"))), rgb.astype(np.uint8))
    rgba16 = rng.integers(0, 65536, (3, 4, 4))
    scenes.write_png(str(tmp_path / "b.png"), rgba16, bits=16)
    assert np.array_equal(ha.read_image(str(tmp_path / "b.png")), rgba16.astype(np.float32) / 65535.0)


def test_pnm_and_pfm(tmp_path):
    rng = np.random.default_rng(2)
    rgb = rng.integers(0, 256, (5, 6, 3), dtype=np.uint8)
    PIL.fromarray(rgb).save(str(tmp_path / "a.ppm"))
    assert np.array_equal(ha.read_image(str(tmp_path / "a.ppm")), rgb.astype(np.float32) / 255.0)
    grey = rng.integers(0, 256, (4, 7), dtype=np.uint8)
    PIL.fromarray(grey).save(str(tmp_path / "a.pgm"))
    assert np.array_equal(ha.read_image(str(tmp_path / "a.pgm"))[..., 0], grey.astype(np.float32) / 255.0)
    # ASCII and 16-bit forms; maxval scaling as OIIO converts integers
    g16 = rng.integers(0, 1001, (3, 4))
    scenes.write_pnm(str(tmp_path / "b.pgm"), g16, maxval=1000)
    assert np.array_equal(ha.read_image(str(tmp_path / "b.pgm"))[..., 0], (g16 / np.float32(1000)).astype(np.float32))
    scenes.write_pnm(str(tmp_path / "c.ppm"), rgb, ascii=True)
    assert np.array_equal(ha.read_image(str(tmp_path / "c.ppm")), rgb.astype(np.float32) / 255.0)
    f = rng.random((4, 5)).astype(np.float32)
    PIL.fromarray(f, "F").save(str(tmp_path / "a.pfm"))
    assert np.array_equal(ha.read_image(str(tmp_path / "a.pfm"))[..., 0], f)
    img = rng.random((3, 4, 3)).astype(np.float32)
    scenes.write_pfm(str(tmp_path / "b.pfm"), img)
    assert np.array_equal(ha.read_image(str(tmp_path / "b.pfm")), img)


def test_exr_reader_and_writer(tmp_path):
    rng = np.random.default_rng(3)
    img = (rng.random((37, 21, 3)) * 5.0).astype(np.float32)
    img[0, 0] = (0.0, 65504.0, 2.0 ** -24)  # largest half, smallest half subnormal
    for comp in ("none", "zips", "zip"):
        for ptype in ("half", "float"):
            p = str(tmp_path / ("%s_%s.exr" % (comp, ptype)))
            scenes.write_exr(p, img, compression=comp, pixel_type=ptype)
            want = img.astype(np.float16).astype(np.float32) if ptype == "half" else img
            assert np.array_equal(ha.read_image(p), want), (comp, ptype)
    # the writer: uncompressed FLOAT, read back exactly
    p = str(tmp_path / "w.exr")
    ha.write_exr(p, img)
    assert np.array_equal(ha.read_image(p), img)
    with open(p, "rb") as f:
        head = f.read(8)
    assert struct.unpack("<II", head) == (20000630, 2)
    # RLE (compression 1, one line per chunk)
    p = str(tmp_path / "rle.exr")
    scenes.write_exr(p, img, compression="rle", pixel_type="half")
    assert np.array_equal(ha.read_image(p), img.astype(np.float16).astype(np.float32))
    flat = np.ones((4, 300, 3), np.float32)  # long runs: run-length codes
    scenes.write_exr(p, flat, compression="rle", pixel_type="half")
    assert os.path.getsize(p) < 2000 and np.array_equal(ha.read_image(p), flat)


def test_reader_refusals(tmp_path):
    bad = tmp_path / "x.bmp"
    bad.write_bytes(b"BM" + bytes(60))
    with pytest.raises(ha.HastyError, match="unsupported image format"):
        ha.read_image(str(bad))
    with pytest.raises(ha.HastyError, match="could not open"):
        ha.read_image(str(tmp_path / "missing.png"))
    trunc = tmp_path / "t.png"
    scenes.write_png(str(trunc), np.zeros((8, 8, 3), np.uint8))
    trunc.write_bytes(trunc.read_bytes()[:40])
    with pytest.raises(ha.HastyError):
        ha.read_image(str(trunc))


def test_reader_refuses_hostile_headers(tmp_path):
    """Sizes whose products would overflow, an EXR chunk offset near 2^64 and
    a zlib stream that inflates past the image are refused before any read
    or write out of bounds."""
    import struct
    import zlib
    big = tmp_path / "big.png"
    scenes.write_png(str(big), np.zeros((2, 2, 3), np.uint8))
    b = bytearray(big.read_bytes())
    b[16:24] = struct.pack(">II", 0xFFFFFFF0, 0xFFFFFFF0)  # IHDR width, height (CRC is not checked)
    big.write_bytes(bytes(b))
    with pytest.raises(ha.HastyError, match="out of range"):
        ha.read_image(str(big))
    (tmp_path / "big.ppm").write_bytes(b"P6 99999999999999999999 2 255\n" + bytes(12))
    with pytest.raises(ha.HastyError, match="out of range"):
        ha.read_image(str(tmp_path / "big.ppm"))
    (tmp_path / "big.pfm").write_bytes(b"PF\n3000000 3000000\n-1.0\n" + bytes(12))
    with pytest.raises(ha.HastyError, match="out of range"):
        ha.read_image(str(tmp_path / "big.pfm"))
    # EXR (2 x 3, uncompressed FLOAT RGB): the two scanline chunks of 8 + 36
    # bytes end the file and the offset table of 2 entries precedes them; its
    # first entry becomes 2^64 - 4
    ex = tmp_path / "a.exr"
    scenes.write_exr(str(ex), np.ones((2, 3, 3), np.float32), compression="none", pixel_type="float")
    e = bytearray(ex.read_bytes())
    first = len(e) - 2 * (8 + 3 * 3 * 4)
    assert struct.unpack_from("<Q", e, first - 16)[0] == first
    e[first - 16:first - 8] = struct.pack("<Q", 2 ** 64 - 4)
    ex.write_bytes(bytes(e))
    with pytest.raises(ha.HastyError, match="offset out of range"):
        ha.read_image(str(ex))
    # PNG whose image data inflates far past W x H
    row = b"\x00" + bytes(6)
    raw = zlib.compress(row * 2 + bytes(1 << 20))
    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 2, 2, 8, 2, 0, 0, 0)) + \
        chunk(b"IDAT", raw) + chunk(b"IEND", b"")
    (tmp_path / "bomb.png").write_bytes(png)
    with pytest.raises(ha.HastyError, match="longer than expected"):
        ha.read_image(str(tmp_path / "bomb.png"))


# ------------------------------------------------------------------ scenes --
QUAD = ["v -1 -1 0", "v 1 -1 0", "v 1 1 0", "v -1 1 0"]  # z = 0, facing +z


def _scene(folder, geometry, material_lines, world, camera, settings):
    """A one-material scene: `geometry` is OBJ lines after "usemtl q"."""
    os.makedirs(folder, exist_ok=True)
    with open(os.path.join(folder, "m.mtl"), "w") as f:
        f.write("newmtl q\n" + material_lines)
    with open(os.path.join(folder, "m.obj"), "w") as f:
        f.write("\n".join(["mtllib m.mtl", "o obj", "usemtl q"] + geometry) + "\n")
    json.dump({"camera": camera, "world": world, "mesh": "m.obj"}, open(os.path.join(folder, "scene.json"), "w"))
    json.dump({"settings": settings, "scene": "scene.json"}, open(os.path.join(folder, "job.json"), "w"))
    return os.path.join(folder, "job.json")


def _uv_quad(uv):
    return QUAD + ["vt %r %r" % (float(a), float(b)) for a, b in uv] + ["f 1/1 2/2 3/3", "f 1/1 3/3 4/4"]


def _texel_mismatches(tmp_path):
    """Albedo film of a uv-mapped quad (1 spp, pinhole) against the texel
    the float64 ray / plane intersection predicts; returns (checked, wrong)."""
    W, H = 5, 3  # non-square, odd
    idx = np.arange(W * H).reshape(H, W)
    tex = np.stack([20 + 15 * idx, 250 - 13 * idx, (37 * idx) % 256], 2).astype(np.uint8)
    folder = str(tmp_path / "lookup")
    os.makedirs(folder, exist_ok=True)
    scenes.write_png(os.path.join(folder, "t.png"), tex)
    au, bu, av, bv = 1.3, 0.6, 1.0, -0.2  # u = 1.3 x + 0.6 in [-0.7, 1.9]; v = y - 0.2 in [-1.2, 0.8]
    corners = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
    uv = [(au * x + bu, av * y + bv) for x, y in corners]
    job = _scene(folder, _uv_quad(uv), "Kd 1 1 1\nNs 0\nmap_Kd t.png\n", {"background_color": [0, 0, 0]},
                 {"position": [0, 0, 2.5], "look_at": [0, 0, 0]},
                 {"width": 48, "height": 32, "num_samples": 1, "max_depth": 1, "use_mis": False})
    o = po.OracleJob(job, seed=0)
    decoded = o.texture(0).reshape(-1, 3)
    want_decoded = np.power(tex.reshape(-1, 3).astype(np.float64) / 255.0, 2.2)
    assert np.allclose(decoded, want_decoded, rtol=2e-6, atol=1e-9)  # sRGB decode, Texture.cpp:43
    (_, _, alb), _ = o.render(0, 1, threads=4)
    rays = o.camera_rays(0).astype(np.float64)
    org, d = rays[..., :3], rays[..., 3:]
    t = -org[..., 2] / d[..., 2]
    x = org[..., 0] + t * d[..., 0]
    y = org[..., 1] + t * d[..., 1]
    u, v = au * x + bu, av * y + bv
    fu = u - np.floor(u)
    fv = (1.0 - v) - np.floor(1.0 - v)
    cu, rv = fu * W, fv * H
    inside = (np.abs(x) < 0.999) & (np.abs(y) < 0.999)
    safe = inside & (np.abs(cu - np.round(cu)) > 1e-3) & (np.abs(rv - np.round(rv)) > 1e-3)
    want = np.floor(rv).astype(int) * W + np.floor(cu).astype(int)
    got = np.full(alb.shape[:2], -1)
    for k in range(W * H):
        got[(alb == decoded[k]).all(axis=2)] = k
    return int(safe.sum()), int((got[safe] != want[safe]).sum())


def test_texel_lookup_known_answers(tmp_path):
    checked, wrong = _texel_mismatches(tmp_path)
    assert checked > 800 and wrong == 0


def test_texel_lookup_pin_fails_without_v_flip(tmp_path):
    with po.mutation(po.mutation.TEX_V_NO_FLIP):
        checked, wrong = _texel_mismatches(tmp_path)
    assert wrong > checked // 10


def _tilted_plane_mean(tmp_path, color):
    folder = str(tmp_path / "tilt")
    os.makedirs(folder, exist_ok=True)
    scenes.write_png(os.path.join(folder, "n.png"), np.array([[color]], np.uint8))
    plane = ["v -50 0 50", "v 50 0 50", "v 50 0 -50", "v -50 0 -50", "vt 0 0", "vt 1 0", "vt 1 1", "vt 0 1",
             "f 1/1 2/2 3/3", "f 1/1 3/3 4/4"]  # y = 0, facing +y
    job = _scene(folder, plane, "Kd 0.5 0.5 0.5\nKs 0 0 0\nNs 0\nbump -bm 1.0 n.png\n",
                 {"background_color": [1, 1, 1]}, {"position": [0, 2, 0.3], "look_at": [0, 0, 0]},
                 {"width": 64, "height": 48, "num_samples": 32, "max_depth": 2, "use_mis": False,
                  "roulette_start_depth": 10})
    o = po.OracleJob(job, seed=4)
    (c, _, _), nan = o.render(0, 32, threads=8)
    assert nan == 0
    return float(c[..., 0].mean() / 32.0)


def test_normal_map_tilt_closed_form(tmp_path):
    color = (200, 128, 230)
    tn = 2.0 * np.array(color) / 255.0 - 1.0
    cos_t = tn[2] / np.linalg.norm(tn)
    want = 0.5 * (1.0 + cos_t) / 2.0
    got = _tilted_plane_mean(tmp_path, color)
    p = (1.0 + cos_t) / 2.0
    se = 0.5 * math.sqrt(p * (1 - p) / (64 * 48 * 32))
    print("normal map: mean %.5f, closed form %.5f (theta %.1f deg, se %.1e)" % (got, want,
                                                                                math.degrees(math.acos(cos_t)), se))
    assert abs(got - want) < 5 * se + 2e-4


def test_normal_map_pin_fails_without_remap(tmp_path):
    color = (200, 128, 230)
    tn = 2.0 * np.array(color) / 255.0 - 1.0
    want = 0.5 * (1.0 + tn[2] / np.linalg.norm(tn)) / 2.0
    with po.mutation(po.mutation.NORMAL_MAP_NO_REMAP):
        got = _tilted_plane_mean(tmp_path, color)
    assert abs(got - want) > 0.01


def test_textured_scene_loader(tmp_path):
    job = scenes.write_textured_scene(str(tmp_path / "tex"))
    o = po.OracleJob(job, seed=0)
    # material order of the MTL: floor, globe, tint, nouv, light
    floor, globe, tint, nouv, light = (o.material_textures(m) for m in range(5))
    assert floor[3] == -1 and min(floor[:3]) >= 0
    assert globe[1] == globe[2] == -1 and globe[0] >= 0 and globe[3] >= 0
    assert tint == (tint[0], -1, -1, -1) and light == (-1, -1, -1, -1)
    # files used twice in one role are loaded once ("-bm 1.0 normal.png" names normal.png)
    assert nouv == (floor[0], floor[1], floor[2], globe[3])
    folder = str(tmp_path / "tex")
    checker = ha.read_image(os.path.join(folder, "checker.png"))
    assert np.allclose(o.texture(floor[0]), np.power(checker.astype(np.float64), 2.2), rtol=2e-6, atol=1e-9)
    assert np.array_equal(o.texture(floor[1]), ha.read_image(os.path.join(folder, "rough.pgm"))[..., 0])
    assert np.array_equal(o.texture(floor[2]), ha.read_image(os.path.join(folder, "metal map.pgm"))[..., 0])
    assert np.array_equal(o.texture(globe[3]), ha.read_image(os.path.join(folder, "normal.png")))
    assert o.texture(globe[0]).shape == (8, 16, 3)  # RGBA: alpha dropped
    # the box without uv reads the NoUV albedo (1, 0, 0.7) (Texture.h:98-101)
    (_, _, alb), _ = o.render(0, 1, threads=8)
    magenta = (alb == np.array([1.0, 0.0, 0.7], np.float32)).all(axis=2)
    assert magenta.sum() > 20


def test_texture_refusals(tmp_path):
    folder = str(tmp_path / "ref")
    os.makedirs(folder, exist_ok=True)
    scenes.write_png(os.path.join(folder, "g.png"), np.full((2, 2), 128, np.uint8))
    scenes.write_png(os.path.join(folder, "c.png"), np.full((2, 2, 3), 128, np.uint8))
    cam, world = {"position": [0, 0, 3]}, {"background_color": [1, 1, 1]}
    st = {"width": 8, "height": 8, "num_samples": 1}
    uv = [(0, 0), (1, 0), (1, 1), (0, 1)]
    job = _scene(folder, _uv_quad(uv), "Kd 1 1 1\nmap_Kd g.png\n", world, cam, st)
    with pytest.raises(RuntimeError, match="expected 3 or 4, found 1"):
        po.OracleJob(job)
    job = _scene(folder, _uv_quad(uv), "Kd 1 1 1\nmap_Ns c.png\n", world, cam, st)
    with pytest.raises(RuntimeError, match="expected 1, found 3"):
        po.OracleJob(job)
    # smooth glass with a roughness texture whose texels are not all 0
    job = _scene(folder, _uv_quad(uv), "Kd 1 1 1\nd 0\nNi 1.5\nNs 1000\nmap_Ns g.png\n", world, cam, st)
    with pytest.raises(RuntimeError, match="not implemented"):
        po.OracleJob(job)
