"""ORACLE (test infrastructure only): ctypes binding of ``liboracle.so``.

Imported by tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline
leg, always as the checker.  The renderer never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(path)
        P = C.c_void_p
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        u64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_job_load.argtypes = [C.c_char_p, C.c_uint32, C.POINTER(P)]
        L.oracle_job_free.argtypes = [P]
        L.oracle_job_set.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int]
        L.oracle_job_info.argtypes = [P, ip]
        L.oracle_render_pt.argtypes = [P, C.c_uint32, C.c_uint32, C.c_int, fp, fp, fp, u64]
        L.oracle_render_pt_counted.argtypes = [P, C.c_uint32, C.c_uint32, fp, fp, fp, u64]
        L.oracle_trace.argtypes = [P, fp, C.c_uint32, C.c_int, ip, fp, ip]
        L.oracle_tri_ids.argtypes = [P, C.c_int32, up]
        L.oracle_material.argtypes = [P, C.c_int, fp]
        L.oracle_material_textures.argtypes = [P, C.c_int, C.POINTER(C.c_int32)]
        L.oracle_texture.argtypes = [P, C.c_int, C.POINTER(C.c_uint32), C.c_void_p]
        L.oracle_light.argtypes = [P, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_float)]
        L.oracle_camera_rays.argtypes = [P, C.c_uint32, fp]
        L.oracle_camera_ray.argtypes = [P, C.c_float, C.c_float, C.c_uint32, fp]
        L.oracle_env_eval.argtypes = [P, fp, C.c_uint32, fp]
        L.oracle_env_pdf_n.argtypes = [P, fp, C.c_uint32, fp]
        L.oracle_tri_pos.argtypes = [P, C.c_uint32, fp]
        L.oracle_set_mutation.argtypes = [C.c_int]
        L.oracle_halton.argtypes = [C.c_uint32, fp]
        L.oracle_philox.argtypes = [up, up, up]
        L.oracle_sample_vndf.argtypes = [fp, C.c_uint32, fp]
        for name in ("oracle_dggx", "oracle_smith"):
            getattr(L, name).argtypes = [C.c_float] * 5
            getattr(L, name).restype = C.c_float
        L.oracle_fresnel.argtypes = [C.c_float] * 4
        L.oracle_fresnel.restype = C.c_float
        L.oracle_refract.argtypes = [fp, fp, C.c_float, C.c_float, fp]
        L.oracle_sampler.argtypes = [C.c_int, C.c_uint32, C.c_uint32, fp, fp, C.c_float, fp]
        L.oracle_env_sample.argtypes = [P, C.c_uint32, C.c_uint32, fp]
        L.oracle_env_pdf.argtypes = [P, fp]
        L.oracle_env_pdf.restype = C.c_float
        L.oracle_bsdf_eval.argtypes = [fp, fp, fp, C.c_uint32, fp]
        L.oracle_bsdf_sample.argtypes = [fp, fp, fp, C.c_uint32, C.c_uint32, fp]
        L.oracle_rotation.argtypes = [fp, fp, fp, fp]
        L.oracle_spherical.argtypes = [fp, fp]
        L.oracle_hm.argtypes = [C.c_int, C.c_float, C.c_float]
        L.oracle_hm.restype = C.c_float
        L.oracle_hm_vec.argtypes = [C.c_int, fp, fp, C.c_uint32, fp]
        L.oracle_sppm_create.argtypes = [C.c_char_p, C.c_uint32, fp, C.POINTER(P)]
        L.oracle_sppm_free.argtypes = [P]
        L.oracle_sppm_set_size.argtypes = [P, C.c_int, C.c_int]
        L.oracle_sppm_pass.argtypes = [P, C.c_uint32, fp, fp, fp]
        L.oracle_sppm_info.argtypes = [P, np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")]
        L.oracle_sppm_pixels.argtypes = [P, fp]
        L.oracle_sppm_initial_radii.argtypes = [P, fp]
        L.oracle_sppm_radii.argtypes = [P, fp]
        L.oracle_get_close.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.oracle_get_close.restype = C.c_float
        L.oracle_radius_neighbors.argtypes = [fp, C.c_uint32, C.c_float, fp, C.c_float,
                                              np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"),
                                              C.c_int64]
        L.oracle_radius_neighbors.restype = C.c_int64
        L.oracle_cell_size.argtypes = [fp, C.c_uint32, C.c_uint32]
        L.oracle_cell_size.restype = C.c_float
        L.oracle_trace_photons.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, fp, C.c_int64]
        L.oracle_trace_photons.restype = C.c_int64
        _LIB = L
    return _LIB


def f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class OracleJob:
    """A loaded renderJob.json with the oracle's BVH2 built."""

    def __init__(self, job_path: str, seed: int = 0):
        L = lib()
        h = C.c_void_p()
        rc = L.oracle_job_load(job_path.encode(), seed, C.byref(h))
        if rc != 0:
            raise RuntimeError(L.oracle_last_error().decode())
        self.h = h
        self.seed = seed

    def __del__(self):
        if getattr(self, "h", None) is not None and _LIB is not None:
            _LIB.oracle_job_free(self.h)
            self.h = None

    def set(self, width=0, height=0, max_depth=-1, rr_start=-1, rr_q=-1.0, use_mis=-1):
        lib().oracle_job_set(self.h, width, height, max_depth, rr_start, rr_q, use_mis)

    def info(self):
        a = np.zeros(8, np.int32)
        lib().oracle_job_info(self.h, a)
        keys = ("width", "height", "spp", "max_depth", "num_tris", "num_lights", "has_env", "rr_start")
        return dict(zip(keys, (int(x) for x in a)))

    def render(self, first_sample: int, n_samples: int, threads: int = 8, films=None):
        info = self.info()
        w, h = info["width"], info["height"]
        if films is None:
            films = tuple(np.zeros((h, w, 3), np.float32) for _ in range(3))
        nan = np.zeros(1, np.uint64)
        rc = lib().oracle_render_pt(self.h, first_sample, n_samples, threads, *films, nan)
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        return films, int(nan[0])

    def render_counted(self, first_sample: int, n_samples: int):
        info = self.info()
        w, h = info["width"], info["height"]
        films = tuple(np.zeros((h, w, 3), np.float32) for _ in range(3))
        out = np.zeros(3, np.uint64)
        rc = lib().oracle_render_pt_counted(self.h, first_sample, n_samples, *films, out)
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        return films, {"rays": int(out[0]), "node_visits": int(out[1]), "tri_tests": int(out[2])}

    def trace(self, rays, brute=False):
        rays = f32(rays).reshape(-1, 6)
        n = rays.shape[0]
        tri = np.zeros(n, np.int32)
        tuv = np.zeros((n, 3), np.float32)
        edge = np.zeros(n, np.int32)
        lib().oracle_trace(self.h, rays, n, 1 if brute else 0, tri, tuv, edge)
        return tri, tuv, edge

    def tri_ids(self, tri):
        out = np.zeros(2, np.uint32)
        lib().oracle_tri_ids(self.h, int(tri), out)
        return int(out[0]), int(out[1])

    def material(self, idx):
        out = np.zeros(14, np.float32)
        if lib().oracle_material(self.h, idx, out) != 0:
            raise IndexError(idx)
        keys = ("albedo", "roughness", "metallic", "specular", "ior", "anisotropy", "transmission", "emission",
                "vary_ior")
        return {"albedo": out[0:3], "roughness": out[3], "metallic": out[4], "specular": out[5], "ior": out[6],
                "anisotropy": out[7], "transmission": out[8], "emission": out[9:12], "vary_ior": bool(out[12])}

    def material_textures(self, idx):
        """(albedo, roughness, metallic, normal map) texture indices, -1 = none."""
        out = (C.c_int32 * 4)()
        if lib().oracle_material_textures(self.h, idx, out) != 0:
            raise IndexError(idx)
        return tuple(out)

    def texture(self, t):
        """Texture t as the loader decoded it: (H, W) or (H, W, 3) float32."""
        whc = (C.c_uint32 * 3)()
        if lib().oracle_texture(self.h, t, whc, None) != 0:
            raise IndexError(t)
        w, h, c = whc
        out = np.zeros((h, w, c), np.float32)
        lib().oracle_texture(self.h, t, whc, out.ctypes.data)
        return out[..., 0] if c == 1 else out

    def light(self, i):
        tri, area = C.c_int32(), C.c_float()
        if lib().oracle_light(self.h, i, C.byref(tri), C.byref(area)) != 0:
            raise IndexError(i)
        return tri.value, area.value

    def camera_rays(self, sample: int):
        info = self.info()
        out = np.zeros((info["height"], info["width"], 6), np.float32)
        lib().oracle_camera_rays(self.h, sample, out)
        return out

    def camera_ray(self, px, py, seed=0):
        """Ray through the continuous pixel position (px, py): origin, direction."""
        out = np.zeros(6, np.float32)
        lib().oracle_camera_ray(self.h, float(px), float(py), seed, out)
        return out[:3].copy(), out[3:].copy()

    def env_eval(self, dirs):
        """Environment radiance along directions (n, 3) -> (n, 3)."""
        d = f32(dirs).reshape(-1, 3)
        out = np.zeros((d.shape[0], 3), np.float32)
        lib().oracle_env_eval(self.h, d, d.shape[0], out)
        return out

    def env_pdf_n(self, dirs):
        """Environment sampling density at directions (n, 3) -> (n,)."""
        d = f32(dirs).reshape(-1, 3)
        out = np.zeros(d.shape[0], np.float32)
        lib().oracle_env_pdf_n(self.h, d, d.shape[0], out)
        return out

    def tri_pos(self, tri):
        out = np.zeros(9, np.float32)
        if lib().oracle_tri_pos(self.h, int(tri), out) != 0:
            raise IndexError(tri)
        return out.reshape(3, 3)

    def env_sample(self, seed, count):
        out = np.zeros((count, 5), np.float32)
        lib().oracle_env_sample(self.h, seed, count, out)
        return out

    def env_pdf(self, d):
        return float(lib().oracle_env_pdf(self.h, f32(d)))


class mutation:
    """Context manager planting one of the oracle's documented mistakes
    (oracle::Mutation) for the duration of a check that must then fail."""

    ENV_PDF_NO_SIN_THETA = 1
    ENV_ROW_FROM_BOTTOM = 2
    CAMERA_LY_BY_HEIGHT = 3
    CAMERA_UP_SIGN = 4
    BEER_NO_DISTANCE = 5
    LENS_FOCAL_ALONG_RAY = 6
    TEX_V_NO_FLIP = 7
    NORMAL_MAP_NO_REMAP = 8

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        lib().oracle_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().oracle_set_mutation(0)
        return False


def sampler(kind, seed, count, normal=(0, 0, 1), wo=(0, 0, 1), param=0.0):
    out = np.zeros((count, 5), np.float32)
    lib().oracle_sampler(kind, seed, count, f32(normal), f32(wo), float(param), out)
    return out


def bsdf_eval(params, n, dirs6):
    """params: albedo rgb, roughness, metallic, specular, ior, anisotropy, transmission."""
    dirs6 = f32(dirs6).reshape(-1, 6)
    out = np.zeros((dirs6.shape[0], 7), np.float32)
    if lib().oracle_bsdf_eval(f32(params), f32(n), dirs6, dirs6.shape[0], out) != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return out


def bsdf_sample(params, n, wo, seed, count):
    out = np.zeros((count, 7), np.float32)
    if lib().oracle_bsdf_sample(f32(params), f32(n), f32(wo), seed, count, out) != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return out


def sample_vndf(rows7):
    rows7 = f32(rows7).reshape(-1, 7)
    out = np.zeros((rows7.shape[0], 3), np.float32)
    lib().oracle_sample_vndf(rows7, rows7.shape[0], out)
    return out


def halton(k):
    out = np.zeros(2, np.float32)
    lib().oracle_halton(k, out)
    return out


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().oracle_philox(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def hm_vec(fn, a, b):
    a, b = f32(a), f32(b)
    out = np.zeros_like(a)
    lib().oracle_hm_vec(fn, a, b, a.size, out)
    return out


def hm(fn, a, b=0.0):
    return float(lib().oracle_hm(fn, float(a), float(b)))


class OracleSppm:
    """The photon mapper's oracle: one job, passes rendered in order."""

    def __init__(self, job_path: str, seed: int = 0, alpha=0.0, footprint_scale=0.0, store_fraction=0.0,
                 emit_fraction=0.0, width=0, height=0, initial_radius=0.0, fixed_point_flux=False,
                 flux_frac_bits=0):
        h = C.c_void_p()
        params = f32([alpha, footprint_scale, store_fraction, emit_fraction, initial_radius,
                      1.0 if fixed_point_flux else 0.0, flux_frac_bits])
        if lib().oracle_sppm_create(job_path.encode(), seed, params, C.byref(h)) != 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        self.h = h
        if width and height:
            lib().oracle_sppm_set_size(self.h, width, height)
        self.width = width
        self.height = height

    def __del__(self):
        if getattr(self, "h", None) is not None and _LIB is not None:
            _LIB.oracle_sppm_free(self.h)
            self.h = None

    def run_pass(self, k, w, h):
        films = tuple(np.zeros((h, w, 3), np.float32) for _ in range(3))
        if lib().oracle_sppm_pass(self.h, k, *films) != 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        return films

    def info(self):
        out = np.zeros(5, np.float64)
        lib().oracle_sppm_info(self.h, out)
        return {"emitted": int(out[0]), "stored": int(out[1]), "cell_size": float(out[2]), "passes": int(out[3]),
                "flux_overflows": int(out[4])}

    def pixels(self, w, h):
        out = np.zeros((h * w, 10), np.float32)
        lib().oracle_sppm_pixels(self.h, out)
        return out

    def initial_radii(self, w, h):
        out = np.zeros((h, w), np.float32)
        lib().oracle_sppm_initial_radii(self.h, out)
        return out

    def radii(self, w, h):
        out = np.zeros((h, w), np.float32)
        lib().oracle_sppm_radii(self.h, out)
        return out

    def trace_photons(self, k, first, count, max_out=1 << 22):
        out = np.zeros((max_out, 11), np.float32)
        n = lib().oracle_trace_photons(self.h, k, first, count, out, max_out)
        return out[:n]


def get_close_non_negative(img, i, j):
    img = f32(img)
    h, w = img.shape
    return float(lib().oracle_get_close(img, w, h, i, j))


def radius_neighbors(points, cell_size, center, r):
    pts = f32(points).reshape(-1, 3)
    out = np.zeros(max(1, len(pts)), np.uint32)
    n = lib().oracle_radius_neighbors(pts, len(pts), cell_size, f32(center), r, out, len(out))
    return out[:n]


def cell_size(radii):
    radii = f32(radii)
    h, w = radii.shape
    return float(lib().oracle_cell_size(radii, w, h))
