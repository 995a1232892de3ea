"""hastyrenderer_amd: MI355X path tracer behind HastyRenderer's render entry points.

The compute path is ``libhasty_amd.so`` (hand-written HIP for gfx950, C ABI in
``include/hasty_amd.h``).  This module binds it with ctypes and mirrors the
reference's host-side render interface:

* :class:`RenderJob` -- ``RenderJob::loadJSON`` / ``fetchSample`` / ``stopFlag``
  (core/include/Hasty/RenderJob.h:42-64, core/src/RenderJob.cpp:21-28);
* :class:`AccumulationBuffer` / :class:`Image3fAccDoubleBuffer` -- the film
  sums and their double buffer (core/include/Hasty/Image.h:18-45, 302-316);
* :func:`renderThread` -- the pass loop of ``renderThread``
  (core/src/PathTracing.cpp:98-160), issuing GPU batches of sample indices;
* :func:`computeEstimate` -- ``computeEstimate`` (apps/ViewerUtils.cpp:95-121).

There is no CPU fallback: if the HIP library is missing every entry point
raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhasty_amd.so")
_lib = None

# Every symbol declared in include/hasty_amd.h.
EXPORTS = (
    "hasty_scene_load",
    "hasty_scene_create",
    "hasty_scene_destroy",
    "hasty_scene_info",
    "hasty_scene_set_settings",
    "hasty_render_pt",
    "hasty_render_pt_device",
    "hasty_trace_closest",
    "hasty_trace_hits",
    "hasty_trace_hits_device",
    "hasty_trace_hits_device_noalloc",
    "hasty_image_read",
    "hasty_image_write_exr",
    "hasty_estimate_device",
    "hasty_present_device",
    "hasty_films_create",
    "hasty_films_destroy",
    "hasty_films_render",
    "hasty_films_snapshot",
    "hasty_films_read",
    "hasty_films_device",
    "hasty_films_estimate",
    "hasty_set_profiling",
    "hasty_kernel_stats",
    "hasty_reset_stats",
    "hasty_counters",
    "hasty_counters_ex",
    "hasty_set_launch",
    "hasty_get_launch",
    "hasty_set_launch_ex",
    "hasty_render_batch_device",
    "hasty_autotune_batch",
    "hasty_autotune",
    "hasty_guide_num_params",
    "hasty_guide_create",
    "hasty_guide_destroy",
    "hasty_guide_set_box",
    "hasty_guide_set_params",
    "hasty_guide_get_params",
    "hasty_guide_infer",
    "hasty_guide_gradient",
    "hasty_guide_train_step",
    "hasty_guide_set_selection",
    "hasty_guide_gradient_ex",
    "hasty_scene_set_guide",
    "hasty_scene_set_guide_training",
    "hasty_scene_set_guide_photons",
    "hasty_scene_set_guide_depth",
    "hasty_scene_guide_stats",
    "hasty_sppm_create",
    "hasty_sppm_destroy",
    "hasty_sppm_pass_device",
    "hasty_sppm_pass",
    "hasty_sppm_info",
    "hasty_sppm_pixels",
    "hasty_sppm_radii",
    "hasty_sppm_photons",
    "hasty_last_error",
)


class HastyError(RuntimeError):
    pass


# HastyHit (include/hasty_amd.h) as a numpy record.
HIT_DTYPE = np.dtype([("geom_id", "<u4"), ("prim_id", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4"),
                      ("ng", "<f4", (3,))])


class RenderSettingsC(C.Structure):
    _fields_ = [
        ("roulette_start_depth", C.c_int32),
        ("roulette_q", C.c_float),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("num_threads", C.c_uint32),
        ("max_depth", C.c_int32),
        ("num_samples", C.c_int32),
        ("use_mis", C.c_int32),
    ]


class MaterialDescC(C.Structure):
    _fields_ = [
        ("albedo", C.c_float * 3),
        ("roughness", C.c_float),
        ("metallic", C.c_float),
        ("specular", C.c_float),
        ("ior", C.c_float),
        ("anisotropy", C.c_float),
        ("transmission", C.c_float),
        ("vary_ior", C.c_int32),
        ("emission", C.c_float * 3),
        ("tinted_emitter", C.c_int32),
    ]


class CameraDescC(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("forward", C.c_float * 3),
        ("right", C.c_float * 3),
        ("up", C.c_float * 3),
        ("fov_degree", C.c_float),
        ("aperture", C.c_float),
        ("focal_distance", C.c_float),
        ("blade_rotation", C.c_float),
        ("num_blades", C.c_int32),
        ("exposure", C.c_float),
    ]


class TextureDescC(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("channels", C.c_uint32),
        ("pixels", C.c_void_p),
    ]


class SceneDescC(C.Structure):
    _fields_ = [
        ("settings", RenderSettingsC),
        ("camera", CameraDescC),
        ("num_vertices", C.c_uint32),
        ("vertices", C.c_void_p),
        ("num_normals", C.c_uint32),
        ("normals", C.c_void_p),
        ("num_texcoords", C.c_uint32),
        ("texcoords", C.c_void_p),
        ("num_triangles", C.c_uint32),
        ("vertex_indices", C.c_void_p),
        ("normal_indices", C.c_void_p),
        ("texcoord_indices", C.c_void_p),
        ("geom_ids", C.c_void_p),
        ("prim_ids", C.c_void_p),
        ("material_ids", C.c_void_p),
        ("smooth", C.c_void_p),
        ("num_materials", C.c_uint32),
        ("materials", C.c_void_p),
        ("has_environment", C.c_int32),
        ("env_width", C.c_uint32),
        ("env_height", C.c_uint32),
        ("env_rgb", C.c_void_p),
        ("background", C.c_float * 3),
        ("num_textures", C.c_uint32),
        ("textures", C.c_void_p),
        ("material_textures", C.c_void_p),
    ]


class SceneInfoC(C.Structure):
    _fields_ = [
        ("settings", RenderSettingsC),
        ("num_triangles", C.c_uint32),
        ("num_shapes", C.c_uint32),
        ("num_materials", C.c_uint32),
        ("num_light_triangles", C.c_uint32),
        ("has_environment", C.c_int32),
        ("bvh_nodes", C.c_uint32),
        ("bvh_depth", C.c_int32),
        ("device_bytes", C.c_uint64),
        ("bvh_width", C.c_int32),
        ("camera_exposure", C.c_float),
    ]


class LaunchC(C.Structure):
    _fields_ = [
        ("path_slots", C.c_uint32),
        ("trace_blocks", C.c_uint32),
        ("shade_blocks", C.c_uint32),
        ("chunk", C.c_uint32),
        ("use_graph", C.c_int32),
        ("bvh_width", C.c_uint32),
    ]


class BatchJobC(C.Structure):
    _fields_ = [
        ("scene", C.c_void_p),
        ("seed", C.c_uint32),
        ("first_sample", C.c_uint32),
        ("num_samples", C.c_uint32),
        ("d_color", C.c_void_p),
        ("d_normal", C.c_void_p),
        ("d_albedo", C.c_void_p),
    ]


class SppmParamsC(C.Structure):
    _fields_ = [
        ("alpha", C.c_float),
        ("footprint_scale", C.c_float),
        ("store_fraction", C.c_double),
        ("emit_fraction", C.c_double),
        ("initial_radius", C.c_float),
        ("fixed_point_flux", C.c_int32),
        ("flux_frac_bits", C.c_int32),
        ("gather_share_candidates", C.c_uint32),
        ("gather_union_factor", C.c_float),
    ]


class SppmInfoC(C.Structure):
    _fields_ = [
        ("emitted", C.c_uint64),
        ("stored", C.c_uint64),
        ("cell_size", C.c_float),
        ("passes_done", C.c_uint32),
        ("emit_limit", C.c_uint32),
        ("store_limit", C.c_uint32),
        ("flux_overflows", C.c_uint32),
        ("gather_queries", C.c_uint64),
        ("gather_candidates", C.c_uint64),
        ("gather_found", C.c_uint64),
    ]


class KernelStatC(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double)]


def lib():
    """Loads libhasty_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HastyError("libhasty_amd.so is not built (run __graft_entry__.build() or make -C hastyrenderer_amd)")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    fp = C.POINTER(C.c_float)
    L.hasty_scene_load.argtypes = [C.c_char_p, C.POINTER(P)]
    L.hasty_scene_create.argtypes = [C.POINTER(SceneDescC), C.POINTER(P)]
    L.hasty_scene_destroy.argtypes = [P]
    L.hasty_scene_destroy.restype = None
    L.hasty_scene_info.argtypes = [P, C.POINTER(SceneInfoC)]
    L.hasty_scene_set_settings.argtypes = [P, C.POINTER(RenderSettingsC)]
    L.hasty_render_pt.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, fp, fp, fp]
    L.hasty_render_pt_device.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, P, P, P]
    L.hasty_trace_closest.argtypes = [P, fp, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_int32), fp]
    L.hasty_trace_hits.argtypes = [P, C.c_void_p, C.c_uint32, C.c_void_p]
    L.hasty_trace_hits_device.argtypes = [P, C.c_void_p, C.c_uint32, C.c_void_p]
    L.hasty_trace_hits_device_noalloc.argtypes = [P, C.c_void_p, C.c_uint32, C.c_void_p]
    u32p = C.POINTER(C.c_uint32)
    L.hasty_image_read.argtypes = [C.c_char_p, u32p, u32p, u32p, fp, C.c_uint64]
    L.hasty_image_write_exr.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, fp]
    L.hasty_estimate_device.argtypes = [P, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), C.c_uint32, C.c_float,
                                        C.c_void_p]
    L.hasty_present_device.argtypes = [P, C.c_void_p, C.c_void_p]
    L.hasty_films_create.argtypes = [P, C.POINTER(P)]
    L.hasty_films_destroy.argtypes = [P]
    L.hasty_films_destroy.restype = None
    L.hasty_films_render.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32]
    L.hasty_films_snapshot.argtypes = [P]
    L.hasty_films_read.argtypes = [P, C.c_int, fp, fp, fp, C.POINTER(C.c_uint32)]
    L.hasty_films_device.argtypes = [P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_uint32)]
    L.hasty_films_estimate.argtypes = [C.POINTER(P), C.c_uint32, C.c_float, fp, fp, fp]
    L.hasty_set_profiling.argtypes = [P, C.c_int]
    L.hasty_kernel_stats.argtypes = [P, C.POINTER(KernelStatC), C.c_int]
    L.hasty_reset_stats.argtypes = [P]
    L.hasty_counters.argtypes = [P, C.POINTER(C.c_uint64)]
    L.hasty_counters_ex.argtypes = [P, C.POINTER(C.c_uint64), C.c_uint32]
    L.hasty_set_launch.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32]
    L.hasty_get_launch.argtypes = [P, C.POINTER(LaunchC)]
    L.hasty_set_launch_ex.argtypes = [P, C.POINTER(LaunchC)]
    L.hasty_render_batch_device.argtypes = [C.POINTER(BatchJobC), C.c_uint32]
    L.hasty_autotune_batch.argtypes = [C.POINTER(BatchJobC), C.c_uint32, C.c_uint32, C.POINTER(LaunchC)]
    L.hasty_autotune.argtypes = [P, C.c_uint32, C.c_uint32, C.POINTER(LaunchC)]
    L.hasty_guide_num_params.restype = C.c_uint32
    L.hasty_guide_create.argtypes = [P, C.c_uint32, C.POINTER(P)]
    L.hasty_guide_destroy.argtypes = [P]
    L.hasty_guide_destroy.restype = None
    L.hasty_guide_set_box.argtypes = [P, fp, fp]
    L.hasty_guide_set_params.argtypes = [P, fp]
    L.hasty_guide_get_params.argtypes = [P, fp]
    L.hasty_guide_infer.argtypes = [P, fp, C.c_uint32, fp]
    L.hasty_guide_gradient.argtypes = [P, fp, C.c_uint32, fp, fp]
    L.hasty_guide_train_step.argtypes = [P, fp, C.c_uint32, C.c_float, fp]
    L.hasty_guide_set_selection.argtypes = [P, C.c_float, C.c_float]
    L.hasty_guide_gradient_ex.argtypes = [P, fp, C.c_uint32, C.c_uint32, fp, fp]
    L.hasty_scene_set_guide.argtypes = [P, P, C.c_float, C.c_float]
    L.hasty_scene_set_guide_training.argtypes = [P, C.c_uint32, C.c_float]
    L.hasty_scene_set_guide_photons.argtypes = [P, C.c_uint32, C.c_float]
    L.hasty_scene_set_guide_depth.argtypes = [P, C.c_uint32, C.c_float]
    L.hasty_scene_guide_stats.argtypes = [P, fp, C.POINTER(C.c_uint32)]
    L.hasty_sppm_create.argtypes = [P, C.c_uint32, C.POINTER(SppmParamsC), C.POINTER(P)]
    L.hasty_sppm_destroy.argtypes = [P]
    L.hasty_sppm_destroy.restype = None
    L.hasty_sppm_pass_device.argtypes = [P, C.c_uint32, P, P, P]
    L.hasty_sppm_pass.argtypes = [P, C.c_uint32, fp, fp, fp]
    L.hasty_sppm_info.argtypes = [P, C.POINTER(SppmInfoC)]
    L.hasty_sppm_pixels.argtypes = [P, fp]
    L.hasty_sppm_radii.argtypes = [P, fp]
    L.hasty_sppm_photons.argtypes = [P, fp, C.c_uint64]
    L.hasty_last_error.restype = C.c_char_p
    _lib = L
    return L


def _err(rc):
    if rc != 0:
        raise HastyError(lib().hasty_last_error().decode())


def _fptr(a: np.ndarray):
    """float* of a C-contiguous float32 array (ValueError otherwise: the C side
    reads and writes through it without knowing its size)."""
    if not isinstance(a, np.ndarray) or a.dtype != np.float32 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("expected a C-contiguous float32 numpy array")
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _check_films(films, h, w):
    if len(films) != 3:
        raise ValueError("expected three films (color, normal, albedo)")
    for f in films:
        if not isinstance(f, np.ndarray) or f.shape != (h, w, 3):
            raise ValueError("film shape %s, expected %s" % (getattr(f, "shape", None), (h, w, 3)))


class Scene:
    """A loaded render job on the GPU (HastyScene*)."""

    def __init__(self, render_job_path: str):
        h = C.c_void_p()
        _err(lib().hasty_scene_load(os.fsencode(render_job_path), C.byref(h)))
        self._h = h

    @classmethod
    def create(cls, settings: dict, camera: dict, vertices, vertex_indices, geom_ids, prim_ids, material_ids,
               materials, normals=None, normal_indices=None, texcoords=None, texcoord_indices=None, smooth=None,
               env_rgb=None, background=(0.0, 0.0, 0.0), textures=None, material_textures=None):
        """hasty_scene_create from arrays (what Scene::Scene holds after
        tinyobj / OIIO): settings as RenderSettingsC fields, camera as
        CameraDescC fields, materials as a list of MaterialDescC-field dicts,
        per-triangle arrays of 3 indices (or one id) per triangle, env_rgb an
        (H, W, 3) image with row 0 at the top or None.  textures: linear
        float images, (H, W) for roughness / metallic or (H, W, 3) for albedo
        / normal maps; material_textures: (num_materials, 4) indices (albedo,
        roughness, metallic, normal map; -1 none)."""
        keep = []

        def arr(a, dtype, cols):
            if a is None:
                return None, 0
            a = np.ascontiguousarray(a, dtype=dtype)
            if a.size % cols:
                raise ValueError("array size %d is not a multiple of %d" % (a.size, cols))
            keep.append(a)
            return a.ctypes.data, a.size // cols

        d = SceneDescC()
        for k, v in settings.items():
            setattr(d.settings, k, v)
        for k, v in camera.items():
            if k in ("position", "forward", "right", "up"):
                setattr(d.camera, k, (C.c_float * 3)(*v))
            else:
                setattr(d.camera, k, v)
        d.vertices, d.num_vertices = arr(vertices, np.float32, 3)
        d.normals, d.num_normals = arr(normals, np.float32, 3)
        d.texcoords, d.num_texcoords = arr(texcoords, np.float32, 2)
        d.vertex_indices, d.num_triangles = arr(vertex_indices, np.int32, 3)
        d.normal_indices, _ = arr(normal_indices, np.int32, 3)
        d.texcoord_indices, _ = arr(texcoord_indices, np.int32, 3)
        for name, a, dt in (("geom_ids", geom_ids, np.uint32), ("prim_ids", prim_ids, np.uint32),
                            ("material_ids", material_ids, np.int32), ("smooth", smooth, np.uint8)):
            ptr, n = arr(a, dt, 1)
            if ptr is not None and n != d.num_triangles:
                raise ValueError("%s has %d entries for %d triangles" % (name, n, d.num_triangles))
            setattr(d, name, ptr)
        mats = (MaterialDescC * max(1, len(materials)))()
        for i, m in enumerate(materials):
            for k, v in m.items():
                setattr(mats[i], k, (C.c_float * 3)(*v) if k in ("albedo", "emission") else v)
        d.num_materials = len(materials)
        d.materials = C.cast(mats, C.c_void_p)
        if env_rgb is not None:
            env = np.ascontiguousarray(env_rgb, np.float32)
            keep.append(env)
            d.has_environment = 1
            d.env_height, d.env_width = env.shape[0], env.shape[1]
            d.env_rgb = env.ctypes.data
        d.background = (C.c_float * 3)(*background)
        if textures:
            tds = (TextureDescC * len(textures))()
            for i, t in enumerate(textures):
                t = np.ascontiguousarray(t, np.float32)
                if t.ndim not in (2, 3) or (t.ndim == 3 and t.shape[2] != 3):
                    raise ValueError("texture %d must be (H, W) or (H, W, 3)" % i)
                keep.append(t)
                tds[i].height, tds[i].width = t.shape[0], t.shape[1]
                tds[i].channels = 1 if t.ndim == 2 else 3
                tds[i].pixels = t.ctypes.data
            keep.append(tds)
            d.num_textures = len(textures)
            d.textures = C.cast(tds, C.c_void_p)
        if material_textures is not None:
            mt = np.ascontiguousarray(material_textures, np.int32)
            if mt.shape != (len(materials), 4):
                raise ValueError("material_textures must be (num_materials, 4)")
            keep.append(mt)
            d.material_textures = mt.ctypes.data
        h = C.c_void_p()
        _err(lib().hasty_scene_create(C.byref(d), C.byref(h)))
        obj = cls.__new__(cls)
        obj._h = h
        return obj

    def close(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.hasty_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        i = SceneInfoC()
        _err(lib().hasty_scene_info(self._h, C.byref(i)))
        s = i.settings
        return {
            "roulette_start_depth": s.roulette_start_depth,
            "roulette_q": s.roulette_q,
            "width": s.width,
            "height": s.height,
            "num_threads": s.num_threads,
            "max_depth": s.max_depth,
            "num_samples": s.num_samples,
            "use_mis": bool(s.use_mis),
            "num_triangles": i.num_triangles,
            "num_shapes": i.num_shapes,
            "num_materials": i.num_materials,
            "num_light_triangles": i.num_light_triangles,
            "has_environment": bool(i.has_environment),
            "bvh_nodes": i.bvh_nodes,
            "bvh_depth": i.bvh_depth,
            "bvh_width": i.bvh_width,
            "camera_exposure": i.camera_exposure,
            "device_bytes": i.device_bytes,
        }

    def set_settings(self, **kw):
        i = SceneInfoC()
        _err(lib().hasty_scene_info(self._h, C.byref(i)))
        s = i.settings
        for k, v in kw.items():
            setattr(s, k, int(v) if k != "roulette_q" else float(v))
        _err(lib().hasty_scene_set_settings(self._h, C.byref(s)))

    def set_launch(self, path_slots=0, trace_blocks=0, shade_blocks=0):
        """Path slots in flight and grid sizes (0 keeps the current value)."""
        _err(lib().hasty_set_launch(self._h, path_slots, trace_blocks, shade_blocks))

    def render_pt(self, seed: int, first_sample: int, num_samples: int, films=None):
        """Adds samples [first, first+n) into host films; returns (color, normal, albedo) sums."""
        info = self.info()
        h, w = info["height"], info["width"]
        if films is None:
            films = tuple(np.zeros((h, w, 3), np.float32) for _ in range(3))
        _check_films(films, h, w)
        _err(lib().hasty_render_pt(self._h, seed, first_sample, num_samples, *(_fptr(f) for f in films)))
        return films

    def render_pt_device(self, seed: int, first_sample: int, num_samples: int, d_color, d_normal, d_albedo):
        """Films are device pointers (ints), e.g. torch tensors' data_ptr()."""
        _err(lib().hasty_render_pt_device(self._h, seed, first_sample, num_samples, C.c_void_p(d_color),
                                          C.c_void_p(d_normal), C.c_void_p(d_albedo)))

    def set_guide(self, guide, fraction=0.5, train_lr=0.0):
        """Guided rendering with `guide` (None: plain); with train_lr > 0,
        online Adam training (one pass of minibatch steps) after every group
        of sample indices of about 2^20 paths."""
        self._guide = guide
        _err(lib().hasty_scene_set_guide(self._h, guide.handle if guide is not None else None, fraction,
                                         train_lr))

    def set_guide_training(self, minibatch=16384, weight_cap=20.0):
        _err(lib().hasty_scene_set_guide_training(self._h, minibatch, weight_cap))

    def set_guide_photons(self, photons=0, share=0.5):
        """Caustic light paths per training step as extra guide training
        samples (0 turns them off), weighted share x the camera samples."""
        _err(lib().hasty_scene_set_guide_photons(self._h, photons, float(share)))

    def set_guide_depth(self, max_depth=255, max_throughput=3.0e38):
        """Guide only vertices of depth below max_depth while the path
        throughput (every channel) is at most max_throughput."""
        _err(lib().hasty_scene_set_guide_depth(self._h, int(max_depth), float(max_throughput)))

    def guide_stats(self):
        loss = np.zeros(1, np.float32)
        n = C.c_uint32()
        _err(lib().hasty_scene_guide_stats(self._h, _fptr(loss), C.byref(n)))
        return float(loss[0]), int(n.value)

    def get_launch(self) -> dict:
        lc = LaunchC()
        _err(lib().hasty_get_launch(self._h, C.byref(lc)))
        return {k: int(getattr(lc, k)) for k, _ in LaunchC._fields_}

    def set_launch_ex(self, **kw):
        """path_slots, trace_blocks, shade_blocks, chunk, use_graph, bvh_width."""
        cur = self.get_launch()
        cur.update(kw)
        lc = LaunchC(*(int(cur[k]) for k, _ in LaunchC._fields_))
        _err(lib().hasty_set_launch_ex(self._h, C.byref(lc)))

    def autotune(self, seed: int = 0, trial_samples: int = 1) -> dict:
        """Times short renders over the launch knobs and keeps the fastest."""
        lc = LaunchC()
        _err(lib().hasty_autotune(self._h, seed, trial_samples, C.byref(lc)))
        return {k: int(getattr(lc, k)) for k, _ in LaunchC._fields_}

    def trace_closest(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = rays.shape[0]
        gp = np.zeros((n, 2), np.uint32)
        tri = np.zeros(n, np.int32)
        tuv = np.zeros((n, 3), np.float32)
        _err(lib().hasty_trace_closest(self._h, _fptr(rays), n, gp.ctypes.data_as(C.POINTER(C.c_uint32)),
                                       tri.ctypes.data_as(C.POINTER(C.c_int32)), _fptr(tuv)))
        return gp, tri, tuv

    def trace_hits(self, rays: np.ndarray) -> np.ndarray:
        """Full hit records (HastyHit) of rays (n, 6): a structured array with
        geom_id, prim_id, t, u, v, ng (3)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        hits = np.zeros(rays.shape[0], HIT_DTYPE)
        _err(lib().hasty_trace_hits(self._h, rays.ctypes.data, rays.shape[0], hits.ctypes.data))
        return hits

    def trace_hits_device(self, d_rays: int, n: int, d_hits: int):
        """Rays (n x HastyRay) and hits (n x HastyHit) in device memory."""
        _err(lib().hasty_trace_hits_device(self._h, C.c_void_p(d_rays), n, C.c_void_p(d_hits)))

    def trace_hits_device_noalloc(self, d_rays: int, n: int, d_hits: int):
        """trace_hits_device without any device allocation (the library's
        static scratch array holds the traversal stack overflow)."""
        _err(lib().hasty_trace_hits_device_noalloc(self._h, C.c_void_p(d_rays), n, C.c_void_p(d_hits)))

    def set_profiling(self, events=True, counters=False, extend_only=False):
        """HIP-event timing of every launch (of the extension launches only
        with extend_only) and/or device counters."""
        _err(lib().hasty_set_profiling(self._h, (1 if events else 0) | (2 if counters else 0) |
                                       (4 if extend_only else 0)))

    def reset_stats(self):
        _err(lib().hasty_reset_stats(self._h))

    def kernel_stats(self) -> dict:
        arr = (KernelStatC * 16)()
        n = lib().hasty_kernel_stats(self._h, arr, 16)
        if n < 0:
            raise HastyError(lib().hasty_last_error().decode())
        return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms)) for i in range(n)}

    def counters(self) -> dict:
        out = (C.c_uint64 * 7)()
        _err(lib().hasty_counters_ex(self._h, out, 7))
        keys = ("extension_rays", "shadow_rays", "node_visits", "triangle_tests", "media_overflow",
                "shadow_node_visits", "shadow_triangle_tests")
        return dict(zip(keys, (int(x) for x in out)))


class DeviceFilms:
    """Color, normal and albedo films in device memory (hipMalloc), zeroed."""

    _hip = None

    def __init__(self, width: int, height: int):
        if DeviceFilms._hip is None:
            h = C.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            DeviceFilms._hip = h
        self.width, self.height = width, height
        self.nbytes = width * height * 3 * 4
        self.ptrs = []
        for _ in range(3):
            p = C.c_void_p()
            if DeviceFilms._hip.hipMalloc(C.byref(p), self.nbytes) != 0:
                raise HastyError("hipMalloc failed")
            self.ptrs.append(p.value)
        self.zero()

    def zero(self):
        for p in self.ptrs:
            if DeviceFilms._hip.hipMemset(C.c_void_p(p), 0, self.nbytes) != 0:
                raise HastyError("hipMemset failed")

    def upload(self, arrays):
        """Copies three host (H, W, 3) float32 arrays into the films."""
        for p, a in zip(self.ptrs, arrays):
            a = np.ascontiguousarray(a, np.float32)
            if a.nbytes != self.nbytes:
                raise ValueError("film of %d bytes expected" % self.nbytes)
            if DeviceFilms._hip.hipMemcpy(C.c_void_p(p), a.ctypes.data, self.nbytes, 1) != 0:  # host to device
                raise HastyError("hipMemcpy failed")

    def download(self):
        out = []
        for p in self.ptrs:
            a = np.zeros((self.height, self.width, 3), np.float32)
            if DeviceFilms._hip.hipMemcpy(a.ctypes.data, C.c_void_p(p), self.nbytes, 2) != 0:  # device to host
                raise HastyError("hipMemcpy failed")
            out.append(a)
        return tuple(out)

    def close(self):
        for p in self.ptrs:
            DeviceFilms._hip.hipFree(C.c_void_p(p))
        self.ptrs = []

    def __del__(self):
        if self.ptrs and DeviceFilms._hip is not None:
            self.close()


def _batch_array(jobs):
    arr = (BatchJobC * len(jobs))()
    for i, (scene, seed, first, n, films) in enumerate(jobs):
        d = films.ptrs if films is not None else (None, None, None)
        arr[i] = BatchJobC(scene.handle, seed, first, n, *d)
    return arr


def render_batch(jobs):
    """jobs: list of (Scene, seed, first_sample, num_samples, DeviceFilms).
    Renders all of them through the first scene's path pool in one batch."""
    arr = _batch_array(jobs)
    _err(lib().hasty_render_batch_device(arr, len(jobs)))


def autotune_batch(jobs, trial_samples: int = 1) -> dict:
    """jobs: list of (Scene, seed, first_sample, num_samples, films-or-None).
    Applies the fastest launch configuration to jobs[0]'s scene."""
    arr = _batch_array(jobs)
    lc = LaunchC()
    _err(lib().hasty_autotune_batch(arr, len(jobs), trial_samples, C.byref(lc)))
    return {k: int(getattr(lc, k)) for k, _ in LaunchC._fields_}


class Guide:
    """The learned path guide of a scene (HastyGuide*): MLP -> vMF mixture."""

    def __init__(self, scene: Scene, seed: int = 0):
        self.scene = scene
        h = C.c_void_p()
        _err(lib().hasty_guide_create(scene.handle, seed, C.byref(h)))
        self._h = h
        self.n_params = int(lib().hasty_guide_num_params())

    def close(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.hasty_guide_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_box(self, lo, hi):
        _err(lib().hasty_guide_set_box(self._h, _fptr(np.ascontiguousarray(lo, np.float32)),
                                       _fptr(np.ascontiguousarray(hi, np.float32))))

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError("expected %d parameters, got %d" % (self.n_params, p.size))
        _err(lib().hasty_guide_set_params(self._h, _fptr(p)))

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        _err(lib().hasty_guide_get_params(self._h, _fptr(p)))
        return p

    def infer(self, queries):
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 9)
        raw = np.zeros((q.shape[0], 48), np.float32)
        _err(lib().hasty_guide_infer(self._h, _fptr(q), q.shape[0], _fptr(raw)))
        return raw

    def set_selection(self, lo=0.0, hi=0.0):
        """Learned guide/BSDF selection probability in [lo, hi] (0 < lo < hi
        < 1); lo = hi = 0 returns to the scene's fixed fraction."""
        _err(lib().hasty_guide_set_selection(self._h, lo, hi))

    def gradient(self, samples):
        """Samples (n, 13): position, normal, wo, wi, F; or (n, 15) with the
        BSDF density of wi and the density wi was drawn with last."""
        sm = np.ascontiguousarray(samples, np.float32)
        stride = sm.shape[1] if sm.ndim == 2 else 13
        sm = sm.reshape(-1, stride)
        g = np.zeros(self.n_params, np.float32)
        loss = np.zeros(1, np.float32)
        _err(lib().hasty_guide_gradient_ex(self._h, _fptr(sm), sm.shape[0], stride, _fptr(g), _fptr(loss)))
        return g, float(loss[0])

    def train_step(self, samples, lr=1e-3):
        sm = np.ascontiguousarray(samples, np.float32).reshape(-1, 13)
        loss = np.zeros(1, np.float32)
        _err(lib().hasty_guide_train_step(self._h, _fptr(sm), sm.shape[0], lr, _fptr(loss)))
        return float(loss[0])


class Sppm:
    """One progressive photon-mapping job on the GPU (HastySppm*): per-pixel
    statistics and the radius image stay on the device between passes."""

    def __init__(self, scene: Scene, seed: int, alpha=0.4, footprint_scale=10.0, store_fraction=0.5,
                 emit_fraction=10.0, initial_radius=0.0, fixed_point_flux=False, flux_frac_bits=0,
                 gather_share_candidates=0, gather_union_factor=0.0):
        self.scene = scene  # keeps the scene alive
        p = SppmParamsC(alpha, footprint_scale, store_fraction, emit_fraction, initial_radius,
                        1 if fixed_point_flux else 0, flux_frac_bits, gather_share_candidates,
                        gather_union_factor)
        h = C.c_void_p()
        _err(lib().hasty_sppm_create(scene.handle, seed, C.byref(p), C.byref(h)))
        self._h = h
        info = scene.info()
        self.width, self.height = info["width"], info["height"]

    def close(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.hasty_sppm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def pass_(self, index: int, films=None):
        """One pass; returns (color, normal, albedo) holding the current estimate."""
        h, w = self.height, self.width
        if films is None:
            films = tuple(np.zeros((h, w, 3), np.float32) for _ in range(3))
        _check_films(films, h, w)
        _err(lib().hasty_sppm_pass(self._h, index, *(_fptr(f) for f in films)))
        return films

    def pass_device(self, index: int, d_color, d_normal, d_albedo):
        _err(lib().hasty_sppm_pass_device(self._h, index, C.c_void_p(d_color), C.c_void_p(d_normal),
                                          C.c_void_p(d_albedo)))

    def info(self) -> dict:
        i = SppmInfoC()
        _err(lib().hasty_sppm_info(self._h, C.byref(i)))
        return {"emitted": i.emitted, "stored": i.stored, "cell_size": i.cell_size, "passes_done": i.passes_done,
                "emit_limit": i.emit_limit, "store_limit": i.store_limit, "flux_overflows": i.flux_overflows,
                "gather_queries": i.gather_queries, "gather_candidates": i.gather_candidates,
                "gather_found": i.gather_found}

    def pixels(self) -> dict:
        """PMRenderJobPixel planes: radius, N, Ne, flux (h, w, 3), LSum (h, w, 3), Nd (int32)."""
        h, w = self.height, self.width
        a = np.zeros((10, h, w), np.float32)
        _err(lib().hasty_sppm_pixels(self._h, _fptr(a)))
        return {"radius": a[0], "N": a[1], "Ne": a[2], "flux": np.stack(a[3:6], -1), "LSum": np.stack(a[6:9], -1),
                "Nd": a[9].view(np.int32)}

    def radii(self) -> np.ndarray:
        a = np.zeros((self.height, self.width), np.float32)
        _err(lib().hasty_sppm_radii(self._h, _fptr(a)))
        return a

    def photons(self) -> np.ndarray:
        """(stored, 10): position, dot(n_g, w_i), w_i, flux in emission order."""
        n = self.info()["stored"]
        a = np.zeros((max(n, 1), 10), np.float32)
        _err(lib().hasty_sppm_photons(self._h, _fptr(a), n))
        return a[:n]


# ------------------------------------------------------------------------------
# Host mirror of the reference's render-thread interface.


class RenderSettings:
    """core/include/Hasty/RenderJob.h:15-30 (defaults) filled from the job file."""

    def __init__(self, info: dict):
        self.rouletteStartDepth = info["roulette_start_depth"]
        self.rouletteQ = info["roulette_q"]
        self.width = info["width"]
        self.height = info["height"]
        self.numThreads = info["num_threads"]
        self.maxDepth = info["max_depth"]
        self.numSamples = info["num_samples"]
        self.useMIS = info["use_mis"]


class RenderJobSample:
    def __init__(self, idx: int):
        self.idx = idx


class RenderJob:
    """RenderJob (RenderJob.h:42-64): settings, the GPU scene, a shared
    sample-index counter and the stop flag."""

    def __init__(self):
        self.scene = None
        self.renderSettings = None
        self.stopFlag = False
        self._next = 0
        self._lock = threading.Lock()

    def loadJSON(self, path: str):
        self.scene = Scene(path)
        self.renderSettings = RenderSettings(self.scene.info())

    def fetchSample(self) -> RenderJobSample:
        with self._lock:
            s = RenderJobSample(self._next)
            self._next += 1
            return s

    def fetchSamples(self, count: int) -> tuple:
        """A contiguous batch of sample indices (one GPU launch)."""
        with self._lock:
            first = self._next
            self._next += count
            return first, count


class RenderThreadData:
    def __init__(self, seed: int):
        self.seed = seed
        self.stoppedFlag = False


class AccumulationBuffer:
    def __init__(self, width: int, height: int):
        self.data = np.zeros((height, width, 3), np.float32)
        self.numSamples = 0


class Image3fAccDoubleBuffer:
    """DoubleBuffer<AccumulationBuffer<Vec3f>> (Image.h:18-45)."""

    def __init__(self, width: int, height: int):
        self._write = AccumulationBuffer(width, height)
        self._read = AccumulationBuffer(width, height)
        self._mutex = threading.Lock()

    def getWriteBuffer(self):
        return self._write

    def getReadBuffer(self):
        return self._read

    def lock(self):
        self._mutex.acquire()

    def unlock(self):
        self._mutex.release()

    def copyBuffer(self):
        with self._mutex:
            self._read.data = self._write.data.copy()
            self._read.numSamples = self._write.numSamples


def renderThread(colorBuffer, normalBuffer, albedoBuffer, job: RenderJob, threadData: RenderThreadData,
                 batch: int = 0):
    """PathTracing.cpp:98-160 with GPU batches: pulls `batch` sample indices at a
    time (default: all remaining), renders them in one call, publishes the sums
    and honours stopFlag between batches."""
    st = job.renderSettings
    for b in (colorBuffer, normalBuffer, albedoBuffer):
        b.getWriteBuffer().data[...] = 0.0
        b.getWriteBuffer().numSamples = 0
    while not job.stopFlag:
        remaining = st.numSamples - job._next
        if remaining <= 0:
            break
        count = min(batch, remaining) if batch > 0 else remaining
        first, count = job.fetchSamples(count)
        count = min(count, st.numSamples - first)
        if count <= 0:
            break
        job.scene.render_pt(threadData.seed, first, count, films=(colorBuffer.getWriteBuffer().data,
                                                                  normalBuffer.getWriteBuffer().data,
                                                                  albedoBuffer.getWriteBuffer().data))
        for b in (colorBuffer, normalBuffer, albedoBuffer):
            b.getWriteBuffer().numSamples += count
            b.copyBuffer()
    threadData.stoppedFlag = True


class PMRenderJob(RenderJob):
    """PMRenderJob (PhotonMapping.h:33-121): a RenderJob whose per-pixel
    progressive statistics live on the GPU (created by the first
    renderPhotonThread)."""

    def __init__(self, alpha: float = 0.4):
        super().__init__()
        self.alpha = alpha
        self.sppm = None


def renderPhotonThread(colorBuffer, normalBuffer, albedoBuffer, job: PMRenderJob, threadData: RenderThreadData):
    """PhotonMapping.cpp:814-933: one photon-map pass per fetched sample index
    until stopFlag or numSamples; each pass publishes the current estimate
    with numSamples = 1."""
    st = job.renderSettings
    if job.sppm is None:
        job.sppm = Sppm(job.scene, threadData.seed, alpha=job.alpha)
    for b in (colorBuffer, normalBuffer, albedoBuffer):
        b.getWriteBuffer().data[...] = 0.0
    sample = job.fetchSample()
    while not job.stopFlag and sample.idx < st.numSamples:
        job.sppm.pass_(sample.idx, films=(colorBuffer.getWriteBuffer().data, normalBuffer.getWriteBuffer().data,
                                          albedoBuffer.getWriteBuffer().data))
        for b in (colorBuffer, normalBuffer, albedoBuffer):
            b.getWriteBuffer().numSamples = 1
            b.copyBuffer()
        sample = job.fetchSample()
    threadData.stoppedFlag = True


def estimate_device(scene: Scene, films, num_samples, exposure: float, out_ptr: int) -> None:
    """hasty_estimate_device: computeEstimate of device films (pointers) with
    num_samples[k] samples each, scaled by 2^exposure, into out_ptr."""
    n = len(films)
    ptrs = (C.c_void_p * n)(*[C.c_void_p(int(p)) for p in films])
    counts = (C.c_uint32 * n)(*[int(c) for c in num_samples])
    _err(lib().hasty_estimate_device(scene.handle, ptrs, counts, n, float(exposure), C.c_void_p(int(out_ptr))))


def present_device(scene: Scene, image_ptr: int, rgba_ptr: int) -> None:
    """hasty_present_device: gamma 1/2.2 and RGBA8 packing of a device image."""
    _err(lib().hasty_present_device(scene.handle, C.c_void_p(int(image_ptr)), C.c_void_p(int(rgba_ptr))))


class Films:
    """Device films of a scene with snapshots for progressive display (HastyFilms*)."""

    def __init__(self, scene: Scene):
        h = C.c_void_p()
        _err(lib().hasty_films_create(scene.handle, C.byref(h)))
        self._h = h
        self.scene = scene
        i = scene.info()
        self.width, self.height = i["width"], i["height"]

    def render(self, seed: int, first_sample: int, num_samples: int):
        _err(lib().hasty_films_render(self._h, seed, first_sample, num_samples))

    def snapshot(self):
        _err(lib().hasty_films_snapshot(self._h))

    def read(self, wait: bool = True, out=None):
        """The last finished snapshot as (color, normal, albedo, samples), or
        None when there is nothing new (or, without wait, it is in flight)."""
        if out is None:
            out = tuple(np.zeros((self.height, self.width, 3), np.float32) for _ in range(3))
        _check_films(out, self.height, self.width)
        n = C.c_uint32()
        rc = lib().hasty_films_read(self._h, 1 if wait else 0, *(_fptr(f) for f in out), C.byref(n))
        if rc == 4:
            return None
        _err(rc)
        return (*out, n.value)

    def device(self):
        """(color, normal, albedo) device pointers and the samples added so far."""
        c, nn, a = C.c_void_p(), C.c_void_p(), C.c_void_p()
        k = C.c_uint32()
        _err(lib().hasty_films_device(self._h, C.byref(c), C.byref(nn), C.byref(a), C.byref(k)))
        return (c.value, nn.value, a.value), k.value

    @staticmethod
    def estimate(film_sets, exposure: float = 0.0):
        """hasty_films_estimate: (color, normal, albedo) estimates of the film
        sets (computeEstimate on the GPU, colour times 2^exposure)."""
        f0 = film_sets[0]
        out = tuple(np.zeros((f0.height, f0.width, 3), np.float32) for _ in range(3))
        hs = (C.c_void_p * len(film_sets))(*[f._h for f in film_sets])
        _err(lib().hasty_films_estimate(hs, len(film_sets), float(exposure), *(_fptr(o) for o in out)))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.hasty_films_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def read_image(path) -> np.ndarray:
    """hasty_image_read: (H, W, C) float32 pixels, row 0 at the top."""
    w, h, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
    p = os.fsencode(str(path))
    _err(lib().hasty_image_read(p, C.byref(w), C.byref(h), C.byref(c), None, 0))
    out = np.empty((h.value, w.value, c.value), np.float32)
    _err(lib().hasty_image_read(p, C.byref(w), C.byref(h), C.byref(c), _fptr(out), out.size))
    return out


def write_exr(path, rgb) -> None:
    """hasty_image_write_exr of an (H, W, 3) image."""
    a = np.ascontiguousarray(rgb, np.float32)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("write_exr takes an (H, W, 3) image")
    _err(lib().hasty_image_write_exr(os.fsencode(str(path)), a.shape[1], a.shape[0], _fptr(a)))


def computeEstimate(buffers) -> np.ndarray:
    """apps/ViewerUtils.cpp:95-121: sum of read buffers / (1e-7 + total samples)."""
    total = 0
    image = None
    for b in buffers:
        b.lock()
        try:
            rb = b.getReadBuffer()
            image = rb.data.copy() if image is None else image + rb.data
            total += rb.numSamples
        finally:
            b.unlock()
    return image / np.float32(1e-7 + float(total))
