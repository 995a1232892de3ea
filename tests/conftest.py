import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def scene_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("scenes")


@pytest.fixture(scope="session")
def cornell_job(scene_dir):
    """Procedural Cornell box at its benchmark settings (256^2, 64 spp, depth 8, RR from 3)."""
    from hastyrenderer_amd import scenes

    return scenes.write_cornell(str(scene_dir / "cornell"))


@pytest.fixture(scope="session")
def ggx_job(scene_dir):
    """The 7x7 GGX sphere grid with its 2048x1024 sky, thin lens and env sampling."""
    from hastyrenderer_amd import scenes

    return scenes.write_ggx_grid(str(scene_dir / "ggx"))


@pytest.fixture(scope="session")
def ggx_small_job(scene_dir):
    """The GGX grid with subdivision-2 spheres and a 256x128 sky (fast checks)."""
    from hastyrenderer_amd import scenes

    return scenes.write_ggx_grid(str(scene_dir / "ggx_small"), width=96, height=96, spp=4, subdiv=2,
                                 env_size=(256, 128))


@pytest.fixture(scope="session")
def cornell_sky_job(scene_dir):
    """Cornell box whose open front sees a sky map: NEE picks surface light or
    environment with probability 1/2 each, environment visibility needs the
    closest hit (a light may be in front)."""
    from hastyrenderer_amd import scenes

    return scenes.write_cornell(str(scene_dir / "cornell_sky"), width=128, height=128, spp=16, sky=True)


@pytest.fixture(scope="session")
def caustic_small_job(scene_dir):
    """Caustic box with a 4,608-triangle torus: dispersive glass prism (media
    stack, hero wavelength, refraction), GGX gold ring, 1% ceiling light."""
    from hastyrenderer_amd import scenes

    return scenes.write_caustic_scene(str(scene_dir / "caustic_small"), width=192, height=108, passes=4,
                                      n_major=96, n_minor=24)


_TK = None


def load_testkernels():
    """tests/libhasty_testkernels.so: device primitives over plain C."""
    global _TK
    if _TK is None:
        path = os.path.join(ROOT, "tests", "libhasty_testkernels.so")
        if not os.path.exists(path):
            raise RuntimeError("tests/libhasty_testkernels.so not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        L.tk_math.argtypes = [C.c_int, fp, fp, C.c_uint32, fp]
        L.tk_philox.argtypes = [up, C.c_uint32, up]
        L.tk_vndf.argtypes = [fp, C.c_uint32, fp]
        L.tk_sampler.argtypes = [C.c_int, C.c_uint32, C.c_uint32, fp, fp, C.c_float, fp]
        L.tk_bsdf_eval.argtypes = [fp, fp, fp, C.c_uint32, fp]
        L.tk_mfma.argtypes = [fp, fp, fp]
        L.tk_bsdf_sample.argtypes = [fp, fp, fp, C.c_uint32, C.c_uint32, fp]
        hp = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
        L.tk_guide_tables.argtypes = [hp, C.c_uint32, fp, C.c_uint32, fp, fp, fp, C.c_uint32, fp, fp]
        L.tk_guide_mixture.argtypes = [fp, C.c_uint32, C.c_uint32, fp, fp, fp, fp]
        _TK = L
    return _TK


def rel_ok(gpu, ref, rtol=1e-3, atol=1e-6):
    """Per-pixel check used for the parity bar: every channel within rtol of the
    oracle (absolute atol for values that are ~0)."""
    diff = np.abs(gpu.astype(np.float64) - ref.astype(np.float64))
    tol = np.maximum(rtol * np.abs(ref.astype(np.float64)), atol)
    return np.all(diff <= tol, axis=-1)
