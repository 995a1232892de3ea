"""hasty_trace_hits_device_noalloc against hasty_trace_hits_device.

The no-allocation entry point runs the same closest-hit kernel with its
traversal stack overflow in the library's static scratch array instead of a
scene buffer, on a smaller fixed grid.  Each ray's traversal is independent
of the grid, so the hit records must be identical byte for byte.  Two calls
in a row on each scene, the second with fewer rays (which must leave the
records past its end untouched), plus the refusals the entry point makes
without touching the GPU.
"""
import ctypes as C

import numpy as np
import pytest

import hastyrenderer_amd as ha

pytestmark = pytest.mark.gpu

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        _HIP = h
    return _HIP


class _Dev:
    def __init__(self, nbytes):
        p = C.c_void_p()
        assert _hip().hipMalloc(C.byref(p), nbytes) == 0
        self.ptr, self.nbytes = p.value, nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert _hip().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def download(self, nbytes):
        out = np.zeros(nbytes, np.uint8)
        assert _hip().hipMemcpy(out.ctypes.data, C.c_void_p(self.ptr), nbytes, 2) == 0  # device to host
        return out

    def fill(self, byte):
        assert _hip().hipMemset(C.c_void_p(self.ptr), byte, self.nbytes) == 0

    def free(self):
        _hip().hipFree(C.c_void_p(self.ptr))


def _noalloc(scene, d_rays, n, d_hits):
    fn = ha.lib().hasty_trace_hits_device_noalloc
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    return fn(scene.handle, C.c_void_p(d_rays), n, C.c_void_p(d_hits))


def _rays(n, rng, lo=-1.2, hi=2.2):
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


@pytest.mark.parametrize("which", ["ggx_small", "cornell"])
def test_noalloc_hits_equal_trace_hits_device(which, cornell_job, ggx_small_job):
    job = ggx_small_job if which == "ggx_small" else cornell_job
    scene = ha.Scene(job)
    rng = np.random.default_rng(31)
    n1, n2 = 150000, 4099
    rays = _Dev(n1 * 24)
    ref = _Dev(n1 * 32)
    got = _Dev(n1 * 32)
    try:
        got.fill(0xA5)
        prev = None
        for n in (n1, n2):
            print("[noalloc] scene %s, %d rays" % (which, n), flush=True)
            r = _rays(n, rng)
            rays.upload(r)
            scene.trace_hits_device(rays.ptr, n, ref.ptr)
            rc = _noalloc(scene, rays.ptr, n, got.ptr)
            assert rc == 0, ha.lib().hasty_last_error().decode()
            a = ref.download(n * 32)
            b = got.download(n1 * 32)
            assert np.array_equal(a, b[: n * 32])
            hits = b[: n * 32].view(ha.HIT_DTYPE)
            assert (hits["geom_id"] != 0xFFFFFFFF).sum() > n // 20
            # the host-pointer query gives the same records
            assert np.array_equal(scene.trace_hits(r).view(np.uint8), a)
            if prev is not None:  # records past the second call's end keep the first call's values
                assert np.array_equal(b[n * 32:], prev[n * 32:])
            prev = b
    finally:
        for d in (rays, ref, got):
            d.free()
        scene.close()


def test_noalloc_refusals(cornell_job):
    """Refused on the host before any launch: misaligned hits, more than
    2^31 - 1 rays, null pointers; zero rays is a no-op."""
    scene = ha.Scene(cornell_job)
    buf = _Dev(4096)
    try:
        assert _noalloc(scene, buf.ptr, 4, buf.ptr + 4) == 1
        assert b"aligned" in ha.lib().hasty_last_error()
        assert _noalloc(scene, buf.ptr, 0x80000000, buf.ptr) == 1
        assert b"2^31" in ha.lib().hasty_last_error()
        assert _noalloc(scene, None, 4, buf.ptr) == 1
        assert _noalloc(scene, buf.ptr, 0, buf.ptr) == 0
    finally:
        buf.free()
        scene.close()
