"""GPU device primitives vs the CPU oracle, bit for bit.

The path tracer's decisions only agree with the oracle if every value that
feeds a branch is computed identically on gfx950 and on the host.  These tests
run the device code (tests/libhasty_testkernels.so, built from the renderer's
own pt_device.h) on the same inputs as the oracle and require equal bits.
"""
import numpy as np
import pytest

from conftest import load_testkernels as _kernels
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(1234)


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("fn,lo,hi,blo,bhi", [
    (0, -40.0, 40.0, 0, 1),      # sin
    (1, -40.0, 40.0, 0, 1),      # cos
    (2, -1.5, 1.5, 0, 1),        # tan
    (3, -5.0, 5.0, -5.0, 5.0),   # atan2
    (4, -1.0, 1.0, 0, 1),        # acos
    (5, 1e-4, 1.0, 0.0, 40.0),   # pow (Beer's law)
    (6, 0.0, 20.0, 0.785398, 0.785399),  # fmod by the 8-blade sector
    (7, -30.0, 30.0, 0, 1),      # atan
    (8, 0.0, 100.0, 0, 1),       # sqrt (correctly rounded)
    (9, -10.0, 10.0, 0.1, 10.0),  # division (correctly rounded)
])
def test_math_bitexact(fn, lo, hi, blo, bhi):
    n = 1 << 18
    a = RNG.uniform(lo, hi, n).astype(np.float32)
    b = RNG.uniform(blo, bhi, n).astype(np.float32)
    # include edge values
    a[:4] = np.array([lo, hi, 0.0, -0.0], np.float32).clip(lo, hi)
    out = np.zeros(n, np.float32)
    assert _kernels().tk_math(fn, a, b, n, out) == 0
    ref = po.hm_vec(fn, a, b)
    same = (_bits(out) == _bits(ref)) | (np.isnan(out) & np.isnan(ref))
    assert same.all(), (fn, a[~same][:5], out[~same][:5], ref[~same][:5])


def test_philox_bitexact():
    n = 4096
    ctr = RNG.integers(0, 2**32, size=(n, 6), dtype=np.uint64).astype(np.uint32)
    ctr[0] = [0, 0, 0, 0, 0, 0]
    ctr[1] = [0xFFFFFFFF] * 6
    ctr[2] = [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344, 0xA4093822, 0x299F31D0]
    out = np.zeros((n, 4), np.uint32)
    assert _kernels().tk_philox(np.ascontiguousarray(ctr), n, out) == 0
    # Random123 known-answer vectors for Philox-4x32-10
    assert list(out[0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(out[1]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(out[2]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]
    for i in range(0, n, 97):
        assert list(out[i]) == list(po.philox(ctr[i, :4], ctr[i, 4:]))


def test_vndf_bitexact():
    n = 1 << 16
    rows = np.zeros((n, 7), np.float32)
    rows[:, 0:2] = RNG.random((n, 2))
    rows[:, 2:4] = RNG.uniform(0.01, 1.0, (n, 2))
    w = RNG.normal(size=(n, 3))
    w[:, 2] = np.abs(w[:, 2])
    rows[:, 4:7] = w / np.linalg.norm(w, axis=1, keepdims=True)
    out = np.zeros((n, 3), np.float32)
    assert _kernels().tk_vndf(rows, n, out) == 0
    ref = po.sample_vndf(rows)
    assert (_bits(out) == _bits(ref)).all()


@pytest.mark.parametrize("kind,param", [(0, 0.0), (1, 0.3), (1, 0.02), (1, 1.0), (3, 8.0), (3, 5.0), (4, 0.0)])
def test_sampler_sequences_bitexact(kind, param):
    count = 4000
    nrm = np.array([0.3, 0.8, -0.52], np.float32)
    nrm /= np.linalg.norm(nrm)
    wo = np.array([0.1, 0.9, 0.3], np.float32)
    wo /= np.linalg.norm(wo)
    out = np.zeros((count, 5), np.float32)
    assert _kernels().tk_sampler(kind, 7, count, nrm, wo, param, out) == 0
    ref = po.sampler(kind, 7, count, nrm, wo, param)
    assert (_bits(out) == _bits(ref)).all()


MATERIALS = [
    # albedo rgb, roughness, metallic, specular, ior, aniso, transmission
    [0.73, 0.73, 0.73, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0],     # Lambertian wall
    [0.9, 0.6, 0.5, 0.3, 1.0, 0.0, 1.0, 0.0, 0.0],        # rough conductor
    [0.9, 0.8, 0.7, 0.1414, 1.0, 0.0, 1.0, 0.0, 0.0],     # alpha 0.02 conductor
    [0.5, 0.4, 0.3, 0.5, 0.0, 0.08, 1.5, 0.0, 0.0],       # opaque dielectric
    [0.8, 0.8, 0.8, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0],        # mirror
    [1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1.5, 0.0, 1.0],        # smooth glass
    [0.6, 0.7, 0.8, 0.6, 0.0, 0.5, 1.0, 0.5, 0.0],        # anisotropic mix
]


@pytest.mark.parametrize("mi", range(len(MATERIALS)))
def test_bsdf_eval_bitexact(mi):
    p = np.array(MATERIALS[mi], np.float32)
    n = np.array([0.0, 0.6, 0.8], np.float32)
    count = 8192
    d = RNG.normal(size=(count, 6))
    d[:, :3] /= np.linalg.norm(d[:, :3], axis=1, keepdims=True)
    d[:, 3:] /= np.linalg.norm(d[:, 3:], axis=1, keepdims=True)
    d = d.astype(np.float32)
    out = np.zeros((count, 7), np.float32)
    assert _kernels().tk_bsdf_eval(p, n, d, count, out) == 0
    ref = po.bsdf_eval(p, n, d)
    same = (_bits(out) == _bits(ref)) | (np.isnan(out) & np.isnan(ref))
    assert same.all()


@pytest.mark.parametrize("mi", range(len(MATERIALS)))
def test_bsdf_sample_bitexact(mi):
    p = np.array(MATERIALS[mi], np.float32)
    n = np.array([0.0, 0.6, 0.8], np.float32)
    wo = np.array([0.3, 0.5, 0.81], np.float32)
    wo /= np.linalg.norm(wo)
    count = 4000
    out = np.zeros((count, 7), np.float32)
    assert _kernels().tk_bsdf_sample(p, n, wo, 11, count, out) == 0
    ref = po.bsdf_sample(p, n, wo, 11, count)
    same = (_bits(out) == _bits(ref)) | (np.isnan(out) & np.isnan(ref))
    assert same.all()
