"""Initial-condition inputs: the library's seeded noise against the reference's compiled
noise core (bitwise) and the reference's own noise invariants (tests/test_noise.c);
the initial flow's normalisation and bands; the seeded banded strip.  CPU only."""
import ctypes as C

import numpy as np
import pytest


def lib_noise(ggb, seed):
    L = ggb.lib()

    class Noise(C.Structure):
        _fields_ = [("perm", C.c_ubyte * 512)]

    class CurlParams(C.Structure):
        _fields_ = [("scale", C.c_double), ("lacunarity", C.c_double), ("gain", C.c_double),
                    ("octaves", C.c_int)]

    L.noise_init.argtypes = [C.POINTER(Noise), C.c_uint]
    L.simplex3.restype = C.c_double
    L.simplex3.argtypes = [C.POINTER(Noise), ggb.CVec3, C.POINTER(ggb.CVec3)]
    L.curl_noise.restype = ggb.CVec3
    L.curl_noise.argtypes = [C.POINTER(Noise), ggb.CVec3, C.POINTER(CurlParams)]
    nz = Noise()
    L.noise_init(C.byref(nz), seed)

    def simplex(p):
        g = ggb.CVec3()
        v = L.simplex3(C.byref(nz), ggb.CVec3(*map(float, p)), C.byref(g))
        return v, np.array([g.x, g.y, g.z])

    def curl(p, scale=3.0, lac=2.0, gain=0.5, octaves=4):
        r = L.curl_noise(C.byref(nz), ggb.CVec3(*map(float, p)), C.byref(CurlParams(scale, lac, gain, octaves)))
        return np.array([r.x, r.y, r.z])

    return np.array(nz.perm[:], np.uint8), simplex, curl


@pytest.mark.parametrize("seed", [1, 42, 123456789])
def test_noise_matches_compiled_reference(ggb, orc, ref, seed):
    perm, simplex, curl = lib_noise(ggb, seed)
    rn = orc.RefNoise(seed)
    assert np.array_equal(perm, rn.perm)
    rng = np.random.default_rng(seed)
    for p in rng.uniform(-10, 10, (300, 3)):
        v, g = simplex(p)
        rv, rg = rn.simplex3(p)
        assert v == rv and np.array_equal(g, rg)
    for p in rng.standard_normal((100, 3)):
        p /= np.linalg.norm(p)
        assert np.array_equal(curl(p), rn.curl(p))


def test_noise_invariants(ggb):
    """tests/test_noise.c:42-78: analytic gradient vs central differences; curl noise is
    divergence-free."""
    _, simplex, curl = lib_noise(ggb, 42)
    rng = np.random.default_rng(7)
    e = 1e-5
    for p in rng.uniform(-10, 10, (300, 3)):
        _, g = simplex(p)
        fd = np.array([(simplex(p + e * d)[0] - simplex(p - e * d)[0]) / (2 * e) for d in np.eye(3)])
        assert np.linalg.norm(g - fd) / (1 + np.linalg.norm(fd)) < 1e-4
    maxdiv = maxmag = 0
    for p in rng.standard_normal((200, 3)):
        p /= np.linalg.norm(p)
        c = curl(p)
        div = sum((curl(p + e * d)[k] - curl(p - e * d)[k]) / (2 * e) for k, d in enumerate(np.eye(3)))
        maxdiv = max(maxdiv, abs(div))
        maxmag = max(maxmag, np.linalg.norm(c))
    assert maxdiv < 1e-3 * maxmag


def test_initial_velocity(ggb):
    """main.c:320-360: tangent noise with area-weighted RMS = noise_amp, plus bands
    u_east = A sin(n colat)."""
    g = ggb.Grid(32)
    p = g.pos.astype(np.float64)
    a = g.area.astype(np.float64)
    vel, vband = g.initial_velocity(seed=3, bands=8, band_amp=1.0, noise_amp=0.5, octaves=3)
    vel = vel.reshape(-1, 3).astype(np.float64)
    vband = vband.reshape(-1, 3).astype(np.float64)
    noise = vel - vband
    assert np.max(np.abs(np.sum(noise * p, 1))) < 1e-6  # tangent
    rms = np.sqrt(np.sum(np.sum(noise ** 2, 1) * a) / g.total_area)
    assert abs(rms - 0.5) < 1e-5
    colat = np.arccos(np.clip(p[:, 1], -1, 1))
    s = np.sin(colat)
    east = np.cross(np.array([0.0, 1.0, 0.0]), p)  # |y x p| = sin(colat)
    want = east * (np.sin(8 * colat) / s)[:, None]
    assert np.max(np.abs(vband - want)) < 1e-6
    again, _ = g.initial_velocity(seed=3, bands=8, band_amp=1.0, noise_amp=0.5, octaves=3)
    assert np.array_equal(again.reshape(-1, 3), vel.astype(np.float32))


def test_banded_strip(ggb):
    img = ggb.banded_strip(9)
    assert img.shape == (1024, 64, 4)
    assert np.all(img[..., 3] == 255)
    assert np.all(img == img[:, :1])  # horizontal bands
    rows = img[:, 0, :3].astype(int)
    change = np.nonzero(np.any(np.diff(rows, axis=0) != 0, axis=1))[0] + 1
    heights = np.diff(np.concatenate([[0], change, [1024]]))
    assert len(heights) <= 24
    assert np.all(heights[:-1] >= 16) and np.all(heights[:-1] <= 64)
    assert np.array_equal(img, ggb.banded_strip(9))
    assert not np.array_equal(img, ggb.banded_strip(10))
