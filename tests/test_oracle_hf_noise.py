"""Procedural height-field noise of the oracle pinned (reference
primitives/heightfield.h:13-98), each pin checked to fail with the planted
mistakes it targets (pin_util.run_pin).

* hf_cell_hash(x, z, seed) is an integer hash evaluated in signed 32-bit
  arithmetic in the reference (x 374761393 + z 668265263 +
  seed 974634097, wrapping; then h = (h ^ h >> 13) 1274126177 and
  h ^= h >> 16 on unsigned bits; value (h & 0xffffff) / 2^24).  Known
  answers come from numpy int32 arithmetic with wrap-around, an independent
  evaluation of those signed products.
* value noise equals the hash at lattice points, the mean of two lattice
  values half way along an edge, and the bilinear mix with smoothstep
  weights s(0.25) = 0.15625, s(0.75) = 0.84375 elsewhere;
* fBm with every octave on lattice points (x freq integer) is the
  amplitude-weighted mean of hashes with octave seeds seed + 131 o;
* hf_eval at the tile corners (local coordinates 0 or 1, where the detail
  octave at 7x is on the lattice too) is
  min(fbm + 0.06 hash(7 lx, 7 lz, seed + 7777) - 0.03, 1) * height;
* a ray cast straight down onto a procedural tile lands on that surface
  within the march's refinement tolerance.
"""
import numpy as np
import pytest

from sheen_amd import contract as K
from pin_util import run_pin
from util import PyScene, f3, material


def _hash_ref(x, z, seed):
    with np.errstate(over="ignore"):
        h = (np.int32(x) * np.int32(374761393) + np.int32(z) * np.int32(668265263) +
             np.int32(seed) * np.int32(974634097))
    h = int(np.uint32(h.view(np.uint32)))
    h = ((h ^ (h >> 13)) * 1274126177) & 0xFFFFFFFF
    h ^= h >> 16
    return np.float32((h & 0xFFFFFF) / 16777216.0)


POINTS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, -1, 3), (12345, -678, 42), (-2147483648, 2147483647, -7),
          (77, 91, 7777), (3, 5, 131), (-40, 17, -1)]


def _smooth(t):
    return t * t * (3 - 2 * t)


def test_hash_value_noise_and_fbm():
    def check(mode):
        L = mode.lib()
        try:
            for (x, z, sd) in POINTS:
                assert L.oracle_hf_cell_hash(x, z, sd) == _hash_ref(x, z, sd), (x, z, sd)
                xi, zi = max(min(x, 10 ** 6), -10 ** 6), max(min(z, 10 ** 6), -10 ** 6)
                assert L.oracle_hf_value_noise(float(xi), float(zi), sd) == _hash_ref(xi, zi, sd)
            for (x0, z0, sd) in [(0, 0, 5), (-3, 7, 11), (40, -9, 2)]:
                h = {(a, b): float(_hash_ref(x0 + a, z0 + b, sd)) for a in (0, 1) for b in (0, 1)}
                got = L.oracle_hf_value_noise(x0 + 0.5, float(z0), sd)
                assert abs(got - 0.5 * (h[0, 0] + h[1, 0])) <= 1e-6, (got, h)
                tx, tz = _smooth(0.25), _smooth(0.75)
                a = h[0, 0] + (h[1, 0] - h[0, 0]) * tx
                b = h[0, 1] + (h[1, 1] - h[0, 1]) * tx
                got = L.oracle_hf_value_noise(x0 + 0.25, z0 + 0.75, sd)
                assert abs(got - (a + (b - a) * tz)) <= 1e-6, (got, a + (b - a) * tz)
            for (x, z, freq, octaves, sd) in [(1.0, 2.0, 3.0, 4, 9), (0.5, -1.5, 2.0, 6, 123), (3.0, 0.0, 1.0, 1, 5)]:
                num = den = 0.0
                for o in range(octaves):
                    amp = 0.5 ** (o + 1)
                    k = freq * 2 ** o
                    num += amp * float(_hash_ref(int(x * k), int(z * k), sd + 131 * o))
                    den += amp
                got = L.oracle_hf_fbm(x, z, freq, octaves, sd)
                assert abs(got - num / den) <= 2e-6, (x, z, got, num / den)
        finally:
            mode.done()

    run_pin(check, ["HF_HASH", "HF_LINEAR", "HF_OCT_SEED"])


def _tile(cx=0.25, cz=-0.5, size=2.0, height=1.5, freq=3.0, octaves=5, seed=17):
    hf = K.HeightField()
    hf.center, hf.size_x, hf.size_z, hf.height = f3(cx, 0.2, cz), size, size, height
    hf.freq, hf.octaves, hf.seed, hf.material_id = freq, octaves, seed, 0
    hf.img_w = hf.img_h = 0
    return hf


def _eval_ref(hf, lx, lz):
    num = den = 0.0
    for o in range(hf.octaves):
        amp = 0.5 ** (o + 1)
        k = hf.freq * 2 ** o
        num += amp * float(_hash_ref(int(lx * k), int(lz * k), hf.seed + 131 * o))
        den += amp
    t = num / den + 0.06 * float(_hash_ref(7 * lx, 7 * lz, hf.seed + 7777)) - 0.03
    return min(t, 1.0) * hf.height


def test_height_at_tile_corners():
    import ctypes as C

    def check(mode):
        L = mode.lib()
        try:
            for seed in (17, -5, 90210):
                hf = _tile(seed=seed)
                for lx, lz in ((0, 0), (1, 1), (1, 0), (0, 1)):
                    x = hf.center.x + (lx - 0.5) * hf.size_x
                    z = hf.center.z + (lz - 0.5) * hf.size_z
                    got = L.oracle_hf_eval(C.byref(hf), x, z)
                    assert abs(got - _eval_ref(hf, lx, lz)) <= 3e-6, (seed, lx, lz, got, _eval_ref(hf, lx, lz))
                # outside the tile the height is 0
                assert L.oracle_hf_eval(C.byref(hf), hf.center.x + hf.size_x, hf.center.z) == 0.0
        finally:
            mode.done()

    run_pin(check, ["HF_HASH", "HF_DETAIL", "HF_OCT_SEED"])


def test_procedural_tile_hits_lie_on_the_surface(oracle_mod):
    """Rays straight down hit at y = centre.y + hf_eval(x, z) within the
    bisection tolerance (march step diag/64, 10 halvings)."""
    import ctypes as C
    hf = _tile()
    sc = PyScene()
    sc.set("materials", K.Material, [material()], "num_materials")
    sc.set("heightfields", K.HeightField, [hf], "num_heightfields")
    o = oracle_mod.Oracle(sc.ptr, threads=1)
    rng = np.random.default_rng(3)
    xz = rng.uniform(-0.9, 0.9, size=(200, 2)) + np.array([hf.center.x, hf.center.z])
    orig = np.stack([xz[:, 0], np.full(200, 5.0), xz[:, 1]], 1).astype(np.float32)
    d = np.tile(np.array([[0, -1, 0]], np.float32), (200, 1))
    res = o.trace(orig, d)
    assert (res["kind"] == 3).all()
    P = res["point"]
    L = oracle_mod.lib()
    h = np.array([L.oracle_hf_eval(C.byref(hf), float(p[0]), float(p[2])) for p in P])
    np.testing.assert_allclose(P[:, 1], hf.center.y + h, atol=2e-3)
