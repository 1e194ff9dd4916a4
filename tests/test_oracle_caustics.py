"""The caustic radiance term pinned by a brute-force gather (reference
render_kernel.cu:288-296: base * flux * caustics / (pi^2 r^2), r = cell / 2).

A glass ball over a Lambert floor under a point light; the oracle's photon
map is read back and a floor point p0 is chosen that takes between 5 and
120 photons within r (no photon within 2e-3 of the boundary).  The gather
visits the at most 8 hash buckets of the cells around p0 (r is half a cell),
so below its 128-photon cap it sums exactly the stored photons within r
whose normal faces the floor's (n . normal >= 0.1).  A nearly orthographic
camera that sees only p0 (within 1e-3) then adds

    albedo * sum(flux) * caustics / (pi^2 r^2)

to every pixel when caustics goes from 0 to c (the photon map is the same:
it depends on the photon count and cell only).  A wrong normalisation, a
missing albedo or a radius other than half a cell fails it.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from sheen_amd import contract as K
from util import PyScene, f3, material

W, H, AA = 8, 8, 2
ALBEDO, CELL, PHOTONS, CAUSTICS = (0.8, 0.6, 0.4), 0.5, 100000, 1.5
PHOTON = np.dtype([("position", "<f4", 3), ("direction", "<f4", 3), ("normal", "<f4", 3), ("flux", "<f4", 3)])


def _scene(p0, caustics):
    ps = PyScene()
    glass = material(color=(1, 1, 1), spec=0.0, amb=0.0)
    glass.transmission, glass.ior = 1.0, 1.5
    ps.set("materials", K.Material, [material(color=ALBEDO, spec=0.0, amb=0.1), glass], "num_materials")
    t = K.Triangle()
    a = 50.0
    t.v0, t.v1, t.v2 = f3(-a, 0, a), f3(a, 0, a), f3(-a, 0, -3 * a)
    t.n0 = t.n1 = t.n2 = f3(0, 0, 0)
    t.material_id = 0
    ps.set("triangles", K.Triangle, [t], "num_triangles")
    o = K.Object()
    o.type, o.material_id, o.param1, o.param2 = 0, 1, f3(0, 1, 0), 0.5
    ps.set("objects", K.Object, [o], "num_objects")
    L = K.Light()
    L.type, L.position, L.color, L.intensity, L.radius, L.flags, L.texture_id = 1, f3(0, 3, 0), f3(1, 1, 1), \
        40.0, 0.0, 1, -1
    ps.set("lights", K.Light, [L], "num_lights")
    ps.s.photons, ps.s.photon_cell, ps.s.caustics = PHOTONS, CELL, float(caustics)
    # looking at p0 from (+3, +3, 0) away, past the ball
    cam = ps.s.camera
    f = np.array([-1.0, -1.0, 0.0]) / math.sqrt(2.0)
    up = np.array([-1.0, 1.0, 0.0]) / math.sqrt(2.0)
    cam.position = f3(*(np.asarray(p0) - 3 * math.sqrt(2.0) * f))
    cam.forward, cam.right, cam.up = f3(*f), f3(0, 0, -1), f3(*up)
    cam.focal_length, cam.sensor_width = 5000.0, 1.0
    return ps


def _photons(oracle):
    n = oracle.photon_stats()[1]
    buf = np.zeros(n, PHOTON)
    got = oracle._L.oracle_photon_map(oracle._h, buf.ctypes.data_as(C.c_void_p), None, None)
    assert got == n
    return buf


def test_caustic_term_matches_brute_force_gather():
    from oracle import Oracle
    threads = min(8, os.cpu_count() or 1)
    o = Oracle(_scene((0, 0, 0), CAUSTICS).ptr, threads=threads)
    ph = _photons(o)
    floor = ph[np.abs(ph["position"][:, 1]) < 1e-3]
    assert len(floor) > 100
    r = CELL * 0.5
    p0 = None
    for cand in floor["position"][::7]:
        c = np.array([cand[0] + 0.013, 0.0, cand[2] - 0.007])
        d2 = ((ph["position"] - c) ** 2).sum(axis=1)
        near = np.abs(np.sqrt(d2) - r) < 2e-3
        take = (d2 <= r * r) & (ph["normal"] @ np.array([0.0, 1.0, 0.0]) >= 0.1)
        if 5 <= take.sum() <= 120 and not near.any():
            p0, flux = c, ph["flux"][take].astype(np.float64).sum(axis=0)
            break
    assert p0 is not None
    on = Oracle(_scene(p0, CAUSTICS).ptr, threads=threads).render(W, H, AA, hits=True)
    off = Oracle(_scene(p0, 0.0).ptr, threads=threads).render(W, H, AA)
    assert (on["kind"] >= 0).all()
    term = on["image"][..., :3].astype(np.float64) - off["image"][..., :3]
    expect = np.array(ALBEDO) * flux * CAUSTICS / (math.pi ** 2 * r * r)
    np.testing.assert_allclose(term, np.broadcast_to(expect, term.shape), rtol=1e-4)


@pytest.mark.gpu
def test_caustic_term_gpu_matches_oracle():
    from oracle import Oracle
    from sheen_amd import sheen
    ps = _scene((0.31, 0.0, -0.17), CAUSTICS)
    with sheen.Context(ps.ptr) as ctx:
        img, _ = ctx.render(W, H, AA)
    ref = Oracle(ps.ptr, threads=min(8, os.cpu_count() or 1)).render(W, H, AA)["image"]
    np.testing.assert_allclose(img[..., :3], ref[..., :3], rtol=1e-4, atol=1e-6)
