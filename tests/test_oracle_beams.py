"""Spot-light beams pinned by closed forms (reference render_kernel.cu:560-652:
six midpoint samples along the camera ray up to the first hit, density
beam * intensity * 1e-3 / (1 + d^2 / 4) at distance d from the spot,
inside its cone and unshadowed).

The spot sits at the camera's position and aims where the camera looks
(straight down at a floor 3 units below); the camera is nearly
orthographic, so every ray runs along the beam with d = t.  With
dt = 3 / 6, the beam adds

    sum_k beam * I * 1e-3 / (1 + t_k^2 / 4) * dt,   t_k = (k + 1/2) dt,

to every pixel whose ray lies inside the cone (blend 0: no soft edge), and
nothing to a pixel whose ray lies outside it, which a very narrow cone
separates by pixel.  The difference of renders with and without the beam
removes the floor's own lighting.  A wrong sample position, density law or
cone test fails it.
"""
import math
import os

import numpy as np
import pytest

from sheen_amd import contract as K
from util import PyScene, f3, material

W, H, AA = 24, 16, 2
FOCAL, SENSOR, CAM_Y = 5000.0, 50.0, 3.0
I, BEAM = 50.0, 2.0


def _scene(beam, angle):
    ps = PyScene()
    ps.set("materials", K.Material, [material(color=(0.8, 0.6, 0.4), spec=0.0, amb=0.1)], "num_materials")
    t = K.Triangle()
    a = 50.0
    t.v0, t.v1, t.v2 = f3(-a, 0, a), f3(a, 0, a), f3(-a, 0, -3 * a)
    t.n0 = t.n1 = t.n2 = f3(0, 0, 0)
    t.material_id = 0
    ps.set("triangles", K.Triangle, [t], "num_triangles")
    cam = ps.s.camera
    cam.position, cam.forward, cam.right, cam.up = f3(0, CAM_Y, 0), f3(0, -1, 0), f3(1, 0, 0), f3(0, 0, -1)
    cam.focal_length, cam.sensor_width = FOCAL, SENSOR
    L = K.Light()
    L.type, L.position, L.direction, L.color, L.intensity = 3, f3(0, CAM_Y, 0), f3(0, -1, 0), f3(1, 1, 1), I
    L.angle, L.blend, L.radius, L.beam, L.flags, L.texture_id = float(angle), 0.0, 0.0, float(beam), 1, -1
    ps.set("lights", K.Light, [L], "num_lights")
    return ps


def _oracle(ps):
    from oracle import Oracle
    out = Oracle(ps.ptr, threads=min(8, os.cpu_count() or 1)).render(W, H, AA, hits=True)
    assert (out["kind"] >= 0).all()
    return out["image"][..., :3].astype(np.float64)


def _beam_sum():
    dt = CAM_Y / 6.0
    return sum(BEAM * I * 1e-3 / (1.0 + 0.25 * ((k + 0.5) * dt) ** 2) * dt for k in range(6))


def _pixel_angles():
    """Smallest and largest angle between the axis and a ray through each
    pixel's footprint (corners, and 0 for the pixels around the centre)."""
    tan_h = 0.5 * SENSOR / FOCAL
    tan_x, tan_y = tan_h, tan_h * H / W
    lo = np.zeros((H, W))
    hi = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            us = [(xx / W * 2 - 1) * tan_x for xx in (x, x + 1)]
            vs = [(1 - yy / H * 2) * tan_y for yy in (y, y + 1)]
            r = [math.atan(math.hypot(u, v)) for u in us for v in vs]
            umin = 0.0 if us[0] <= 0 <= us[1] else min(abs(u) for u in us)
            vmin = 0.0 if min(vs) <= 0 <= max(vs) else min(abs(v) for v in vs)
            lo[y, x], hi[y, x] = math.atan(math.hypot(umin, vmin)), max(r)
    return lo, hi


def test_beam_closed_form():
    wide = 0.5
    img = _oracle(_scene(BEAM, wide)) - _oracle(_scene(0.0, wide))
    np.testing.assert_allclose(img, np.full(img.shape, _beam_sum()), rtol=1e-4)
    narrow = 0.0025
    lo, hi = _pixel_angles()
    inside, outside = hi < narrow * 0.999, lo > narrow * 1.001
    assert inside.sum() >= 4 and outside.sum() >= 40
    img = _oracle(_scene(BEAM, narrow)) - _oracle(_scene(0.0, narrow))
    np.testing.assert_allclose(img[inside], np.full(img[inside].shape, _beam_sum()), rtol=1e-4)
    np.testing.assert_allclose(img[outside], 0.0, atol=1e-7)


@pytest.mark.gpu
def test_beam_gpu_matches_oracle():
    from sheen_amd import sheen
    for angle in (0.5, 0.0025):
        ps = _scene(BEAM, angle)
        with sheen.Context(ps.ptr) as ctx:
            img, _ = ctx.render(W, H, AA)
        np.testing.assert_allclose(img[..., :3], _oracle(ps), rtol=1e-4, atol=1e-6)
