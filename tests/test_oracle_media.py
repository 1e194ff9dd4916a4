"""Homogeneous media pinned by closed forms (reference render_kernel.cu,
media_march: n jittered samples over the box chord, in-scatter
sigma_s * radiance / (4 pi) weighted by exp(-tau) before each step).

A box medium between heights 1 and 2 over a Lambert floor, seen by a nearly
orthographic camera looking straight down from height 3 (every ray crosses
the box along a chord of 1 within 1e-5), lit by a sun shining straight
down.  The sun's radiance is the same at every sample, so with
S = sigma_s * radiance / (4 pi), dt = 1 / n:

    pixel = floor * exp(-sigma_t) + S dt (1 - exp(-sigma_t)) / (1 - exp(-sigma_t dt))

(S for sigma_t = 0).  Jitter drops out.  A transmittance taken after the
step, a missing phase factor or dt, or an attenuation that skips the
surface colour fails it.
"""
import math
import os

import numpy as np
import pytest

from sheen_amd import contract as K
from util import PyScene, f3, material

W, H, AA = 24, 16, 2
ALBEDO, AMB, RAD = (0.8, 0.6, 0.4), 0.1, 2.0


def _scene(sigma_t=None, sigma_s=(0.3, 0.5, 0.7), samples=16, sun_on=True):
    ps = PyScene()
    ps.set("materials", K.Material, [material(color=ALBEDO, spec=0.0, amb=AMB)], "num_materials")
    t = K.Triangle()
    a = 50.0
    t.v0, t.v1, t.v2 = f3(-a, 0, a), f3(a, 0, a), f3(-a, 0, -3 * a)
    t.n0 = t.n1 = t.n2 = f3(0, 0, 0)
    t.material_id = 0
    ps.set("triangles", K.Triangle, [t], "num_triangles")
    cam = ps.s.camera
    cam.position, cam.forward, cam.right, cam.up = f3(0, 3, 0), f3(0, -1, 0), f3(1, 0, 0), f3(0, 0, -1)
    cam.focal_length, cam.sensor_width = 5000.0, 50.0
    L = K.Light()
    L.type, L.direction, L.color, L.intensity = 0, f3(0, -1, 0), f3(1, 1, 1), RAD if sun_on else 0.0
    L.angle, L.flags, L.texture_id = 0.0, 1, -1
    ps.set("lights", K.Light, [L], "num_lights")
    if sigma_t is not None:
        m = K.Medium()
        m.center, m.half, m.sigma_s = f3(0, 1.5, 0), f3(10, 0.5, 10), f3(*sigma_s)
        m.sigma_t, m.samples = float(sigma_t), int(samples)
        ps.set("media", K.Medium, [m], "num_media")
    return ps


def _oracle(ps):
    from oracle import Oracle
    out = Oracle(ps.ptr, threads=min(8, os.cpu_count() or 1)).render(W, H, AA, hits=True)
    assert (out["kind"] >= 0).all()
    return out["image"][..., :3].astype(np.float64)


@pytest.mark.parametrize("sigma_t,samples", [(0.8, 16), (2.5, 7), (0.0, 5)])
def test_media_closed_form(sigma_t, samples):
    floor = _oracle(_scene())
    np.testing.assert_allclose(floor, np.broadcast_to(np.array(ALBEDO) * (AMB + RAD), floor.shape), rtol=1e-5)
    img = _oracle(_scene(sigma_t=sigma_t, samples=samples))
    S = np.array((0.3, 0.5, 0.7)) * RAD / (4.0 * math.pi)
    dt = 1.0 / samples
    series = dt * (1 - math.exp(-sigma_t)) / (1 - math.exp(-sigma_t * dt)) if sigma_t > 0 else 1.0
    np.testing.assert_allclose(img, floor * math.exp(-sigma_t) + S * series, rtol=1e-4)
    # no scattering: transmittance only
    np.testing.assert_allclose(_oracle(_scene(sigma_t=sigma_t, sigma_s=(0, 0, 0), samples=samples)),
                               floor * math.exp(-sigma_t), rtol=1e-4)


@pytest.mark.gpu
def test_media_gpu_matches_oracle():
    from sheen_amd import sheen
    ps = _scene(sigma_t=0.8, samples=16)
    with sheen.Context(ps.ptr) as ctx:
        img, _ = ctx.render(W, H, AA)
    np.testing.assert_allclose(img[..., :3], _oracle(ps), rtol=1e-4, atol=1e-6)
