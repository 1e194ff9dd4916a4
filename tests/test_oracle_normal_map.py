"""Normal maps pinned by closed forms (reference render_kernel.cu:119-136).

A ground quad of two triangles whose UVs run along +x (u) and +z (v), so
the tangent frame is tan_u = +x, tan_v = +z around the normal +y; a 1x1
normal map (8-bit texel (r, g)) perturbs the normal to
normalize(n + tan_u * (2 r/255 - 1) s + tan_v * (2 g/255 - 1) s) for
strength s.  The material is pure Lambert (no specular), lit by one sun
with no angular spread, so a pixel is base * ambient + base * radiance *
max(N.L, 0):

* texel (255, 0), s = 1: N = (1, 1, -1)/sqrt 3; a sun along N lights the
  quad exactly as an overhead sun lights it with no map (N.L = 1);
* texel (0, 255): N.L = -1/3 for that sun, so only the ambient term stays,
  as with the sun switched off;
* texel (0, 0): N.L = 1/3, a third of the full direct term;
* texel (255, 0) at s = 0.5 with the sun along (0.5, 1, -0.5): N.L = 1.

A swapped channel, a flipped sign, a dropped strength factor or a map read
as r/255 rather than 2 r/255 - 1 each fails one of these.  The GPU path is
checked against the oracle on the same scenes.
"""
import ctypes as C
import os

import numpy as np
import pytest

from sheen_amd import contract as K
from util import PyScene, f3, material

W, H, AA = 48, 32, 2
AMB, ALBEDO, RAD = 0.1, (0.8, 0.6, 0.4), 2.0


def _scene(texel=None, strength=1.0, sun=(0.0, 1.0, 0.0), intensity=RAD):
    ps = PyScene()
    m = material(color=ALBEDO, spec=0.0, amb=AMB)
    if texel is not None:
        m.normal_id, m.normal_strength = 0, float(strength)
        px = (C.c_ubyte * 3)(texel[0], texel[1], 128)
        ps.keep.append(px)
        ps.s.texture_pixels = ps._arr(C.POINTER(C.c_ubyte), [C.cast(px, C.POINTER(C.c_ubyte))])
        ps.s.texture_widths = ps._arr(C.c_int, [1])
        ps.s.texture_heights = ps._arr(C.c_int, [1])
        ps.s.num_textures = 1
    ps.set("materials", K.Material, [m], "num_materials")
    a = 50.0
    quad = [((-a, 0, -a), (a, 0, a), (a, 0, -a), (0, 0), (1, 1), (1, 0)),
            ((-a, 0, -a), (-a, 0, a), (a, 0, a), (0, 0), (0, 1), (1, 1))]
    tris = []
    for v0, v1, v2, t0, t1, t2 in quad:
        t = K.Triangle()
        t.v0, t.v1, t.v2 = f3(*v0), f3(*v1), f3(*v2)
        t.n0 = t.n1 = t.n2 = f3(0, 0, 0)
        t.uv0, t.uv1, t.uv2 = K.Float2(*t0), K.Float2(*t1), K.Float2(*t2)
        t.material_id = 0
        tris.append(t)
    ps.set("triangles", K.Triangle, tris, "num_triangles")
    cam = ps.s.camera
    cam.position, cam.forward, cam.right, cam.up = f3(0, 3, 0), f3(0, -1, 0), f3(1, 0, 0), f3(0, 0, -1)
    cam.focal_length, cam.sensor_width = 50.0, 36.0
    L = K.Light()
    d = np.asarray(sun, np.float64)
    d = d / np.linalg.norm(d)
    L.type, L.direction, L.color, L.intensity = 0, f3(*(-d)), f3(1, 1, 1), float(intensity)
    L.angle, L.flags, L.texture_id = 0.0, 1, -1
    ps.set("lights", K.Light, [L], "num_lights")
    return ps


def _oracle(ps):
    from oracle import Oracle
    out = Oracle(ps.ptr, threads=min(8, os.cpu_count() or 1)).render(W, H, AA, hits=True)
    assert (out["kind"] >= 0).all()  # every sample hits the quad
    return out["image"][..., :3].astype(np.float64)


def test_normal_map_closed_forms():
    base = np.array(ALBEDO)
    plain = _oracle(_scene())                      # overhead sun, no map
    dark = _oracle(_scene(intensity=0.0))          # ambient only
    np.testing.assert_allclose(dark, np.broadcast_to(base * AMB, dark.shape), rtol=1e-6)
    np.testing.assert_allclose(plain, np.broadcast_to(base * (AMB + RAD), plain.shape), rtol=1e-6)
    toward = _oracle(_scene(texel=(255, 0), sun=(1, 1, -1)))
    np.testing.assert_allclose(toward, plain, rtol=1e-5)
    away = _oracle(_scene(texel=(0, 255), sun=(1, 1, -1)))
    np.testing.assert_allclose(away, dark, rtol=1e-6)
    third = _oracle(_scene(texel=(0, 0), sun=(1, 1, -1)))
    np.testing.assert_allclose(third, dark + (plain - dark) / 3.0, rtol=1e-5)
    half = _oracle(_scene(texel=(255, 0), strength=0.5, sun=(0.5, 1, -0.5)))
    np.testing.assert_allclose(half, plain, rtol=1e-5)
    # the plain map read as r/255 (no 2x-1) would tilt toward (1, 1, 0)
    assert np.abs(_oracle(_scene(texel=(255, 0), sun=(1, 1, 0))) - plain).max() > 0.05


@pytest.mark.gpu
@pytest.mark.parametrize("texel,strength", [((255, 0), 1.0), ((0, 0), 1.0), ((200, 40), 0.7)])
def test_normal_map_gpu_matches_oracle(texel, strength):
    from sheen_amd import sheen
    ps = _scene(texel=texel, strength=strength, sun=(1, 1, -1))
    with sheen.Context(ps.ptr) as ctx:
        img, _ = ctx.render(W, H, AA)
    ref = _oracle(ps)
    np.testing.assert_allclose(img[..., :3], ref, rtol=1e-4, atol=1e-6)
