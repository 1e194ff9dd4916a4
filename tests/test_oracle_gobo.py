"""Spot-light gobos pinned by closed forms (reference render_kernel.cu, the
spot branch of compute_lighting: shading.h spot cone and gobo lookup).

A spot at (0, 2, 0) aims straight down (axis -y) over a Lambert ground quad
seen by a camera looking down from (0, 3, 0).  For that axis the gobo frame
is t1 = +z, t2 = -x, and a floor point p is looked up at
u = 1/2 + (ts . t1) / (2 tan(angle)), ts = normalize(p - spot).  The gobo is
8 texels wide: black in columns 0-3 and white in 4-7 (or the mirror image);
the bilinear lookup (shading.h:26-58) reads columns floor(7 u) and the next,
so rows of the image whose whole footprint maps to u >= 4/7 see only white
texels and rows with u < 3/7 only black ones.  Those rows must equal the
render without a gobo and the render with the spot switched off
respectively; a uniform grey gobo scales the direct
term by exactly 128/255.  A swapped (u, v), a flipped frame axis or a gobo
that does not multiply the radiance fails one of these.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from sheen_amd import contract as K
from util import PyScene, f3, material

W, H, AA = 48, 32, 2
ANGLE = 0.7
SPOT_Y, CAM_Y = 2.0, 3.0
FOCAL, SENSOR = 50.0, 36.0


def _scene(gobo=None, intensity=60.0):
    ps = PyScene()
    ps.set("materials", K.Material, [material(color=(0.8, 0.6, 0.4), spec=0.0, amb=0.1)], "num_materials")
    if gobo is not None:
        px = (C.c_ubyte * gobo.size)(*gobo.reshape(-1).tolist())
        ps.keep.append(px)
        ps.s.texture_pixels = ps._arr(C.POINTER(C.c_ubyte), [C.cast(px, C.POINTER(C.c_ubyte))])
        ps.s.texture_widths = ps._arr(C.c_int, [gobo.shape[1]])
        ps.s.texture_heights = ps._arr(C.c_int, [gobo.shape[0]])
        ps.s.num_textures = 1
    a = 50.0
    tris = []
    for v0, v1, v2 in (((-a, 0, -a), (a, 0, a), (a, 0, -a)), ((-a, 0, -a), (-a, 0, a), (a, 0, a))):
        t = K.Triangle()
        t.v0, t.v1, t.v2 = f3(*v0), f3(*v1), f3(*v2)
        t.n0 = t.n1 = t.n2 = f3(0, 0, 0)
        t.material_id = 0
        tris.append(t)
    ps.set("triangles", K.Triangle, tris, "num_triangles")
    cam = ps.s.camera
    cam.position, cam.forward, cam.right, cam.up = f3(0, CAM_Y, 0), f3(0, -1, 0), f3(1, 0, 0), f3(0, 0, -1)
    cam.focal_length, cam.sensor_width = FOCAL, SENSOR
    L = K.Light()
    L.type, L.position, L.direction, L.color, L.intensity = 3, f3(0, SPOT_Y, 0), f3(0, -1, 0), f3(1, 1, 1), \
        float(intensity)
    L.angle, L.blend, L.radius, L.flags = ANGLE, 0.0, 0.0, 1
    L.texture_id = 0 if gobo is not None else -1
    ps.set("lights", K.Light, [L], "num_lights")
    return ps


def _gobo(white_cols):
    g = np.zeros((2, 8, 3), np.uint8)
    g[:, white_cols, :] = 255
    return g


def _row_u_range():
    """Bounds of the gobo u over each image row's footprint (pixel corners,
    the whole image width)."""
    tan_h = 0.5 * SENSOR / FOCAL
    tan_x, tan_y = tan_h, tan_h * H / W
    out = []
    for y in range(H):
        us = []
        for yy in (y, y + 1):
            v = 1.0 - yy / H * 2.0
            z = -v * tan_y * CAM_Y  # camera up = -z
            for xx in (0, W):
                x = (xx / W * 2.0 - 1.0) * tan_x * CAM_Y
                ts = np.array([x, -SPOT_Y, z]) / math.sqrt(x * x + SPOT_Y ** 2 + z * z)
                us.append(0.5 + 0.5 * ts[2] / math.tan(ANGLE))
        out.append((min(us), max(us)))
    return out


def _oracle(ps):
    from oracle import Oracle
    out = Oracle(ps.ptr, threads=min(8, os.cpu_count() or 1)).render(W, H, AA, hits=True)
    assert (out["kind"] >= 0).all()
    return out["image"][..., :3].astype(np.float64)


def test_gobo_closed_forms():
    spot = _oracle(_scene())
    dark = _oracle(_scene(intensity=0.0))
    assert (spot > dark + 1e-3).all()  # the whole view lies inside the cone
    rng = _row_u_range()
    white_rows = [y for y, (lo, hi) in enumerate(rng) if lo >= 4 / 7 + 1e-3 and hi < 1.0]
    black_rows = [y for y, (lo, hi) in enumerate(rng) if hi <= 3 / 7 - 1e-3 and lo > 0.0]
    assert len(white_rows) >= 4 and len(black_rows) >= 4
    right = _oracle(_scene(_gobo([4, 5, 6, 7])))
    left = _oracle(_scene(_gobo([0, 1, 2, 3])))
    for img, lit, unlit in ((right, white_rows, black_rows), (left, black_rows, white_rows)):
        np.testing.assert_allclose(img[lit], spot[lit], rtol=1e-5)
        np.testing.assert_allclose(img[unlit], dark[unlit], rtol=1e-6)
    grey = np.full((2, 8, 3), 128, np.uint8)
    np.testing.assert_allclose(_oracle(_scene(grey)), dark + (spot - dark) * (128.0 / 255.0), rtol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("cols", [[4, 5, 6, 7], [0, 2, 5]])
def test_gobo_gpu_matches_oracle(cols):
    from sheen_amd import sheen
    ps = _scene(_gobo(cols))
    with sheen.Context(ps.ptr) as ctx:
        img, _ = ctx.render(W, H, AA)
    np.testing.assert_allclose(img[..., :3], _oracle(ps), rtol=1e-4, atol=1e-6)
