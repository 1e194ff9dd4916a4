"""Lathes (reference primitives/lathe.h:14-95): surfaces of revolution about
the y axis as chains of cone frusta, the centre ignored by the hit test and
the bounds +-1e30 in x/z, so every ray whose y-slab passes tests every
lathe.  The HIP traversal skips a frustum whose axial coordinate stays
outside the segment over [t_min, running best]; primary hits (kind, index,
distance) must still equal the oracle's bit for bit and radiance within
1e-4, for lathes that overlap, widen and narrow, touch the axis, carry a
negative-radius point, and sit at a centre off the origin, seen from above,
from the side and at grazing angles."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from util import PyScene, f3, material  # noqa: E402

from sheen_amd import contract as K  # noqa: E402

# lathes beyond this count are traversed through their tree; up to it they
# are scanned in one fixed order (dev_trace.h LATHE_SCAN_MAX)
LATHE_SCAN_MAX = 64

pytestmark = pytest.mark.gpu

CAMERAS = [  # position, forward, up
    ((0.0, 1.5, 9.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0)),
    ((0.0, 12.0, 0.5), (0.0, -0.99875, -0.0499376), (0.0, 0.0499376, -0.99875)),
    ((7.0, 0.3, 7.0), (-0.7071068, 0.0, -0.7071068), (0.0, 1.0, 0.0)),
]


def _scene(cam, n_lathes=8, seed=5, plane_at_top=False):
    rng = np.random.default_rng(seed)
    ps = PyScene()
    ps.set("materials", K.Material, [material(color=(0.8, 0.7, 0.6), spec=0.0), material(color=(0.3, 0.6, 0.9))],
           "num_materials")
    pos, fwd, up = cam
    c = ps.s.camera
    right = np.cross(np.array(fwd), np.array(up))
    c.position, c.forward, c.up, c.right = f3(*pos), f3(*fwd), f3(*up), f3(*right)
    c.focal_length, c.sensor_width = 28.0, 36.0
    L = K.Light()
    L.type, L.position, L.color, L.intensity, L.radius, L.flags, L.texture_id = 1, f3(3, 9, 4), f3(1, 1, 1), 300.0, \
        0.0, 1, -1
    ps.set("lights", K.Light, [L], "num_lights")
    p = K.Plane()
    p.normal, p.d, p.material_id = f3(0, 1, 0), 2.0, 0
    ps.set("planes", K.Plane, [p], "num_planes")
    pts, lathes = [], []
    for k in range(n_lathes):
        la = K.Lathe()
        la.center = f3(*(rng.uniform(-2, 2, 3) if k % 3 else (0.0, 0.0, 0.0)))
        la.point_offset, la.material_id = len(pts), k % 2
        n = int(rng.integers(6, 34))
        y = np.sort(rng.uniform(-2.0, 4.0, n))
        rad = np.abs(np.cumsum(rng.normal(0, 0.3, n))) + 0.2 * (k % 8)
        if k == 3:
            rad[n // 2] = -0.4  # a negative profile radius
        if k == 5:
            rad[0] = 0.0  # closed at the bottom
        for yy, rr in zip(y, rad):
            lp = K.LathePoint()
            lp.y, lp.radius = float(yy), float(rr)
            pts.append(lp)
        la.point_count = n
        lathes.append(la)
    ps.set("lathe_points", K.LathePoint, pts, "num_lathe_points")
    ps.set("lathes", K.Lathe, lathes, "num_lathes")
    if plane_at_top:
        # the floor at the exact top of the first lathe's profile: segments
        # that end on it stop at that height (the lathe's y range bound)
        p.d = -float(np.float32(pts[lathes[0].point_count - 1].y))
        ps.set("planes", K.Plane, [p], "num_planes")
    return ps


@pytest.mark.parametrize("cam,n_lathes,top", [(c, 8, False) for c in range(len(CAMERAS))] +
                         [(0, LATHE_SCAN_MAX, False), (2, LATHE_SCAN_MAX + 8, False), (0, 8, True), (2, 8, True)])
def test_lathe_hits_match_oracle(cam, n_lathes, top):
    """Scanned (up to LATHE_SCAN_MAX lathes) and tree-traversed lathes, and a
    floor at the top height of a lathe's profile."""
    from oracle import Oracle
    from sheen_amd import sheen
    ps = _scene(CAMERAS[cam], n_lathes, plane_at_top=top)
    W, H, AA = 160, 120, 2
    with sheen.Context(ps.ptr) as ctx:
        img, _ = ctx.render(W, H, AA)
        kind, idx, t = ctx.primary_hits(W, H, AA)
    ref = Oracle(ps.ptr, threads=min(16, os.cpu_count() or 1)).render(W, H, AA, hits=True)
    assert int(np.sum(ref["kind"] == 5)) > 2000  # lathe hits
    assert np.array_equal(kind, ref["kind"]) and np.array_equal(idx, ref["index"])
    assert np.array_equal(t.view(np.uint32), ref["t"].view(np.uint32))
    rel = np.abs(img[..., :3] - ref["image"][..., :3]) / np.maximum(np.abs(ref["image"][..., :3]), 1e-3)
    assert float(np.mean(rel.max(-1) > 1e-4)) <= 0.001, rel.max()


def test_lathe_scenes_batched_equal_their_own_renders():
    """Three lathe scenes in one batch: the shadow and secondary queues mix
    their rays in a wave, where the lathes are traversed through each scene's
    tree; each batched image equals the scene's own render (scanned lathes)."""
    from sheen_amd.batch import Batch
    sc = [_scene(CAMERAS[0], 8, seed=5), _scene(CAMERAS[2], 12, seed=6), _scene(CAMERAS[1], 5, seed=7)]
    sizes = [(72, 56, 2), (64, 48, 3), (40, 40, 2)]
    with Batch([s.ptr for s in sc], sizes) as b:
        b.render()
        b.render_serial()
        for k in range(b.n):
            assert np.array_equal(b.image(k), b.image(k, serial=True)), k
