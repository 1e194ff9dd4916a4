"""Superellipsoids (reference primitives/superellipsoid.h:12-161): the entry
is the first of 32 uniform samples inside the surface (the reference's
entry bisection starts with both ends at that sample).  The HIP hit test
stops at that sample, and earlier when the samples pass the traversal's
running best; primary hits (kind, index, distance) must equal the oracle's
bit for bit for overlapping superellipsoids of exponents 0.2-10 (convex
and not), flat and elongated, seen from several sides and at grazing
angles."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from util import PyScene, f3, material  # noqa: E402

from sheen_amd import contract as K  # noqa: E402

pytestmark = pytest.mark.gpu

CAMERAS = [  # position, forward, up
    ((0.0, 1.0, 10.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0)),
    ((0.0, 11.0, 0.5), (0.0, -0.99875, -0.0499376), (0.0, 0.0499376, -0.99875)),
    ((8.0, -0.2, 6.0), (-0.8, 0.0, -0.6), (0.0, 1.0, 0.0)),
]


def _scene(cam):
    rng = np.random.default_rng(9)
    ps = PyScene()
    ps.set("materials", K.Material, [material(color=(0.8, 0.7, 0.6)), material(color=(0.3, 0.6, 0.9))], "num_materials")
    pos, fwd, up = cam
    c = ps.s.camera
    right = np.cross(np.array(fwd), np.array(up))
    c.position, c.forward, c.up, c.right = f3(*pos), f3(*fwd), f3(*up), f3(*right)
    c.focal_length, c.sensor_width = 28.0, 36.0
    L = K.Light()
    L.type, L.position, L.color, L.intensity, L.radius, L.flags, L.texture_id = 1, f3(4, 9, 5), f3(1, 1, 1), 300.0, \
        0.0, 1, -1
    ps.set("lights", K.Light, [L], "num_lights")
    ses = []
    for k in range(40):
        se = K.Superellipsoid()
        se.center = f3(*rng.uniform(-3.5, 3.5, 3))
        se.a, se.b, se.c = (float(v) for v in rng.uniform(0.2, 1.6, 3) * (np.array([4.0, 0.3, 1.0]) if k % 7 == 0 else 1.0))
        se.n = float([0.2, 0.5, 1.0, 2.0, 4.0, 10.0][k % 6])
        se.material_id = k % 2
        ses.append(se)
    ps.set("superellipsoids", K.Superellipsoid, ses, "num_superellipsoids")
    return ps


@pytest.mark.parametrize("cam", range(len(CAMERAS)))
def test_superellipsoid_hits_match_oracle(cam):
    from oracle import Oracle
    from sheen_amd import sheen
    ps = _scene(CAMERAS[cam])
    W, H, AA = 160, 120, 2
    with sheen.Context(ps.ptr) as ctx:
        kind, idx, t = ctx.primary_hits(W, H, AA)
    ref = Oracle(ps.ptr, threads=min(16, os.cpu_count() or 1)).render(W, H, AA, hits=True)
    assert int(np.sum(ref["kind"] == 4)) > 3000  # superellipsoid hits
    assert np.array_equal(kind, ref["kind"]) and np.array_equal(idx, ref["index"])
    assert np.array_equal(t.view(np.uint32), ref["t"].view(np.uint32))
