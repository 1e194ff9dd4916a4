"""Optical depth of the strands along shadow rays (reference strand_bvh_tau,
bvh.h:928-960): the HIP path finds the strands through the strand BVH and
sums the nonzero depths in strand index order, falling back to the sum over
every strand when a ray meets more than a few; the oracle sums over every
strand in index order.  A sparse field of strands (rays meeting 0-3 of
them) next to a dense bundle (rays meeting many) between a point light and
a ground plane; Lambert shading only, so radiance must equal the oracle's
within 1e-4 on every pixel and primary hits bit for bit."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from util import PyScene, f3, material  # noqa: E402

from sheen_amd import contract as K  # noqa: E402

pytestmark = pytest.mark.gpu


def _scene(seed):
    rng = np.random.default_rng(seed)
    ps = PyScene()
    ps.set("materials", K.Material, [material(color=(0.8, 0.8, 0.8)), material(color=(0.7, 0.5, 0.3))],
           "num_materials")
    cam = ps.s.camera
    cam.position, cam.forward, cam.right, cam.up = f3(0, 6, 6), f3(0, -0.7071068, -0.7071068), f3(1, 0, 0), \
        f3(0, 0.7071068, -0.7071068)
    cam.focal_length, cam.sensor_width = 30.0, 36.0
    L = K.Light()
    L.type, L.position, L.color, L.intensity, L.radius, L.flags, L.texture_id = 1, f3(0.3, 8, 0.2), f3(1, 1, 1), \
        400.0, 0.0, 1, -1
    ps.set("lights", K.Light, [L], "num_lights")
    p = K.Plane()
    p.normal, p.d, p.material_id = f3(0, 1, 0), 0.0, 0
    ps.set("planes", K.Plane, [p], "num_planes")
    pts, strands = [], []

    def strand(c, spread, n):
        s = K.SweepStrand()
        s.point_offset, s.point_count, s.material_id = len(pts), n, 1
        x = c + rng.uniform(-spread, spread, 3)
        for _ in range(n):
            sp = K.SweepPoint()
            sp.position, sp.radius = f3(*x), float(rng.uniform(0.02, 0.06))
            pts.append(sp)
            x = x + rng.normal(0, 0.25, 3)
        strands.append(s)

    for _ in range(300):  # sparse field over the plane
        strand(np.array([0.0, 2.5, 0.0]), np.array([3.0, 1.5, 3.0]), int(rng.integers(3, 6)))
    for _ in range(150):  # dense bundle under the light
        strand(np.array([0.0, 3.0, 0.0]), np.array([0.6, 1.0, 0.6]), int(rng.integers(3, 6)))
    order = rng.permutation(len(strands))  # index order unrelated to position
    strands = [strands[i] for i in order]
    ps.set("sweep_points", K.SweepPoint, pts, "num_sweep_points")
    ps.set("sweep_strands", K.SweepStrand, strands, "num_sweep_strands")
    return ps


@pytest.mark.parametrize("seed", [1, 2])
def test_strand_optical_depth_matches_oracle(seed):
    from oracle import Oracle
    from sheen_amd import sheen
    ps = _scene(seed)
    W, H, AA = 128, 96, 2
    with sheen.Context(ps.ptr) as ctx:
        img, _ = ctx.render(W, H, AA)
        kind, idx, t = ctx.primary_hits(W, H, AA)
    ref = Oracle(ps.ptr, threads=min(16, os.cpu_count() or 1)).render(W, H, AA, hits=True)
    assert np.array_equal(kind, ref["kind"]) and np.array_equal(idx, ref["index"])
    assert np.array_equal(t.view(np.uint32), ref["t"].view(np.uint32))
    a, b = img[..., :3], ref["image"][..., :3]
    # plane pixels under the bundle are in deep strand shadow, others partly or not at all
    plane = b.mean(-1)[ref["kind"][..., 0] == 8]
    assert plane.size > 1000 and plane.min() < 0.25 * plane.max()
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
    assert rel.max() <= 1e-4, (rel.max(), np.argwhere(rel.max(-1) > 1e-4)[:5])
