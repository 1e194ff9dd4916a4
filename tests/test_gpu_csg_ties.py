"""CSG unions whose subtrees return spans with equal entries: the hit takes
the material of the first such span in the order a full in-order evaluation
merges them (stable sort), which the GPU reproduces by subtree rank while it
visits the root's union spine near-first.  Identical spheres with different
materials sit in different subtrees (leaves, intersections, differences,
nested unions, a superellipsoid) in several arrangements; radiance and
primary hits must equal the oracle's full-tree evaluation."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from util import PyScene, f3, material  # noqa: E402

from sheen_amd import contract as K  # noqa: E402

pytestmark = pytest.mark.gpu

COLORS = [(0.9, 0.1, 0.1), (0.1, 0.9, 0.1), (0.1, 0.1, 0.9), (0.9, 0.9, 0.1), (0.1, 0.9, 0.9), (0.9, 0.1, 0.9)]


def _node(op=0, kind=-1, index=-1, left=-1, right=-1):
    n = K.CSGNode()
    n.op, n.leaf_kind, n.leaf_index, n.left, n.right = op, kind, index, left, right
    return n


def _sphere(c, r, m):
    o = K.Object()
    o.type, o.material_id, o.param1, o.param2 = 1, m, f3(*c), r  # CSG operand only
    return o


def _scene(arrangement):
    ps = PyScene()
    ps.set("materials", K.Material, [material(color=c, spec=0.2, shin=16.0) for c in COLORS], "num_materials")
    cam = ps.s.camera
    cam.position, cam.forward, cam.right, cam.up = f3(0, 0, 6), f3(0, 0, -1), f3(1, 0, 0), f3(0, 1, 0)
    cam.focal_length, cam.sensor_width = 50.0, 36.0
    L = K.Light()
    L.type, L.position, L.color, L.intensity, L.radius, L.flags, L.texture_id = 1, f3(3, 4, 5), f3(1, 1, 1), 60.0, 0.0, 1, -1
    ps.set("lights", K.Light, [L], "num_lights")
    # identical spheres at the origin with materials 0..3, a bigger one to cut, a small one to subtract
    objs = [_sphere((0, 0, 0), 1.0, m) for m in range(4)] + [_sphere((0, 0, 0), 1.5, 4), _sphere((0.6, 0.6, 1.2), 0.5, 5)]
    ps.set("objects", K.Object, objs, "num_objects")
    se = K.Superellipsoid()
    se.center, se.a, se.b, se.c, se.n, se.material_id = f3(0, 0, 0), 1.0, 1.0, 1.0, 2.0, 5
    ps.set("superellipsoids", K.Superellipsoid, [se], "num_superellipsoids")
    leaf = lambda i: _node(kind=0, index=i)  # noqa: E731
    if arrangement == 0:
        # union(union(leaf0, leaf1), union(leaf2, leaf3))
        nodes = [_node(0, left=1, right=2), _node(0, left=3, right=4), _node(0, left=5, right=6),
                 leaf(0), leaf(1), leaf(2), leaf(3)]
    elif arrangement == 1:
        # union(intersection(big, leaf2), union(difference(leaf1, small), leaf0))
        nodes = [_node(0, left=1, right=2), _node(1, left=3, right=4), _node(0, left=5, right=6),
                 leaf(4), leaf(2), _node(2, left=7, right=8), leaf(0), leaf(1), leaf(5)]
    elif arrangement == 2:
        # union(union(se, leaf3), intersection(leaf1, big)) with the SE sharing the sphere's surface
        nodes = [_node(0, left=1, right=2), _node(0, left=3, right=4), _node(1, left=5, right=6),
                 _node(kind=1, index=0), leaf(3), leaf(1), leaf(4)]
    else:
        # a right-leaning chain: union(leaf3, union(leaf2, union(leaf1, leaf0)))
        nodes = [_node(0, left=1, right=2), leaf(3), _node(0, left=3, right=4), leaf(2), _node(0, left=5, right=6),
                 leaf(1), leaf(0)]
    ps.set("csg_nodes", K.CSGNode, nodes, "num_csg_nodes")
    ps.s.csg_root = 0
    return ps


@pytest.mark.parametrize("arrangement", [0, 1, 2, 3])
def test_equal_entries_take_the_first_subtree(arrangement):
    from oracle import Oracle
    from sheen_amd import sheen
    ps = _scene(arrangement)
    W, H, AA = 96, 72, 2
    with sheen.Context(ps.ptr) as ctx:
        img, _ = ctx.render(W, H, AA)
        kind, idx, t = ctx.primary_hits(W, H, AA)
    ref = Oracle(ps.ptr, threads=min(16, os.cpu_count() or 1)).render(W, H, AA, hits=True)
    assert np.array_equal(kind, ref["kind"]) and np.array_equal(t.view(np.uint32), ref["t"].view(np.uint32))
    assert int(np.sum(ref["kind"] == 7)) > 1000
    rel = np.abs(img[..., :3] - ref["image"][..., :3]) / np.maximum(np.abs(ref["image"][..., :3]), 1e-3)
    assert float(np.mean(rel.max(-1) > 1e-4)) <= 0.002, rel.max()
