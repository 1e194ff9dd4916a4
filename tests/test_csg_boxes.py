"""CSG node boxes used to skip subtrees a ray's line cannot meet: each box
must contain every span its subtree can return (union: union of the
children's boxes; intersection and difference: the left operand's box;
missing children give nothing, cycles are not bounded), and the CSG root's
union spine: the subtrees below the unions reachable from the root through
unions, in in-order rank with their frame depth, those at depth 16 or more
dropped (the evaluation stack has 16 frames), and a hierarchy over them."""
import ctypes as C
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(__file__))
from util import PyScene, material, spheres  # noqa: E402

from sheen_amd import contract as K  # noqa: E402


def _node(op=0, kind=-1, index=-1, left=-1, right=-1):
    n = K.CSGNode()
    n.op, n.leaf_kind, n.leaf_index, n.left, n.right = op, kind, index, left, right
    return n


def _boxes(nodes, objs, ses=()):
    from sheen_amd import native
    ps = PyScene()
    ps.set("materials", K.Material, [material()], "num_materials")
    ps.set("objects", K.Object, objs, "num_objects")
    if ses:
        ps.set("superellipsoids", K.Superellipsoid, list(ses), "num_superellipsoids")
    ps.set("csg_nodes", K.CSGNode, nodes, "num_csg_nodes")
    ps.s.csg_root = 0
    out = np.zeros((len(nodes), 6), np.float32)
    assert native.host().sheen_host_csg_boxes(ps.ptr, out.ctypes.data_as(C.c_void_p)) == len(nodes)
    return out


def test_box_rules():
    objs = spheres([(0, 0, 0), (1.5, 0, 0)], [1.0, -0.5], [0, 0])
    se = K.Superellipsoid()
    se.center, se.a, se.b, se.c, se.n = K.Float3(0, 3, 0), 0.5, -0.25, 2.0, 2.0
    leaves = [_node(kind=0, index=0), _node(kind=0, index=1), _node(kind=1, index=0)]
    # 0: union(1, 2), 1: intersection(3, 4), 2: difference(5, 3), 3..5 leaves
    nodes = [_node(0, left=1, right=2), _node(1, left=3, right=4), _node(2, left=5, right=3)] + leaves
    b = _boxes(nodes, objs, [se])
    s0 = [-1, -1, -1, 1, 1, 1]
    s1 = [1, -0.5, -0.5, 2, 0.5, 0.5]  # |radius|
    e0 = [-0.5, 2.75, -2, 0.5, 3.25, 2]  # |half extents|
    np.testing.assert_array_equal(b[3], s0)
    np.testing.assert_array_equal(b[4], s1)
    np.testing.assert_array_equal(b[5], e0)
    np.testing.assert_array_equal(b[1], s0)  # intersection: the left operand
    np.testing.assert_array_equal(b[2], e0)  # difference: the left operand
    np.testing.assert_array_equal(b[0], [-1, -1, -2, 1, 3.25, 2])  # union of 1 and 2


def test_missing_children_cycles_and_unknown_ops():
    objs = spheres([(0, 0, 0)], [1.0], [0])
    inf = np.float32(np.inf)
    # 0: union(1, missing) -> leaf box; 1: leaf; 2: intersection(missing, 1) -> empty
    b = _boxes([_node(0, left=1, right=-1), _node(kind=0, index=0), _node(1, left=-1, right=1), _node(7, left=1, right=1)],
               objs)
    np.testing.assert_array_equal(b[0], [-1, -1, -1, 1, 1, 1])
    assert np.all(b[2][:3] > b[2][3:])  # empty
    assert np.all(b[3][:3] == inf) and np.all(b[3][3:] == -inf)  # unknown op: no spans
    # a cycle 0 -> 1 -> 0: unbounded (never culled)
    b = _boxes([_node(0, left=1, right=2), _node(0, left=0, right=2), _node(kind=0, index=0)], objs)
    assert np.all(b[0][:3] == -inf) and np.all(b[0][3:] == inf)


def test_analytic_scene_boxes_follow_the_rules():
    from sheen_amd import native, scenes
    sc = scenes.generate(3)
    c = sc.scene.contents
    n = c.num_csg_nodes
    out = np.zeros((n, 6), np.float32)
    assert native.host().sheen_host_csg_boxes(sc.scene, out.ctypes.data_as(C.c_void_p)) == n
    for i in range(n):
        nd = c.csg_nodes[i]
        if nd.leaf_kind == 0:
            o = c.objects[nd.leaf_index]
            ctr = np.array([o.param1.x, o.param1.y, o.param1.z], np.float32)
            np.testing.assert_array_equal(out[i], np.concatenate([ctr - abs(o.param2), ctr + abs(o.param2)]))
            continue
        a, b = out[nd.left], out[nd.right]
        if nd.op == 0:
            want = np.concatenate([np.minimum(a[:3], b[:3]), np.maximum(a[3:], b[3:])])
        else:
            want = a
        np.testing.assert_array_equal(out[i], want)


def _spine(scene):
    from sheen_amd import native
    info = (C.c_int * 4)()
    assert native.host().sheen_host_csg_spine(scene, info, None, None) == 0
    sub = np.zeros((info[1], 4), np.int32)
    nodes = np.zeros((info[3], 8), np.int32)
    assert native.host().sheen_host_csg_spine(scene, info, sub.ctypes.data_as(C.c_void_p),
                                              nodes.ctypes.data_as(C.c_void_p)) == 0
    return info[0], sub, info[2], nodes


def test_spine_ranks_depths_and_truncation():
    objs = spheres([(i, 0, 0) for i in range(20)], [0.4] * 20, [0] * 20)
    # a left-leaning union chain 20 deep: union_k = union(union_{k+1}, leaf_k)
    nodes = []
    for k in range(19):
        nodes.append(_node(0, left=k + 1, right=19 + k))
    nodes[-1].left = 19 + 19  # the deepest union's left child: the last leaf
    leaves = [_node(kind=0, index=k) for k in range(20)]
    nodes = nodes + leaves  # unions 0..18, leaves 19..38
    nodes[18].left = 38
    ps = PyScene()
    ps.set("materials", K.Material, [material()], "num_materials")
    ps.set("objects", K.Object, objs, "num_objects")
    ps.set("csg_nodes", K.CSGNode, nodes, "num_csg_nodes")
    ps.s.csg_root = 0
    spine, sub, always, hier = _spine(ps.ptr)
    assert spine == 1 and always == 0
    by_rank = sorted(sub.tolist(), key=lambda x: x[1])
    # in-order: the deepest subtrees first; union k sits at depth k, its leaf at depth k + 1;
    # subtrees at depth >= 16 are never evaluated
    assert all(d < 16 for _, _, d, _ in by_rank)
    assert [n for n, _, _, _ in by_rank] == [19 + k for k in range(14, -1, -1)]
    assert [d for _, _, d, _ in by_rank] == [k + 1 for k in range(14, -1, -1)]
    assert [r for _, r, _, _ in by_rank] == list(range(len(by_rank)))
    # every subtree is reached exactly once through the hierarchy leaves
    seen = []
    for nd in hier:
        a, b = int(nd[3]), int(nd[7])
        if a < 0:
            seen += list(range(-a - 1, -a - 1 + b))
    assert sorted(seen) == list(range(len(sub)))


def test_spine_of_the_analytic_scene():
    from sheen_amd import scenes
    sc = scenes.generate(3)
    spine, sub, always, hier = _spine(sc.scene)
    assert spine == 1 and always == 0
    # unions inside the 624 eight-sphere trees belong to the spine as well
    assert len(sub) >= 624
    assert sorted(sub[:, 1].tolist()) == list(range(len(sub)))
    c = sc.scene.contents
    assert all(c.csg_nodes[int(n)].leaf_kind >= 0 or c.csg_nodes[int(n)].op != 0 for n in sub[:, 0])
