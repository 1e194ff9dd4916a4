"""Device BVH build (sheen_context_create) against the host builder
(SHEEN_CTX_HOST_BVH): every category's wide nodes, leaf slots, stack bound
and depth must be identical byte for byte, and frames rendered from the two
contexts identical bit for bit (the triangle leaf copies included).

Cases cover the builder's paths: a root small enough for one wavefront,
sizes just past the wavefront limit (512) and the workgroup chunk (4096),
coincident centroids (split by index), duplicated and degenerate triangles,
flat scenes (zero centroid extent on an axis), skipped primitives (CSG-only
superellipsoids, non-sphere objects), the analytic kinds with +-1e30 lathe
bounds, and the 1M-triangle mesh.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from util import PyScene, material, triangles  # noqa: E402

pytestmark = pytest.mark.gpu


def _compare(scene, render=None):
    from sheen_amd import sheen
    with sheen.Context(scene) as dev, sheen.Context(scene, host_bvh=True) as host:
        for cat in range(7):
            dn, dp, ds, dd = dev.bvh(cat)
            hn, hp, hs, hd = host.bvh(cat)
            assert dn.shape == hn.shape, (cat, dn.shape, hn.shape)
            if len(hn):
                bad = np.flatnonzero(np.any(dn != hn, axis=1))
                assert bad.size == 0, (cat, bad[:8], len(hn))
            assert np.array_equal(dp, hp), cat
            assert (ds, dd) == (hs, hd), cat
        if render:
            w, h, aa = render
            a, _ = dev.render(w, h, aa)
            b, _ = host.render(w, h, aa)
            assert np.array_equal(a, b)
        return dev.build_ms, host.build_ms


def _tri_scene(v):
    ps = PyScene()
    ps.set("materials", type(material()), [material()], "num_materials")
    ps.set("triangles", type(triangles([((0, 0, 0), (1, 0, 0), (0, 1, 0))])[0]), triangles(v), "num_triangles")
    return ps


def _soup(n, seed, flat=False):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-4, 4, (n, 3))
    if flat:
        c[:, 1] = 0.0
    e = rng.normal(scale=0.05, size=(n, 2, 3))
    if flat:
        e[:, :, 1] = 0.0
    return [(tuple(p), tuple(p + d[0]), tuple(p + d[1])) for p, d in zip(c, e)]


@pytest.mark.parametrize("n", [1, 3, 4, 5, 200, 512, 513, 4096, 4097, 20000])
def test_triangle_soup(n):
    ps = _tri_scene(_soup(n, n))
    _compare(ps.ptr, render=(64, 48, 1) if n >= 200 else None)


def test_flat_and_coincident():
    v = _soup(3000, 7, flat=True)
    v += [((1.0, 0.0, 1.0), (1.5, 0.0, 1.0), (1.0, 0.0, 1.5))] * 700  # one centroid 700 times
    v += [((2.0, 0.0, 2.0), (2.0, 0.0, 2.0), (2.0, 0.0, 2.0))] * 40  # degenerate points
    _compare(_tri_scene(v).ptr, render=(64, 48, 1))


@pytest.mark.parametrize("cfg,kw", [(1, {}), (2, {"detail": 60}), (5, {"detail": 3000}), (101, {}), (102, {}),
                                    (103, {}), (104, {}), (105, {})])
def test_generated_scenes(cfg, kw):
    from sheen_amd import scenes
    sc = scenes.generate(cfg, **kw)
    _compare(sc.scene, render=(96, 64, 1))


def test_scene_files_with_skipped_primitives():
    from sheen_amd import scenes
    for name in ("objects", "media"):
        sf = scenes.load_sheen(os.path.join(os.path.dirname(__file__), "scenes", name + ".sheen"))
        _compare(sf.scene, render=(64, 48, 1))


def test_batch_scenes():
    from sheen_amd import scenes
    for sc in scenes.batch(256)[:256:37]:
        _compare(sc.scene)


def test_million_triangle_mesh():
    from sheen_amd import scenes
    sc = scenes.generate(2)
    dev_ms, host_ms = _compare(sc.scene)
    print("1M-triangle mesh: device build %.1f ms, host build %.1f ms" % (dev_ms, host_ms))
