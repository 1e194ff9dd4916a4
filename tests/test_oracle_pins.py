"""Pins of the CPU oracle against closed forms, brute force and the
reference's recorded results (the reference holds no numeric golden images).
"""
import ctypes as C

import numpy as np
import pytest

from sheen_amd import contract as K
from tests.util import PyScene, f3, material, rand_dirs, spheres, triangles

M32 = 0xFFFFFFFF


def _unxorshift(x, s):
    y = x
    for _ in range(32 // s + 1):
        y = x ^ (y >> s)
    return y & M32


def _inverse_hash(h):
    # inverse of the documented steps of aa_hash (reference sampling/aa.h:32-37)
    x = _unxorshift(h, 16)
    x = (x * pow(0x846CA68B, -1, 1 << 32)) & M32
    x = _unxorshift(x, 15)
    x = (x * pow(0x7FEB352D, -1, 1 << 32)) & M32
    return _unxorshift(x, 16)


def test_hash_is_the_documented_bijection(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(1)
    xs = [0, 1, 2, 0xFFFFFFFF, 0x9E3779B9] + [int(v) for v in rng.integers(0, 1 << 32, 2000, dtype=np.uint64)]
    hs = [L.oracle_aa_hash(x) for x in xs]
    assert [_inverse_hash(h) for h in hs] == xs
    assert len(set(hs)) == len(set(xs))
    # aa_rand = low 24 bits / 2^24 (aa.h:39-41)
    for x, h in zip(xs[:200], hs[:200]):
        assert L.oracle_aa_rand(x) == (h & 0xFFFFFF) / 16777216.0


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 9, 16, 17, 64])
def test_sample_layout_strata(oracle_mod, n):
    """aa.h:15-71: n<=4 jitters +-0.25 around the 2x2 cell centres;
    n>4 stratifies a side x side grid (side = ceil(sqrt(n)))."""
    L = oracle_mod.lib()
    ox, oy = C.c_float(), C.c_float()
    side = 2
    while side * side < n:
        side += 1
    centres = [(0.25, 0.25), (0.75, 0.25), (0.25, 0.75), (0.75, 0.75)]
    for (x, y) in [(0, 0), (5, 3), (1023, 511), (77, 901)]:
        for s in range(n):
            L.oracle_sample_offset(s, n, x, y, C.byref(ox), C.byref(oy))
            if n <= 4:
                cx, cy = centres[min(s, 3)]
                assert cx - 0.25 <= ox.value < cx + 0.25 and cy - 0.25 <= oy.value < cy + 0.25
            else:
                cx, cy = s % side, s // side
                assert cx / side <= ox.value <= (cx + 1) / side and cy / side <= oy.value <= (cy + 1) / side


def _one_prim_scene(kind, **kw):
    sc = PyScene()
    sc.set("materials", K.Material, [material()], "num_materials")
    if kind == "sphere":
        sc.set("objects", K.Object, spheres([kw["c"]], [kw["r"]], [0]), "num_objects")
    elif kind == "tris":
        sc.set("triangles", K.Triangle, triangles(kw["verts"]), "num_triangles")
    elif kind == "plane":
        p = K.Plane()
        p.normal, p.d, p.material_id = f3(*kw["n"]), kw["d"], 0
        sc.set("planes", K.Plane, [p], "num_planes")
    return sc


def test_sphere_against_quadratic(oracle_mod):
    rng = np.random.default_rng(2)
    c, r = np.array([0.3, -0.2, 1.1]), 0.9
    sc = _one_prim_scene("sphere", c=c, r=r)
    o = oracle_mod.Oracle(sc.ptr, threads=1)
    orig = (rng.uniform(-4, 4, size=(500, 3)) + np.array([0, 0, 6])).astype(np.float32)
    tgt = c + rng.uniform(-1.2, 1.2, size=(500, 3))
    d = (tgt - orig)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    res = o.trace(orig, d)
    oc = orig.astype(np.float64) - c
    b = np.sum(oc * d, 1)
    disc = b * b - (np.sum(oc * oc, 1) - r * r)
    hit = disc > 1e-6
    t = -b - np.sqrt(np.maximum(disc, 0))
    sure = (np.abs(disc) > 1e-3) & ((t > 1e-3) | ~hit)
    assert np.array_equal(res["kind"][sure] == 0, hit[sure])
    h = sure & hit
    np.testing.assert_allclose(res["t"][h], t[h], rtol=2e-5)
    n = (res["point"][h] - c) / r
    np.testing.assert_allclose(res["normal"][h], n, atol=1e-4)


def test_plane_two_sided(oracle_mod):
    sc = _one_prim_scene("plane", n=(0, 1, 0), d=2.0)  # y = -2
    o = oracle_mod.Oracle(sc.ptr, threads=1)
    orig = np.array([[0, 0, 0], [1, -5, 2]], np.float32)
    d = np.array([[0, -1, 0], [0, 1, 0]], np.float32)
    res = o.trace(orig, d)
    np.testing.assert_allclose(res["t"], [2.0, 3.0])
    np.testing.assert_allclose(res["normal"], [[0, 1, 0], [0, -1, 0]])  # flipped toward the ray


def test_triangle_against_barycentric_solve(oracle_mod):
    rng = np.random.default_rng(3)
    a, b, c = np.array([0, 0, 0.]), np.array([2, 0.2, 0.1]), np.array([0.3, 1.7, -0.2])
    sc = _one_prim_scene("tris", verts=[(a, b, c)])
    o = oracle_mod.Oracle(sc.ptr, threads=1)
    orig = (rng.uniform(-1, 3, size=(2000, 3)) + np.array([0, 0, 4])).astype(np.float32)
    tgt = rng.uniform(-0.5, 2.5, size=(2000, 3)) * np.array([1, 1, 0])
    d = (tgt - orig)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    res = o.trace(orig, d)
    for i in range(len(orig)):
        A = np.stack([b - a, c - a, -d[i].astype(np.float64)], 1)
        u, v, t = np.linalg.solve(A, orig[i] - a)
        margin = min(u, v, 1 - u - v)
        if abs(margin) < 1e-4 or t < 1e-2:
            continue
        assert (res["kind"][i] == 2) == (margin > 0), i
        if margin > 0:
            assert abs(res["t"][i] - t) <= 2e-5 * max(1.0, t)


def test_bvh_closest_equals_brute_force(oracle_mod):
    """The reference's own invariant: BVH result == linear scan (tools/bench.cu:236-248)."""
    rng = np.random.default_rng(4)
    n = 400
    cen = rng.uniform(-3, 3, size=(n, 3))
    verts = [(p, p + rng.normal(scale=0.4, size=3), p + rng.normal(scale=0.4, size=3)) for p in cen]
    sc = _one_prim_scene("tris", verts=verts)
    sc.set("objects", K.Object, spheres(rng.uniform(-3, 3, size=(60, 3)), rng.uniform(0.05, 0.4, 60), [0] * 60),
           "num_objects")
    o = oracle_mod.Oracle(sc.ptr, threads=2)
    orig = rng.uniform(-5, 5, size=(3000, 3)).astype(np.float32)
    d = rand_dirs(rng, 3000)
    res = o.trace(orig, d)
    for cat in (0, 2):
        idx, t = o.trace_brute(orig, d, cat)
        sel = res["kind"] == cat
        assert np.array_equal(res["index"][sel], idx[sel])
        assert np.array_equal(res["t"][sel], t[sel])
    # every brute-force hit of either category is at or beyond the scene's closest hit
    idx0, t0 = o.trace_brute(orig, d, 0)
    idx2, t2 = o.trace_brute(orig, d, 2)
    best = np.where(idx0 >= 0, t0, np.inf)
    best = np.minimum(best, np.where(idx2 >= 0, t2, np.inf))
    hit = res["kind"] >= 0
    assert np.array_equal(hit, np.isfinite(best))
    np.testing.assert_array_equal(res["t"][hit], best[hit].astype(np.float32))


def test_caustics_scene_matches_reference_photon_counts(oracle_mod, scenes_mod):
    """scenes/caustics.sheen: the reference stored 10,868 of 20,000 photons
    (ROADMAP.md:313, HANDOFF.md:12) and its map/brute probes agreed with an
    average of 277-282 (ROADMAP.md:324-325, HANDOFF.md:14)."""
    sc = scenes_mod.generate(103)
    o = oracle_mod.Oracle(sc.scene, threads=4)
    emitted, stored = o.photon_stats()
    assert emitted == 20000
    assert stored == 10868
    m = o.photon_probe(97, 0.125)
    b = o.photon_probe(97, 0.125, brute=True)
    assert np.array_equal(m, b)
    assert 270 <= int(m.mean()) <= 290


def test_mips_box_filter(oracle_mod):
    rng = np.random.default_rng(5)
    for (w, h) in [(8, 8), (7, 5), (1, 6), (16, 3)]:
        px = rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8)
        out = np.zeros(4 * w * h * 3, np.uint8)
        mips = C.c_int()
        n = oracle_mod.lib().oracle_build_mips(px.ctypes.data, w, h, out.ctypes.data, C.byref(mips))
        levels = [px.astype(np.int64)]
        while levels[-1].shape[0] > 1 or levels[-1].shape[1] > 1:
            p = levels[-1]
            sh, sw = p.shape[:2]
            nh, nw = max(sh // 2, 1), max(sw // 2, 1)
            ys0 = np.minimum(2 * np.arange(nh), sh - 1)
            ys1 = np.minimum(2 * np.arange(nh) + 1, sh - 1)
            xs0 = np.minimum(2 * np.arange(nw), sw - 1)
            xs1 = np.minimum(2 * np.arange(nw) + 1, sw - 1)
            s = p[ys0][:, xs0] + p[ys0][:, xs1] + p[ys1][:, xs0] + p[ys1][:, xs1]
            levels.append((s + 2) // 4)
        ref = np.concatenate([l.reshape(-1) for l in levels]).astype(np.uint8)
        assert mips.value == len(levels)
        assert n == ref.size and np.array_equal(out[:n], ref)
