"""The .sheen scene loader (reference api/scene_io.cpp, scene.h builders):
directive semantics, host tessellation (Catmull-Rom sweeps, cubic Bezier
profiles, cube UVs) against closed forms, the file loaders, and the
reference's own scene files where the reference tree is present."""
import ctypes as C
import glob
import os
import re
import struct

import numpy as np
import pytest

REF_SCENES = "/root/reference/scenes"


def _load(scenes_mod, tmp_path, text, name="s.sheen"):
    p = tmp_path / name
    p.write_text(text)
    return scenes_mod.load_sheen(p)


def _pts(sc, strand):
    c = sc.scene.contents
    s = c.sweep_strands[strand]
    return np.array([[c.sweep_points[s.point_offset + k].position.x, c.sweep_points[s.point_offset + k].position.y,
                      c.sweep_points[s.point_offset + k].position.z, c.sweep_points[s.point_offset + k].radius]
                     for k in range(s.point_count)], np.float64)


def test_catmull_rom_sweep_passes_through_its_points(scenes_mod, tmp_path):
    ctrl = np.array([[0, 0, 0, 0.1], [1, 0.5, 0, 0.2], [2, 0, 1, 0.15], [3, 1, 1, 0.05]], np.float64)
    sub = 5
    sc = _load(scenes_mod, tmp_path, "material 1 1 1 0 8 0.1\nsweep_smooth 0 %d %s\nsweep 0 0 %s\n"
               % (sub, " ".join("%g" % v for v in ctrl.ravel()), " ".join("%g" % v for v in ctrl.ravel())))
    pts = _pts(sc, 0)
    assert len(pts) == (len(ctrl) - 1) * sub + 1
    np.testing.assert_allclose(pts[::sub], ctrl, atol=1e-6)
    # interior point: uniform Catmull-Rom with mirrored phantom end points
    ph0 = 2 * ctrl[0] - ctrl[1]
    p = np.vstack([ph0, ctrl, 2 * ctrl[-1] - ctrl[-2]])
    u = 2.0 / sub
    i = 1  # segment from ctrl[0] to ctrl[1]
    q = 0.5 * (2 * p[i] + (p[i + 1] - p[i - 1]) * u + (2 * p[i - 1] - 5 * p[i] + 4 * p[i + 1] - p[i + 2]) * u * u
               + (-p[i - 1] + 3 * p[i] - 3 * p[i + 1] + p[i + 2]) * u ** 3)
    np.testing.assert_allclose(pts[2], q, atol=1e-6)
    # a linear sweep keeps its points
    np.testing.assert_allclose(_pts(sc, 1), ctrl, atol=1e-7)


def _bezier(ctrl, sub):
    out = [ctrl[0]]
    for s in range(0, len(ctrl) // 4 * 4, 4):
        for k in range(1, sub + 1):
            t = k / sub
            u = 1 - t
            out.append(u ** 3 * ctrl[s] + 3 * u * u * t * ctrl[s + 1] + 3 * u * t * t * ctrl[s + 2] + t ** 3 * ctrl[s + 3])
    return np.array(out)


def test_bezier_profiles(scenes_mod, tmp_path):
    ctrl = np.array([[0, 0.5], [0.3, 0.9], [0.7, 0.2], [1, 0.4], [1.2, 0.6], [1.5, 0.1], [1.8, 0.3], [2, 0.05], [9, 9]],
                    np.float64)
    flat = " ".join("%g" % v for v in ctrl.ravel())
    sc = _load(scenes_mod, tmp_path, "material 1 1 1 0 8 0.1\nlathe_bezier 0 0 0 0 4 %d %s\n"
               "prism_bezier 0 0 0 0 1 0 0 3 %d %s\n" % (len(ctrl), flat, len(ctrl), flat))
    c = sc.scene.contents
    lt = c.lathes[0]
    lp = np.array([[c.lathe_points[lt.point_offset + k].y, c.lathe_points[lt.point_offset + k].radius]
                   for k in range(lt.point_count)])
    want = _bezier(ctrl[:8], 4)  # the ninth point is an incomplete group, dropped
    assert lp.shape == want.shape
    np.testing.assert_allclose(lp, want, atol=1e-6)
    pr = c.prisms[0]
    pp = np.array([[c.prism_points[pr.point_offset + k].x, c.prism_points[pr.point_offset + k].z]
                   for k in range(pr.point_count)])
    want = _bezier(ctrl[:8], 3)
    assert len(pp) == len(want) + 1 and np.array_equal(pp[-1], pp[0])  # contour closed explicitly
    np.testing.assert_allclose(pp[:-1], want, atol=1e-6)


def test_cube_has_per_face_uvs(scenes_mod, tmp_path):
    sc = _load(scenes_mod, tmp_path, "material 1 1 1 0 8 0.1\ncube 1 2 3 0.5 0\n")
    c = sc.scene.contents
    assert c.num_triangles == 12
    for i in range(12):
        t = c.triangles[i]
        v = np.array([[t.v0.x, t.v0.y, t.v0.z], [t.v1.x, t.v1.y, t.v1.z], [t.v2.x, t.v2.y, t.v2.z]])
        uv = np.array([[t.uv0.x, t.uv0.y], [t.uv1.x, t.uv1.y], [t.uv2.x, t.uv2.y]])
        axis = int(np.argmin(np.ptp(v, 0)))  # the face's constant axis
        assert axis == i // 4
        others = [a for a in range(3) if a != axis]
        # u follows z on x faces and x elsewhere; v follows z on y faces and y elsewhere
        ua = 2 if axis == 0 else 0
        va = 2 if axis == 1 else 1
        assert ua in others and va in others
        lo = np.array([1, 2, 3]) - 0.5
        np.testing.assert_allclose(uv[:, 0], (v[:, ua] - lo[ua]) / 1.0, atol=1e-6)
        np.testing.assert_allclose(uv[:, 1], (v[:, va] - lo[va]) / 1.0, atol=1e-6)


def test_file_loaders(scenes_mod, tmp_path):
    # OBJ: a quad (fan -> 2 triangles) with uv and normals, negative indices
    (tmp_path / "q.obj").write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nvn 0 0 1\n"
                                    "f -4/1/1 -3/2/1 -2/3/1 -1/4/1\n")
    # SMESH1: 3 vertices, 1 normal, 0 uvs, 1 triangle of material 2
    with open(tmp_path / "t.smesh", "wb") as f:
        f.write(b"SMESH1\0\0" + struct.pack("<4I", 3, 1, 0, 1))
        f.write(struct.pack("<9f", 0, 0, 0, 2, 0, 0, 0, 2, 0) + struct.pack("<3f", 0, 0, 1))
        f.write(struct.pack("<10i", 0, 1, 2, 0, 0, 0, -1, -1, -1, 2))
    # P6 texture 3x2, P5 height map 2x2
    px = bytes(range(18))
    (tmp_path / "t.ppm").write_bytes(b"P6\n3 2\n255\n" + px)
    (tmp_path / "h.pgm").write_bytes(b"P5\n2 2\n255\n" + bytes([0, 64, 128, 255]))
    sc = _load(scenes_mod, tmp_path,
               "material 1 1 1 0 8 0.1\nmaterial 1 0 0 0 8 0.1\nmaterial 0 1 0 0 8 0.1\n"
               "obj file=q.obj mat=1 scale=2 offset=1,0,0\nmesh file=t.smesh mat0=0 scale=0.5\n"
               "texture file=t.ppm\nhf_image 0 0 0 2 2 1 h.pgm 1\n")
    c = sc.scene.contents
    assert c.num_triangles == 3
    t0, t1, t2 = c.triangles[0], c.triangles[1], c.triangles[2]
    assert (t0.v1.x, t0.v2.y, t0.material_id) == (3.0, 2.0, 1) and (t0.uv2.x, t0.uv2.y) == (1.0, 1.0)
    assert (t0.n0.z, t1.v2.x, t1.v2.y) == (1.0, 1.0, 2.0)
    assert (t2.v1.x, t2.v2.y, t2.material_id, t2.n0.z) == (1.0, 1.0, 2, 1.0)
    assert c.num_textures == 1 and (c.texture_widths[0], c.texture_heights[0]) == (3, 2)
    assert bytes(c.texture_pixels[0][k] for k in range(18)) == px
    hf = c.heightfields[0]
    assert (hf.img_w, hf.img_h, hf.seed) == (2, 2, -1)
    assert [hf.image[k] for k in range(4)] == [0, 64, 128, 255]


def test_csg_prefix_tree_and_hidden_operands(scenes_mod, tmp_path):
    sc = _load(scenes_mod, tmp_path, "material 1 1 1 0 8 0.1\nsphere 0 0 0 1 0\nse 2 0 0 1 1 1 4 0\n"
               "se_csg 0 0 0 0.5 2 0.5 2 0\ncsg_begin 2\ncsg_leaf 0 0\ncsg_begin 0\ncsg_leaf 1 1\ncsg_leaf 0 0\n"
               "csg_end\ncsg_end\n")
    c = sc.scene.contents
    n = [(c.csg_nodes[i].op, c.csg_nodes[i].leaf_kind, c.csg_nodes[i].leaf_index, c.csg_nodes[i].left,
          c.csg_nodes[i].right) for i in range(c.num_csg_nodes)]
    assert c.csg_root == 0
    assert n == [(2, -1, -1, 1, 2), (0, 0, 0, -1, -1), (0, -1, -1, 3, 4), (0, 1, 1, -1, -1), (0, 0, 0, -1, -1)]
    assert [c.superellipsoids[i]._pad0 for i in range(2)] == [0, 1]
    from sheen_amd import native
    info = (C.c_int * 6)()
    assert native.host().sheen_host_bvh_check(sc.scene, 4, info) == 0
    assert info[5] == 1  # only the visible superellipsoid is in the SE BVH


def test_directive_semantics(scenes_mod, tmp_path):
    sc = _load(scenes_mod, tmp_path,
               "resolution width=320 height=200 aa=3 out=x.png\n"
               "camera loc=1,2,3 fwd=0,0,-1 lens=35 sensor=24\n"
               "material diffuse=0.5,0.6,0.7 spec=0.2 shin=20 amb=0.05 tex=0 nstr=0.5\n"
               "material 0.1 0.2 0.3 0.4 5 0.06 0.7 0.2 0.3 10 0.4 20 0.5\n"
               "reflect 0 0.25\nglass 1 0.9 1.33\nemit 1 1 0.5 0.25 3\n"
               "light 0 -2 0 1 1 1 2\nsun_angle 0.05\nfill_light 1 -1 0 1 1 1 0.5\n"
               "point 1 2 3 1 1 1 100 0.2\nspot 0 5 0 0 -2 0 1 1 1 300 1.0 0.3 0.1 0.5\n"
               "area 0 4 0 2 1 1 1 1 50\n"
               "grid 3 2 1.5 0.5 0.25 2\nphotons 123\nphoton_cell 0.5\ncaustics 0.2\n")
    assert (sc.width, sc.height, sc.aa, sc.output) == (320, 200, 3, "x.png")
    c = sc.scene.contents
    cam = c.camera
    assert (cam.position.y, cam.focal_length, cam.sensor_width, cam.forward.z) == (2.0, 35.0, 24.0, -1.0)
    m0, m1 = c.materials[0], c.materials[1]
    assert (m0.texture_id, m0.mask_id, m0.normal_strength, m0.reflect, np.float32(m0.ior)) == (0, -1, 0.5, 0.25, np.float32(1.45))
    assert (m1.anisotropy, m1.hair_backlight, np.float32(m1.transmission), np.float32(m1.ior), m1.emission) == \
        (np.float32(0.7), np.float32(0.5), np.float32(0.9), np.float32(1.33), 3.0)
    L = [c.lights[i] for i in range(c.num_lights)]
    assert [l.type for l in L] == [0, 0, 1, 3, 2]
    assert L[0].direction.y == -1.0 and np.float32(L[0].angle) == np.float32(0.05) and L[0].flags == 1 and L[1].flags == 0
    assert np.float32(L[3].angle) == np.float32(0.5) and L[3].direction.y == -1.0 and L[3].beam == 0.5
    assert (L[4].u_axis.x, L[4].v_axis.z) == (1.0, 0.5)
    assert c.num_objects == 6 and [c.objects[i].material_id for i in range(6)] == [0, 1, 0, 1, 0, 1]
    assert (c.photons, c.photon_cell, np.float32(c.caustics)) == (123, 0.5, np.float32(0.2))


def test_hair_is_deterministic_and_errors_name_the_line(scenes_mod, tmp_path):
    text = "material 0.3 0.2 0.1 0.2 16 0.1\nhair 0 1.6 0 0.55 40 8 3 0\n"
    a = _pts(_load(scenes_mod, tmp_path, text, "a.sheen"), 7)
    b = _pts(_load(scenes_mod, tmp_path, text, "b.sheen"), 7)
    assert np.array_equal(a, b) and len(a) == (8 - 1) * 2 + 1
    with pytest.raises(ValueError, match="line 2: unknown directive 'teapot'"):
        _load(scenes_mod, tmp_path, "material 1 1 1 0 8 0.1\nteapot 1 2 3\n")
    with pytest.raises(ValueError, match="line 1: texture: cannot open"):
        _load(scenes_mod, tmp_path, "texture file=nowhere.ppm\n")


@pytest.mark.skipif(not os.path.isdir(REF_SCENES), reason="the reference tree is not mounted")
def test_reference_scene_files(scenes_mod):
    """Every scene file of the reference parses (except the two whose assets
    live in /tmp), with one record per directive line."""
    for f in sorted(glob.glob(os.path.join(REF_SCENES, "*.sheen"))):
        lines = [ln.split("#")[0].split() for ln in open(f, encoding="utf-8")]
        cmds = [ln[0] for ln in lines if ln]
        try:
            sc = scenes_mod.load_sheen(f)
        except ValueError as e:
            assert "/tmp/" in str(e), (f, e)
            continue
        c = sc.scene.contents
        assert c.num_materials == cmds.count("material"), f
        assert c.num_planes == cmds.count("plane"), f
        assert c.num_lights == sum(cmds.count(k) for k in ("light", "fill_light", "point", "spot", "area")), f
        assert c.num_superellipsoids == cmds.count("se") + cmds.count("se_csg"), f
        assert c.num_csg_nodes == cmds.count("csg_begin") + cmds.count("csg_leaf"), f
        if "hair" not in cmds and "strands" not in cmds:
            assert c.num_sweep_strands == cmds.count("sweep") + cmds.count("sweep_smooth"), f


@pytest.mark.skipif(not os.path.isdir(REF_SCENES), reason="the reference tree is not mounted")
def test_reference_caustics_scene_stores_10868_photons(scenes_mod, oracle_mod):
    """caustics.sheen: 10,868 of 20,000 photons stored (ROADMAP.md:313)."""
    sc = scenes_mod.load_sheen(os.path.join(REF_SCENES, "caustics.sheen"))
    assert oracle_mod.Oracle(sc.scene, threads=4).photon_stats() == (20000, 10868)
