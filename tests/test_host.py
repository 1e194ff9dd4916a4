"""CPU tests of the product's host side: ABI layout, exported symbols,
error paths that return before any device call, wide-BVH invariants,
scene generators, sRGB/PNG output."""
import ctypes as C
import os
import re
import subprocess
import sys
import zlib

import numpy as np
import pytest

from sheen_amd import contract as K
from sheen_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def test_ctypes_layout_matches_reference_sizes():
    for name, size in K.SIZES.items():
        assert C.sizeof(getattr(K, name)) == size, name
    assert K.SheenScene.photons.offset == 316
    assert K.SheenScene.photon_cell.offset == 320
    assert K.SheenScene.caustics.offset == 324
    assert K.SheenScene.media.offset == 328
    assert K.SheenScene.num_media.offset == 336
    assert K.Material.emission_color.offset == 72
    assert K.Material.normal_strength.offset == 96


@pytest.mark.parametrize("lang,compiler", [("c", "gcc"), ("c++", "g++")])
def test_headers_static_asserts_compile(tmp_path, lang, compiler):
    src = tmp_path / ("t." + ("c" if lang == "c" else "cpp"))
    src.write_text('#include "sheen_c_api.h"\n#include "sheen_scenes.h"\nint main(void){return 0;}\n')
    std = "-std=c11" if lang == "c" else "-std=c++17"
    subprocess.run([compiler, std, "-I", INC, "-c", str(src), "-o", str(tmp_path / "t.o")], check=True)


def _declared(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(sheen_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_c_api_library_exports_every_declared_symbol():
    lib = native.capi()  # loads without touching the device
    names = _declared("sheen_c_api.h")
    assert "sheen_render_scene" in names and len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_host_library_exports_scene_symbols():
    lib = native.host()
    for n in _declared("sheen_scenes.h"):
        assert hasattr(lib, n), n


def test_reference_error_paths_without_device():
    """api/sheen_c_api.cu:189-193 validation; all return before device work."""
    lib = native.capi()
    sc = K.SheenScene()
    sc.csg_root = -1
    assert lib.sheen_render_scene(None, 8, 8, 1, b"/tmp/x.png") == -1
    assert native.last_error() == "null scene"
    for (w, h, aa, path) in [(0, 8, 1, b"/tmp/x.png"), (8, -1, 1, b"/tmp/x.png"), (8, 8, 0, b"/tmp/x.png"),
                             (8, 8, 65, b"/tmp/x.png"), (8, 8, 1, None)]:
        assert lib.sheen_render_scene(C.pointer(sc), w, h, aa, path) == -1
        assert native.last_error() == "bad render params"
    assert lib.sheen_render_scene(C.pointer(sc), 8, 8, 1, b"/tmp/x.png") == -1
    assert native.last_error() == "no materials"


def test_python_module_mirror_errors():
    from sheen_amd import sheen
    with pytest.raises(TypeError):
        sheen.render_scene(12345, 8, 8, 1, "/tmp/x.png")
    with pytest.raises(TypeError):
        sheen.render_scene(bytearray(100), 8, 8, 1, "/tmp/x.png")
    with pytest.raises(RuntimeError, match="bad render params"):
        sheen.render_scene(bytearray(352), 0, 8, 1, "/tmp/x.png")


@pytest.mark.parametrize("cfg,kw", [(1, {}), (2, {"detail": 24}), (5, {"detail": 3000}), (100, {}), (101, {}),
                                    (102, {}), (103, {})])
def test_wide_bvh_is_conservative_and_complete(scenes_mod, cfg, kw):
    sc = scenes_mod.generate(cfg, **kw)
    checked = 0
    for cat in range(7):
        info = (C.c_int * 6)()
        assert native.host().sheen_host_bvh_check(sc.scene, cat, info) == 0
        nodes, max_stack, depth, leaves, violations, refs = list(info)
        if nodes:
            checked += 1
            assert violations == 0
            assert max_stack <= 128
    assert checked >= 1


def test_generators_are_deterministic(scenes_mod):
    a = scenes_mod.generate(2, detail=16)
    b = scenes_mod.generate(2, detail=16)
    na = a.scene.contents.num_triangles
    assert na == 20 * 16 * 16
    ta = C.string_at(a.scene.contents.triangles, na * 112)
    tb = C.string_at(b.scene.contents.triangles, na * 112)
    assert ta == tb
    c = scenes_mod.generate(2, detail=16, seed=99)
    assert C.string_at(c.scene.contents.triangles, na * 112) != ta


def test_geodesic_mesh_is_watertight(scenes_mod):
    """Shared vertices are bit-identical across faces: every edge is used twice."""
    sc = scenes_mod.generate(2, detail=12)
    s = sc.scene.contents
    buf = np.ctypeslib.as_array(C.cast(s.triangles, C.POINTER(C.c_float)), shape=(s.num_triangles * 28,))
    v = buf.reshape(s.num_triangles, 28)[:, :9].reshape(-1, 3, 3)
    edges = {}
    for tri in v:
        keys = [tuple(p.tobytes() for p in (tri[i], tri[(i + 1) % 3])) for i in range(3)]
        for a, b in keys:
            k = (min(a, b), max(a, b))
            edges[k] = edges.get(k, 0) + 1
    assert set(edges.values()) == {2}


def test_srgb_encoding_closed_form():
    """IEC 61966-2-1 curve then byte = min(255, enc*255 + 0.5) (png_color.h:11-21)."""
    x = np.concatenate([np.linspace(-0.5, 1.5, 2001), [0.0031308, 0.0031309, np.nan]]).astype(np.float32)
    rgba = np.zeros((1, len(x), 4), np.float32)
    rgba[0, :, 0] = x
    from sheen_amd import sheen
    out = sheen.encode_srgb(rgba)[0, :, 0]
    c = np.clip(np.nan_to_num(x.astype(np.float64), nan=0.0), 0, 1)
    enc = np.where(c <= 0.0031308, c * 12.92, 1.055 * c ** (1 / 2.4) - 0.055)
    ref = np.minimum(enc * 255 + 0.5, 255).astype(np.int64)
    assert np.abs(out.astype(np.int64) - ref).max() <= 1
    assert np.mean(out.astype(np.int64) == ref) > 0.995


def test_png_writer_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, size=(13, 17, 3), dtype=np.uint8)
    p = tmp_path / "a.png"
    assert native.capi().sheen_write_png(str(p).encode(), 17, 13, img.ctypes.data) == 0
    data = p.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "big")
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        assert zlib.crc32(typ + body) == int.from_bytes(data[pos + 8 + n:pos + 12 + n], "big")
        if typ == b"IHDR":
            assert body[:8] == (17).to_bytes(4, "big") + (13).to_bytes(4, "big") and body[8:10] == b"\x08\x02"
        if typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(13, 1 + 17 * 3)
    assert (raw[:, 0] == 0).all()
    assert np.array_equal(raw[:, 1:].reshape(13, 17, 3), img)


def _host_tree(sc, cat):
    H = native.host()
    info = (C.c_int * 5)()
    assert H.sheen_host_bvh_nodes(sc.scene, cat, None, None, None, None, info) == 0
    n_nodes, n_prims, count = info[0], info[1], info[2]
    nodes = np.zeros((n_nodes, 128), np.uint8)
    prims = np.zeros(n_prims, np.uint32)
    bounds = np.zeros(6 * count, np.float32)
    skip = np.zeros(count, np.uint8)
    assert H.sheen_host_bvh_nodes(sc.scene, cat, native.as_ptr(nodes), native.as_ptr(prims), native.as_ptr(bounds),
                                  native.as_ptr(skip), info) == 0
    return nodes, prims, bounds, skip


def _violations(nodes, prims, bounds, skip):
    r = (C.c_int * 2)()
    native.host().sheen_host_wbvh_check(native.as_ptr(nodes), len(nodes), native.as_ptr(prims), len(prims),
                                        native.as_ptr(bounds), len(skip), native.as_ptr(skip), r)
    return r[1]


def test_wide_bvh_check_catches_planted_boxes(scenes_mod):
    """A child box that reaches past its parent's decoded box on either side,
    a leaf that loses a primitive, a primitive referenced twice, and slots out
    of the internal-first order are each counted as violations (the check
    itself is pinned, not only the trees)."""
    sc = scenes_mod.generate(2, detail=24)
    nodes, prims, bounds, skip = _host_tree(sc, 2)
    assert len(nodes) > 8
    assert _violations(nodes, prims, bounds, skip) == 0
    meta = nodes[:, 24:32]
    q = nodes[:, 32:128].view(np.uint16).reshape(-1, 6, 8)
    org = nodes[:, 0:12].view(np.float32).reshape(-1, 3).astype(np.float64)
    scale = 2.0 ** (nodes[:, 12:15].astype(np.float64) - 127.0)
    qv = q.view(np.float16).astype(np.float64)
    # an internal child whose x-max plane, moved to the top of its grid,
    # lies well outside the box its parent decodes for it
    plant = None
    for parent in range(len(nodes)):
        for slot in range(8):
            m = int(meta[parent, slot])
            if m == 0xFF or (m >> 5):
                continue
            child = int(nodes[parent, 16:20].view(np.uint32)[0]) + (m & 31)
            parent_hi = org[parent, 0] + qv[parent, 1, slot] * scale[parent, 0]
            top = org[child, 0] + 2048.0 * scale[child, 0]
            for j in range(8):
                if meta[child, j] != 0xFF and top > parent_hi + 8 * scale[child, 0]:
                    plant = (child, j)
                    break
            if plant:
                break
        if plant:
            break
    assert plant is not None
    child, j = plant
    bad = nodes.copy()
    qb = bad[:, 32:128].view(np.uint16).reshape(-1, 6, 8)
    qb[child, 1, j] = np.float16(2048.0).view(np.uint16)
    assert _violations(bad, prims, bounds, skip) > 0
    # lower side: the child's origin moved far below its parent's box
    bad = nodes.copy()
    o = bad[child, 0:12].view(np.float32)
    o[1] -= 1000.0
    assert _violations(bad, prims, bounds, skip) > 0
    # a primitive referenced twice and one never referenced
    badp = prims.copy()
    badp[1] = badp[0]
    assert _violations(nodes, badp, bounds, skip) > 0
    # slot layout: a node's last internal child and first leaf swapped (the
    # traversal takes slot k for internal child k), and a wrong internal count
    n = next(i for i in range(len(nodes)) if 0 < (nodes[i, 15] >> 4) < (nodes[i, 15] & 15))
    ni = int(nodes[n, 15] >> 4)
    bad = nodes.copy()
    mb = bad[:, 24:32]
    qb = bad[:, 32:128].view(np.uint16).reshape(-1, 6, 8)
    mb[n, [ni - 1, ni]] = mb[n, [ni, ni - 1]]
    qb[n, :, [ni - 1, ni]] = qb[n, :, [ni, ni - 1]]
    assert _violations(bad, prims, bounds, skip) > 0
    bad = nodes.copy()
    bad[n, 15] = (bad[n, 15] & 15) | ((ni - 1) << 4)
    assert _violations(bad, prims, bounds, skip) > 0


def test_one_hip_runtime_when_torch_is_imported_first():
    """torch's bundled libamdhip64.so and the system ROCm one share the soname
    libamdhip64.so.7, so once torch is loaded the library's HIP calls bind to
    torch's runtime: one runtime per process, and torch's stream handles mean
    the same thing to the library (the adjoint entry points rely on it)."""
    script = ("import torch, sheen_amd.native as n; n.capi(); "
              "print(sorted({l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}))")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, cwd=root, timeout=300)
    assert r.returncode == 0, r.stderr
    libs = eval(r.stdout.strip().splitlines()[-1])
    assert len(libs) == 1, libs
