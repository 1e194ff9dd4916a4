"""GPU parity of the HIP render path against the CPU oracle.

Primary hits (category, index, distance) must match bit for bit.  Radiance
must agree within 1e-4 relative per pixel (relative error
|gpu - oracle| / max(|oracle|, 1e-3), worst channel): every case bounds both
the fraction of pixels beyond 1e-4 and the largest relative error, and
prints both.  Transcendentals on the shading, light-sampling, media and
photon paths are evaluated in double and rounded once on both sides, so the
measured bounds are zero excursions for every case.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (config, overrides, width, height, aa, max fraction of pixels beyond 1e-4, max relative error)
CASES = [
    ("cfg1", 1, {}, 256, 256, 1, 0.0, 1e-4),
    ("mesh", 2, {"detail": 40}, 192, 108, 4, 0.0, 1e-4),
    ("lights", 100, {}, 160, 120, 4, 0.0, 1e-4),
    ("meshes", 101, {}, 160, 120, 4, 0.0, 1e-4),
    ("textures", 102, {}, 160, 120, 4, 0.0, 1e-4),
    ("caustics", 103, {}, 160, 120, 2, 0.0, 1e-4),
    ("shading", 5, {"detail": 2000}, 128, 128, 2, 0.0, 1e-4),
    ("analytic", 104, {}, 160, 90, 4, 0.0, 1e-4),
    ("features", 105, {}, 160, 120, 4, 0.0, 1e-4),
]


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-3)


@pytest.fixture(scope="module")
def renders():
    from sheen_amd import scenes, sheen
    from oracle import Oracle
    out = {}
    for name, cfg, kw, w, h, aa, max_frac, max_rel in CASES:
        print("case", name, flush=True)
        sc = scenes.generate(cfg, **kw)
        with sheen.Context(sc.scene) as ctx:
            img, st = ctx.render(w, h, aa)
            kind, idx, t = ctx.primary_hits(w, h, aa)
        ref = Oracle(sc.scene, threads=min(16, os.cpu_count() or 1)).render(w, h, aa, hits=True)
        out[name] = dict(img=img, st=st, kind=kind, idx=idx, t=t, ref=ref, max_frac=max_frac, max_rel=max_rel)
    return out


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_primary_hits_bit_exact(renders, name):
    r = renders[name]
    ref = r["ref"]
    assert np.array_equal(r["kind"], ref["kind"])
    assert np.array_equal(r["idx"], ref["index"])
    assert np.array_equal(r["t"].view(np.uint32), ref["t"].view(np.uint32))


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_radiance_within_1e4(renders, name):
    r = renders[name]
    m = _rel(r["img"][..., :3], r["ref"]["image"][..., :3]).max(-1)
    frac = float(np.mean(m > 1e-4))
    print(name, "max rel %.3g, beyond 1e-4: %.6f" % (m.max(), frac), flush=True)
    assert frac <= r["max_frac"]
    assert m.max() <= r["max_rel"]


@pytest.mark.parametrize("name", ["cfg1", "mesh", "meshes", "shading"])
def test_ray_counts_match_oracle(renders, name):
    st, ref = renders[name]["st"], renders[name]["ref"]["stats"]
    for k in ("primary_rays", "reflection_rays", "refraction_rays", "shadow_rays"):
        assert st[k] == ref[k], k


def test_chunking_does_not_change_the_image():
    from sheen_amd import scenes, sheen
    sc = scenes.generate(101)
    with sheen.Context(sc.scene) as ctx:
        a, _ = ctx.render(160, 120, 4)
        ctx.set_chunk(64 * 4 * 7)
        b, _ = ctx.render(160, 120, 4)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("cfg", [1, 101, 102, 105])
def test_launch_shape_does_not_change_the_image(cfg):
    """Queue order depends on the grid, the shading workgroup size and the
    queue-order knobs; every queue entry carries its sample slot, so the image
    is bit-identical for any launch shape."""
    from sheen_amd import scenes, sheen, autotune
    sc = scenes.generate(cfg)
    shapes = [{k: v for k, v in autotune.DEFAULTS.items()},
              {"primary_grid": 1, "sec_grid": 1, "shadow_grid": 1, "shade_grid": 1, "shade_sec_grid": 1},
              {"primary_grid": 0, "sec_grid": 0, "shadow_grid": 0, "shade_grid": 0, "shade_sec_grid": 0,
               "chunk": 64 * 4 * 5},
              {"shade_block": 256, "shadow_order": 0, "sec_order": 0, "chunk": 0},
              {"shade_block": 64, "shadow_order": 1, "sec_order": 1, "light_runs": 0}]
    with sheen.Context(sc.scene) as ctx:
        imgs = []
        for p in shapes:
            autotune.apply(ctx, p)
            imgs.append(ctx.render(160, 120, 4)[0])
    for img in imgs[1:]:
        assert np.array_equal(img, imgs[0])


@pytest.mark.parametrize("cfg", [2, 101, 103, 104, 105])
def test_side_stream_shadows_do_not_change_the_image(cfg):
    """The primary shading's shadow requests traced on a second stream while
    the secondary rays run give the same image and ray counts as one shadow
    launch after them (the default), across chunk boundaries too, and so does
    the shadow queue kept in one run per light (the default) or one run shared
    by the lights."""
    from sheen_amd import scenes, sheen
    sc = scenes.generate(cfg, **({"detail": 40} if cfg == 2 else {}))
    out = {}
    with sheen.Context(sc.scene) as ctx:
        for runs in (1, 0):
            ctx.set_param("light_runs", runs)
            for overlap in (0, 1):
                ctx.set_param("overlap", overlap)
                for chunk in (0, 64 * 4 * 5):
                    ctx.set_chunk(chunk)
                    img, st = ctx.render(160, 120, 4)
                    out[runs, overlap, chunk] = (img, st)
    ref, st0 = out[1, 0, 0]
    for key, (img, st) in out.items():
        assert np.array_equal(img, ref), key
        for k in ("shadow_rays", "reflection_rays", "refraction_rays", "total_rays"):
            assert st[k] == st0[k], (key, k)


@pytest.mark.parametrize("cfg", [2, 101, 103, 104])
def test_sorted_secondary_rays_do_not_change_the_image(cfg):
    """Secondary rays sorted by kind and direction before each trace round
    (sec_sort 1: octant, 2: octant and dominant axis) give the same image and
    ray counts as the queue in shading order, within one chunk and across
    chunk boundaries, with the side-stream shadows on and off."""
    from sheen_amd import scenes, sheen
    sc = scenes.generate(cfg, **({"detail": 40} if cfg == 2 else {}))
    out = {}
    with sheen.Context(sc.scene) as ctx:
        for mode in (0, 1, 2):
            ctx.set_param("sec_sort", mode)
            for overlap, chunk in ((0, 0), (1, 64 * 4 * 5)):
                ctx.set_param("overlap", overlap)
                ctx.set_chunk(chunk)
                out[mode, overlap, chunk] = ctx.render(160, 120, 4)
        with pytest.raises(RuntimeError, match="sec_sort"):
            ctx.set_param("sec_sort", 3)
    ref, st0 = out[0, 0, 0]
    assert st0["reflection_rays"] + st0["refraction_rays"] > 0
    for key, (img, st) in out.items():
        assert np.array_equal(img, ref), key
        for k in ("shadow_rays", "reflection_rays", "refraction_rays", "total_rays"):
            assert st[k] == st0[k], (key, k)


@pytest.mark.parametrize("cfg", [1, 2, 100, 101, 102, 104, 105])
def test_skipping_dark_shadow_rays_keeps_the_image(cfg):
    """With `skip_dark`, shadow requests whose unshadowed contribution is
    exactly zero are not traced: the image is bit-identical (they add
    0 x transmittance) while fewer shadow rays are counted."""
    from sheen_amd import scenes, sheen
    sc = scenes.generate(cfg, **({"detail": 40} if cfg == 2 else {}))
    w, h = (160, 120) if cfg != 1 else (128, 128)
    with sheen.Context(sc.scene) as ctx:
        ref, st0 = ctx.render(w, h, 4)
        ctx.set_param("skip_dark", 1)
        img, st1 = ctx.render(w, h, 4)
        ctx.set_chunk(64 * 4 * 5)
        img2, _ = ctx.render(w, h, 4)
    assert np.array_equal(img, ref) and np.array_equal(img2, ref)
    assert st1["shadow_rays"] <= st0["shadow_rays"]
    assert st1["primary_rays"] == st0["primary_rays"] and st1["reflection_rays"] == st0["reflection_rays"]


def test_autotuner_picks_from_its_space_and_keeps_the_image():
    from sheen_amd import scenes, sheen, autotune
    sc = scenes.generate(2, detail=40)
    space = {"shadow_grid": (8, 64), "shade_grid": (1, 4)}
    with sheen.Context(sc.scene) as ctx:
        ref, _ = ctx.render(192, 108, 4)
        res = autotune.tune(ctx, 192, 108, 4, space=space, reps=1)
        img, _ = ctx.render(192, 108, 4)
        with pytest.raises(RuntimeError, match="unknown parameter"):
            ctx.set_param("no_such_knob", 1)
        with pytest.raises(RuntimeError, match="negative"):
            ctx.set_param("shadow_grid", -1)
        with pytest.raises(RuntimeError, match="shade_block"):
            ctx.set_param("shade_block", 100)
    assert res["params"]["shadow_grid"] in (8, 64) and res["params"]["shade_grid"] in (1, 4)
    assert len(res["trials"]) == 4 and res["frame_ms"] > 0
    assert np.array_equal(img, ref)


def test_photon_map_matches_oracle_and_reference_counts():
    """caustics.sheen: 10,868 of 20,000 photons stored (ROADMAP.md:313);
    hash-map gather equals brute force (ROADMAP.md:324-325).  Photon positions
    come through sin/cos of the spot sampling, so they agree with the oracle to
    rounding, not bit for bit (a photon on a cell boundary may change cell)."""
    from sheen_amd import scenes, sheen
    from sheen_amd import contract as K
    from oracle import Oracle
    import ctypes as C
    sc = scenes.generate(103)
    with sheen.Context(sc.scene) as ctx:
        emitted, stored = ctx.photon_stats()
        m = ctx.photon_probe(97, 0.125)
        b = ctx.photon_probe(97, 0.125, brute=True)
        gpu = (K.Photon * stored)()
        assert ctx._lib.sheen_context_photon_map(ctx._c, C.cast(gpu, C.c_void_p), None, None) == stored
    o = Oracle(sc.scene, threads=4)
    assert (emitted, stored) == o.photon_stats() == (20000, 10868)
    assert np.array_equal(m, b)
    assert 270 <= int(m.mean()) <= 290
    ref = (K.Photon * stored)()
    assert o._L.oracle_photon_map(o._h, C.cast(ref, C.c_void_p), None, None) == stored
    pg = np.ctypeslib.as_array(C.cast(gpu, C.POINTER(C.c_float)), shape=(stored * 12,)).reshape(stored, 12)
    pr = np.ctypeslib.as_array(C.cast(ref, C.POINTER(C.c_float)), shape=(stored * 12,)).reshape(stored, 12)
    kg = pg[np.lexsort(pg[:, ::-1].T)]
    kr = pr[np.lexsort(pr[:, ::-1].T)]
    np.testing.assert_allclose(kg, kr, rtol=1e-4, atol=1e-5)


def test_reference_module_writes_png(tmp_path):
    """sheen.render_scene (reference api/sheen_py_module.cpp) -> sRGB PNG."""
    import zlib
    from sheen_amd import scenes, sheen
    from oracle import Oracle
    sc = scenes.generate(100)
    path = tmp_path / "out.png"
    sheen.init(0, 0)
    sheen.render_scene(sc.scene.contents, 64, 48, 2, str(path))
    data = path.read_bytes()
    pos, idat = 8, b""
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "big")
        if data[pos + 4:pos + 8] == b"IDAT":
            idat += data[pos + 8:pos + 8 + n]
        pos += 12 + n
    px = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(48, 1 + 64 * 3)[:, 1:].reshape(48, 64, 3)
    ref = Oracle(sc.scene, threads=4).render(64, 48, 2)["image"]
    ref8 = sheen.encode_srgb(ref)
    diff = np.abs(px.astype(int) - ref8.astype(int))
    assert diff.max() <= 255 and np.mean(diff > 1) < 0.01
