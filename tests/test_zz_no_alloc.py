"""sheen_context_render_noalloc: the entry point for callers that cannot
allocate device memory.  Its work buffers live in the library's static
scratch array and the frame goes to a device buffer the caller owns; its
image must equal sheen_context_render's bit for bit, on two calls in a row
(the second smaller, so the scratch holds the first call's contents), for a
frame that needs many chunks of the scratch and for a scene with media,
beams and caustics.  The process's device allocations (torch's own aside)
are not what is checked here; the entry point's source calls no allocator.
"""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _noalloc(ctx, w, h, aa):
    import torch
    from sheen_amd import native
    out = torch.full((h, w, 4), float("nan"), device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    st = native.SheenRayStats()
    print("noalloc render %dx%d x%d" % (w, h, aa), flush=True)
    rc = ctx._lib.sheen_context_render_noalloc(ctx._c, w, h, aa, C.c_void_p(out.data_ptr()), C.byref(st))
    assert rc == 0, native.last_error()
    return out.cpu().numpy(), st.as_dict()


def _check(scene, sizes, min_chunks=1):
    """min_chunks: the first size must take at least that many chunks of the
    scratch (one shadow launch per chunk)."""
    from sheen_amd import sheen
    with sheen.Context(scene) as ctx:
        for k, (w, h, aa) in enumerate(sizes):
            img, st = _noalloc(ctx, w, h, aa)
            ref, rst = ctx.render(w, h, aa)
            assert np.array_equal(img, ref), (w, h, aa, int(np.sum(img != ref)))
            for key in ("samples", "primary_rays", "shadow_rays", "reflection_rays", "refraction_rays", "total_rays"):
                assert st[key] == rst[key], key
            assert st["shadow_launches"] >= (min_chunks if k == 0 else 1), st["shadow_launches"]


def test_noalloc_matches_render_over_many_chunks():
    """1080p x 4 spp with 3 lights needs about 20 chunks of the 256 MiB
    scratch; then a smaller frame in the same scratch."""
    from sheen_amd import scenes
    sc = scenes.generate(2, detail=40)
    _check(sc.scene, [(1920, 1080, 4), (320, 200, 2)], min_chunks=16)


def test_noalloc_full_scene_with_media_beams_and_caustics():
    from sheen_amd import scenes
    f = scenes.load_sheen(os.path.join(HERE, "scenes", "media.sheen"))
    _check(f.scene, [(f.width * 2, f.height * 2, f.aa), (f.width, f.height, 1)])


def test_noalloc_rejects_what_the_scratch_cannot_hold():
    """A scene with so many lights that one 8x8 tile of 64 samples needs more
    than the scratch: refused with an error, nothing launched."""
    from sheen_amd import contract as K
    from sheen_amd import native, sheen
    from util import PyScene, material, spheres, f3
    ps = PyScene()
    ps.set("materials", K.Material, [material()], "num_materials")
    ps.set("objects", K.Object, spheres([(0, 0, 0)], [1.0], [0]), "num_objects")
    lights = []
    for i in range(600):
        L = K.Light()
        L.type, L.position, L.color, L.intensity = 1, f3(0, 5, i * 0.01), f3(1, 1, 1), 1.0
        lights.append(L)
    ps.set("lights", K.Light, lights, "num_lights")
    import torch
    out = torch.empty((8, 8, 4), device="cuda")
    with sheen.Context(ps.ptr) as ctx:
        rc = ctx._lib.sheen_context_render_noalloc(ctx._c, 8, 8, 64, C.c_void_p(out.data_ptr()), None)
        assert rc == -1
        assert "beyond the" in native.last_error() and "scratch" in native.last_error()
        rc = ctx._lib.sheen_context_render_noalloc(ctx._c, 8, 8, 1, None, None)
        assert rc == -1 and native.last_error() == "null argument"
