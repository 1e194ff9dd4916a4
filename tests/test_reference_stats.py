"""Pixel statistics the reference recorded for its own scene file, reproduced
by the oracle (CPU; skipped where the reference's scenes/ is not mounted).

reference ROADMAP.md:229-237 on scenes/emission_demo.sheen (640x480, AA 4,
three emissive lamp spheres over a half-mirror floor, reflect 0.35): with
`emit 2 1.0 0.80 0.45 3.0` the lamp centres are (255, 255, 255) and without
it (17, 18, 21); the floor's reflection of a lamp is (255, 239, 188)
against (41, 43, 50); 1,624 pixels change.  Measured here: every lamp's
reflection with emission is (255, 239, 188) exactly and every lamp centre
(255, 255, 255); without emission the middle lamp's centre is (19, 20, 23)
and its reflection (42, 43, 51), within 2 levels of the recorded 8-bit
values (the reference ran CUDA's transcendentals on a GTX 1080 Ti, the
oracle glibc's; the side lamps read 1-2 levels darker);
1,558 pixels change, 4 % fewer than recorded (the superellipsoid back wall
also mirrors the lamps, and its powf-bisected surface moves silhouette
pixels between the two maths libraries).  The lamp centre and its mirror
image in the floor y = 0 are projected through the file's legacy camera
(focal length 0: direction forward + u right + v up, tan = 1 on both axes).
Pins the emission term, the reflection blend weight and the sRGB bytes;
shown to fail with emission scaled by the ambient term or the reflection
weights exchanged.
"""
import ctypes as C
import os

import numpy as np
import pytest

from pin_util import run_pin

SCENE = "/root/reference/scenes/emission_demo.sheen"
LAMPS = [(-0.8, 3.2, 0.0), (0.0, 3.35, 0.0), (0.8, 3.2, 0.0)]

pytestmark = pytest.mark.skipif(not os.path.exists(SCENE), reason="reference scenes are not mounted")


def _srgb(img, w, h):
    from sheen_amd import native
    out = np.zeros((h, w, 3), np.uint8)
    rgba = np.ascontiguousarray(img, np.float32)
    native.capi().sheen_encode_srgb(rgba.ctypes.data_as(C.c_void_p), w, h, out.ctypes.data_as(C.c_void_p))
    return out


def _project(cam, p, w, h):
    pos = np.array([cam.position.x, cam.position.y, cam.position.z])
    fw = np.array([cam.forward.x, cam.forward.y, cam.forward.z])
    d = np.asarray(p) - pos
    lam = d[2] / fw[2]
    u = d[0] / lam - fw[0]
    v = d[1] / lam - fw[1]
    return int((u + 1) / 2 * w), int((1 - v) / 2 * h)


def test_emission_demo_statistics():
    from sheen_amd import scenes
    f = scenes.load_sheen(SCENE)
    W, H, AA = f.width, f.height, f.aa
    assert (W, H, AA) == (640, 480, 4)
    cam = f.scene.contents.camera
    assert cam.focal_length == 0.0
    assert (cam.right.x, cam.up.y, cam.forward.z) == (1.0, 1.0, -1.0)
    mats = f.scene.contents.materials
    assert mats[2].emission == 3.0

    def check(mode):
        mats[2].emission = 3.0
        on = _srgb(mode.render(f, W, H, AA, raw=True), W, H)
        mats[2].emission = 0.0
        off = _srgb(mode.render(f, W, H, AA, raw=True), W, H)
        mats[2].emission = 3.0
        for k, c in enumerate(LAMPS):
            x, y = _project(cam, c, W, H)
            xm, ym = _project(cam, (c[0], -c[1], c[2]), W, H)
            assert tuple(on[y, x]) == (255, 255, 255), (c, on[y, x])
            assert tuple(on[ym, xm]) == (255, 239, 188), (c, on[ym, xm])
            if k == 1:  # the middle lamp
                assert np.all(np.abs(off[y, x].astype(int) - (17, 18, 21)) <= 2), (c, off[y, x])
                assert np.all(np.abs(off[ym, xm].astype(int) - (41, 43, 50)) <= 2), (c, off[ym, xm])
        changed = int(np.any(on != off, -1).sum())
        assert abs(changed - 1624) <= 0.06 * 1624, changed

    run_pin(check, ["EMISSION_LIT", "REFL_SWAP"])


BEAM_SCENE = "/root/reference/scenes/beam_shadow.sheen"


@pytest.mark.skipif(not os.path.exists(BEAM_SCENE), reason="reference scenes are not mounted")
def test_beam_shadow_statistics():
    """reference ROADMAP.md:118-134 on scenes/beam_shadow.sheen (640x480,
    AA 4, a spot with beam 0.9 through a dark hall, a shield at the lamp and
    a column in the beam): tracing a shadow ray per beam step made 13,977
    pixels (4.55 % of the frame) darker than the beam without them, by at
    most 149/255, and took the mean brightness of those pixels from 117.5 to
    66.0.  The oracle, against its own beam without shadow rays (planted
    mistake BEAM_NO_SHADOW, the reference's earlier beam, ROADMAP.md:101-111):
    14,356 darker pixels (4.67 %, 2.7 % more than recorded: penumbra edges
    from jittered disk samples, where CUDA's transcendentals and ours round
    differently), darkest by exactly 149 levels in one channel, mean
    brightness (mean of the sRGB channels) 116.0 -> 65.8.  The GPU matches
    the oracle on this scene file (tests/test_gpu_reference_scenes.py)."""
    from sheen_amd import scenes
    from pin_util import Mode, P
    f = scenes.load_sheen(BEAM_SCENE)
    W, H, AA = f.width, f.height, f.aa
    assert (W, H, AA) == (640, 480, 4)
    bare = Mode(P["BEAM_NO_SHADOW"])
    try:
        off = _srgb(bare.render(f, W, H, AA, raw=True), W, H).astype(int)
    finally:
        bare.done()

    def check(mode):
        on = _srgb(mode.render(f, W, H, AA, raw=True), W, H).astype(int)
        darker = np.any(on < off, -1)
        n = int(darker.sum())
        assert abs(n - 13977) <= 0.04 * 13977, n
        assert int((off - on).max()) == 149, int((off - on).max())
        before, after = off.mean(-1)[darker].mean(), on.mean(-1)[darker].mean()
        assert abs(before - 117.5) <= 2.0 and abs(after - 66.0) <= 1.0, (before, after)

    run_pin(check, ["BEAM_NO_SHADOW"])
