"""Sky, reflection, refraction and the glass passes of shadow rays, pinned by
closed forms (reference render_kernel.cu:243-268 shadow passes, 329-334
sky_color, 672-733 reflection, 736-807 refraction; integrator/shading.h:
92-108 mask transmission and IOR), each pin checked to fail with the planted
mistakes it targets (pin_util.run_pin).  Near-orthographic cameras: every
sample of a pixel follows the same path up to ~1e-6.

* sky: a ray that meets nothing returns (1 - t) (1, 1, 1) + t (0.5, 0.7, 1),
  t = (d.y + 1) / 2;
* mirror floor of reflect w under an overhead sun: c (1 - min(w, 1)) +
  r min(w, 1), c = A a + A I, r = sky(d - 2 (d.n) n) or, when the
  reflected ray meets an emissive wall, the wall's own shading;
* glass slab (thickness T, IOR eta) over a floor lit by a point light: a ray
  entering at theta_i leaves parallel to itself and reaches the floor
  shifted by T tan(theta_t) + h tan(theta_i), sin(theta_t) =
  sin(theta_i) / eta; the pixel is w1 w2 L_floor(x) + c_top (1 - min(w1, 1))
  with face transmissions w1, w2 (a mask map sets them to
  transmission * R / 255 and the IOR to 1 + G / 255);
* right-angle prism: a vertical ray is totally reflected at the 45 degree
  face (critical angle asin(1/1.5) = 41.8 degrees) and leaves through the
  vertical face horizontally, so the pixel is the emission of the wall it
  then meets, not of the floor below;
* shadow rays pass at most three faces of transmission > 0.5 (a fourth face
  is never tested, so a blocker behind three glass sheets does not block),
  and keep their original t_max while passing: a blocker just beyond a point
  light shadows it when glass lies between.
"""
import math

import numpy as np
import pytest

from pin_util import camera, f32, gpu_render, hquad, light, mat, quad, run_pin, scene, texture, unit

A = np.array([0.8, 0.6, 0.4])
AMB = 0.1


def _sky(d):
    d = f32(d)
    t = 0.5 * (d[1] + 1.0)
    return (1 - t) * np.ones(3) + t * np.array([0.5, 0.7, 1.0])


SKY_DIRS = [(0, 1, 0), (0, -1, 0), (1, 0, 0), unit((1, 1, 1)), unit((-0.3, 0.2, 0.9)), unit((0.5, -0.7, 0.1))]


def _sky_scene(d):
    ps = scene([mat()])
    camera(ps, (0, 0, 0), d, dist=1.0)
    return ps


def test_sky_gradient():
    def check(mode):
        for d in SKY_DIRS:
            img = mode.render(_sky_scene(d))
            np.testing.assert_allclose(img, np.broadcast_to(_sky(unit(d)), img.shape), rtol=1e-5, err_msg=str(d))

    run_pin(check, ["SKY_X", "SKY_SWAP"])


P0 = np.array([0.3, 0.0, -0.2])
VIEW = unit((0.0, -1.0, -1.0))
WALL_E = np.array([0.2, 1.5, 0.7])


def _sun_overhead(inten=1.0):
    return light("sun", direction=(0.0, -1.0, 0.0), intensity=inten, angle=0.0)


def _mirror_scene(w, wall):
    mats = [mat(color=A, amb=AMB, reflect=w), mat(color=(0.5, 0.5, 0.5), amb=0.2, emission_color=WALL_E, emission=1.0)]
    tris = hquad(0.0, -50, 50, -50, 50, 0)
    if wall:  # a vertical wall at z = -5 facing +z
        tris += quad((-20, 0, -5), (20, 0, -5), (20, 20, -5), (-20, 20, -5), 1)
    ps = scene(mats, tris, [_sun_overhead(1.3)])
    camera(ps, P0, VIEW, dist=1.0)
    return ps


def test_mirror_blend():
    def check(mode):
        d = f32(VIEW)
        rd = unit(d - 2 * d[1] * np.array([0, 1, 0]))
        c = A * AMB + A * 1.3
        for wall in (False, True):
            r = (np.array([0.5, 0.5, 0.5]) * 0.2 + WALL_E) if wall else _sky(rd)
            for w in (0.3, 1.0, 1.7):
                img = mode.render(_mirror_scene(w, wall))
                ww = min(w, 1.0)
                np.testing.assert_allclose(img, np.broadcast_to(c * (1 - ww) + r * ww, img.shape), rtol=1e-5,
                                           err_msg=str((wall, w)))

    run_pin(check, ["REFL_NO_CLAMP", "REFL_SWAP", "SKY_X"])


LIGHT = np.array([3.0, 0.2, 0.0])
LIGHT_I = 6.0
FLOOR_AMB = 0.01  # small, so the floor radiance is mostly the light's, which varies with x
THETA_I = 0.6
TOP, BOT = 1.0, 0.5


def _floor_radiance(x):
    p = np.array([x, 0.0, 0.0])
    v = f32(LIGHT) - p
    d2 = float(v @ v)
    return A * FLOOR_AMB + A * LIGHT_I * (v[1] / math.sqrt(d2)) / (4 * math.pi * d2)


def _slab_scene(top_mat, bot_mat, mask=None):
    mats = [mat(color=A, amb=FLOOR_AMB), top_mat, bot_mat]
    tris = hquad(0.0, -50, 50, -50, 50, 0)
    tris += hquad(TOP, -50, 50, -50, 50, 1, uv=True) + hquad(BOT, -50, 50, -50, 50, 2, uv=True)
    ps = scene(mats, tris, [light("point", position=tuple(LIGHT), intensity=LIGHT_I)])
    if mask is not None:
        texture(ps, [mask])
    d = (math.sin(THETA_I), -math.cos(THETA_I), 0.0)
    camera(ps, (0.0, TOP, 0.0), d, dist=0.5)
    return ps


def _slab_x(eta):
    d = f32((math.sin(THETA_I), -math.cos(THETA_I), 0.0))
    d = d / np.linalg.norm(d)
    sin_i = d[0]
    sin_t = sin_i / eta
    tan_t = sin_t / math.sqrt(1 - sin_t * sin_t)
    tan_i = sin_i / math.sqrt(1 - sin_i * sin_i)
    return (TOP - BOT) * tan_t + BOT * tan_i


def test_glass_slab_shift_and_face_weights():
    B = np.array([0.3, 0.5, 0.7])

    def check(mode):
        # clear glass, second face transmission 0.8
        ps = _slab_scene(mat(color=B, amb=0.2, transmission=1.0, ior=1.5),
                         mat(color=B, amb=0.2, transmission=0.8, ior=1.5))
        img = mode.render(ps)
        want = 0.8 * _floor_radiance(_slab_x(1.5))
        np.testing.assert_allclose(img, np.broadcast_to(want, img.shape), rtol=2e-5, err_msg="clear")
        # first face transmission 0.7: its own colour keeps 0.3
        ps = _slab_scene(mat(color=B, amb=0.2, transmission=0.7, ior=1.5),
                         mat(color=B, amb=0.2, transmission=0.8, ior=1.5))
        img = mode.render(ps)
        want = B * 0.2 * 0.3 + 0.7 * 0.8 * _floor_radiance(_slab_x(1.5))
        np.testing.assert_allclose(img, np.broadcast_to(want, img.shape), rtol=2e-5, err_msg="tinted")
        # mask map: transmission 1 * 204/255 on both faces, IOR 1 + 128/255
        mask = np.zeros((4, 4, 3), np.uint8)
        mask[..., 0], mask[..., 1] = 204, 128
        gm = mat(color=B, amb=0.2, transmission=1.0, ior=1.2, mask_id=0)
        img = mode.render(_slab_scene(gm, gm, mask))
        tr = float(np.float32(204) / np.float32(255))
        eta = 1.0 + float(np.float32(128) / np.float32(255))
        want = B * 0.2 * (1 - tr) + tr * tr * _floor_radiance(_slab_x(eta))
        np.testing.assert_allclose(img, np.broadcast_to(want, img.shape), rtol=2e-5, err_msg="mask")

    run_pin(check, ["ETA_FIXED", "W_ONCE", "IOR_RED", "TRANS_NO_MASK"])


FLOOR_E = np.array([0.9, 0.1, 0.1])
PRISM_WALL_E = np.array([0.1, 0.2, 1.1])


def _prism_scene():
    g = mat(color=(0.5, 0.5, 0.5), amb=0.3, transmission=1.0, ior=1.5)
    dark = dict(color=(0, 0, 0), amb=0.0)
    mats = [g, mat(emission_color=FLOOR_E, emission=1.0, **dark), mat(emission_color=PRISM_WALL_E, emission=1.0, **dark)]
    z0, z1 = -5.0, 5.0
    tris = quad((0, 1, z0), (1, 1, z0), (1, 1, z1), (0, 1, z1), 0)       # top, y = 1
    tris += quad((1, 1, z0), (1, 0, z0), (1, 0, z1), (1, 1, z1), 0)      # vertical, x = 1
    tris += quad((1, 0, z0), (0, 1, z0), (0, 1, z1), (1, 0, z1), 0)      # 45 degree face x + y = 1
    tris += hquad(-1.0, -50, 50, -50, 50, 1)                             # floor
    tris += quad((3, -5, -20), (3, 5, -20), (3, 5, 20), (3, -5, 20), 2)  # wall, x = 3
    ps = scene(mats, tris)
    camera(ps, (0.7, 1.0, 0.0), (0, -1, 0), dist=0.5)
    return ps


def test_total_internal_reflection_in_a_prism():
    def check(mode):
        img = mode.render(_prism_scene())
        np.testing.assert_allclose(img, np.broadcast_to(PRISM_WALL_E, img.shape), rtol=1e-6)

    run_pin(check, ["TIR_SIGN"])


L45 = unit((1.0, 1.0, 0.0))


def _sheet_scene(n_glass, opaque=True, tr=1.0):
    """Sun along (1, 1, 0); sheets at heights 1, 2, ... across the shadow ray
    of the floor point (0, 0, 0), out of the camera's view."""
    mats = [mat(color=A, amb=AMB), mat(transmission=tr, ior=1.5), mat(color=(0.2, 0.2, 0.2))]
    tris = hquad(0.0, -50, 50, -50, 50, 0)
    for k in range(n_glass):
        y = 1.0 + k
        tris += hquad(y, y - 0.5, y + 0.5, -0.5, 0.5, 1)
    if opaque:
        y = 1.0 + n_glass
        tris += hquad(y, y - 0.5, y + 0.5, -0.5, 0.5, 2)
    sun = light("sun", direction=tuple(-np.float32(L45)), intensity=1.0, angle=0.0)
    ps = scene(mats, tris, [sun])
    camera(ps, (0, 0, 0), (0, -1, 0), dist=0.5)
    return ps


def test_shadow_rays_pass_at_most_three_glass_faces():
    lit = A * AMB + A * float(f32(L45)[1])
    dark = A * AMB

    def check(mode):
        for n_glass, opaque, tr, want in ((0, True, 1.0, dark), (1, True, 1.0, dark), (2, True, 1.0, dark),
                                          (3, True, 1.0, lit), (3, False, 1.0, lit), (1, False, 0.4, dark),
                                          (1, False, 0.6, lit)):
            img = mode.render(_sheet_scene(n_glass, opaque, tr))
            np.testing.assert_allclose(img, np.broadcast_to(want, img.shape), rtol=1e-5,
                                       err_msg=str((n_glass, opaque, tr)))

    run_pin(check, ["SHADOW_4", "SHADOW_THRESH"])


def _beyond_light_scene(glass):
    mats = [mat(color=A, amb=AMB), mat(transmission=1.0, ior=1.5), mat(color=(0.2, 0.2, 0.2))]
    tris = hquad(0.0, -50, 50, -50, 50, 0) + hquad(2.4, -0.5, 0.5, -0.5, 0.5, 2)
    if glass:
        tris += hquad(1.0, -0.5, 0.5, -0.5, 0.5, 1)
    ps = scene(mats, tris, [light("point", position=(0.0, 2.0, 0.0), intensity=5.0)])
    camera(ps, (0, 0, 0), (0, -1, 0), dist=0.5)
    return ps


def test_glass_pass_keeps_t_max():
    def check(mode):
        img = mode.render(_beyond_light_scene(False))
        lit = A * AMB + A * 5.0 / (4 * math.pi * 4.0)
        np.testing.assert_allclose(img, np.broadcast_to(lit, img.shape), rtol=1e-5)
        img = mode.render(_beyond_light_scene(True))
        np.testing.assert_allclose(img, np.broadcast_to(A * AMB, img.shape), rtol=1e-5)

    run_pin(check, ["SHADOW_TMAX"])


@pytest.mark.gpu
def test_secondary_pins_gpu_matches_oracle():
    from pin_util import Mode
    mask = np.zeros((4, 4, 3), np.uint8)
    mask[..., 0], mask[..., 1] = 204, 128
    B = (0.3, 0.5, 0.7)
    gm = mat(color=B, amb=0.2, transmission=1.0, ior=1.2, mask_id=0)
    scenes = [_sky_scene(d) for d in SKY_DIRS[:3]]
    scenes += [_mirror_scene(0.3, False), _mirror_scene(1.7, True)]
    scenes += [_slab_scene(mat(color=B, amb=0.2, transmission=0.7, ior=1.5),
                           mat(color=B, amb=0.2, transmission=0.8, ior=1.5)), _slab_scene(gm, gm, mask)]
    scenes += [_prism_scene(), _sheet_scene(3, True), _sheet_scene(2, True), _sheet_scene(1, False, 0.4),
               _beyond_light_scene(True), _beyond_light_scene(False)]
    for k, ps in enumerate(scenes):
        got = gpu_render(ps, 8, 8, 2)
        ref = Mode().render(ps, 8, 8, 2)
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-7, err_msg="scene %d" % k)
