"""Direct-lighting terms of the oracle pinned by closed forms (reference
render_kernel.cu:152-240, compute_lighting's light loop, and
integrator/shading.h:85-89,158-175), each pin checked to fail with the
planted mistakes it targets (pin_util.run_pin).

Every scene is a Lambert floor (albedo A, ambient a) seen by a
near-orthographic camera (rays within 2e-6 rad of the axis), so every sample
shades the same floor point p up to ~1e-6 and the pixel is the closed form
at p:
* point light at P:   A a + A C I cos(theta) / (4 pi d^2), d = |P - p|,
  cos(theta) = n . (P - p) / d;
* point light of radius R straight above p at height h: the disk sample is
  uniform over the disk perpendicular to the light direction, so the mean
  over N samples is A C I (2 h / R^2) (1/h - 1/sqrt(h^2 + R^2)) / (4 pi)
  within its Monte Carlo error;
* area light (centre c, half-axes u, v, n = normalize(u x v)): a vanishing
  quad gives the point-light value times cos_l = n . L, and nothing when
  cos_l <= 0; a large quad's mean is the integral of the point-light value
  times cos_l over the quad (uniform in the two axis parameters);
* spot light: zero outside cos(angle), the linear ramp
  (c - cos_half) / (cos_soft - cos_half) between cos_half and
  cos_soft = cos(angle (1 - blend)), full inside;
* Blinn-Phong under a sun along L (no cone) and the view V = +y:
  A a + A I (n.L) + I s (n.H)^shin, H = normalize(L + V), nothing specular
  when n.L <= 0; with a mask map, shin = clamp(2 / r^4 - 2, 1, 4096),
  r = max(B / 255, 0.02);
* emission: A a + E e, unlit and unshadowed.
"""
import math

import numpy as np
import pytest

from pin_util import camera, gpu_render, hquad, light, mat, run_pin, scene, texture, f32, unit

A = np.array([0.8, 0.6, 0.4])
AMB = 0.1
P0 = np.array([0.3, 0.0, -0.2])  # the floor point every sample sees


def _floor_scene(lights, m=None, uv=False):
    ps = scene([m if m is not None else mat(color=A, amb=AMB)], hquad(0.0, -50, 50, -50, 50, 0, uv=uv), lights)
    camera(ps, P0, (0, -1, 0), dist=1.0)
    return ps


def _point_value(pos, intensity, color=(1, 1, 1)):
    v = f32(pos) - P0
    d2 = float(v @ v)
    cos = v[1] / math.sqrt(d2)
    return A * AMB + A * np.asarray(color) * intensity * max(cos, 0.0) / (4 * math.pi * d2)


POINT_CASES = [((0.3, 2.0, -0.2), 3.0), ((1.8, 2.0, -0.2), 5.0), ((0.3, 1.0, 2.8), 7.0), ((-2.7, 0.5, -0.2), 2.0)]


def test_point_light_inverse_square_and_cosine():
    col = (1.0, 0.9, 0.7)

    def check(mode):
        for pos, inten in POINT_CASES:
            ps = _floor_scene([light("point", position=pos, intensity=inten, color=col)])
            img = mode.render(ps)
            np.testing.assert_allclose(img, np.broadcast_to(_point_value(pos, inten, col), img.shape), rtol=2e-5,
                                       err_msg=str(pos))

    run_pin(check, ["POINT_DIST", "NO_COS", "TWO_PI"])


def _disk_mean(h, R):
    return (2 * h / R ** 2) * (1 / h - 1 / math.sqrt(h * h + R * R)) / (4 * math.pi)


def _disk_sq(h, R):
    return (2 * h * h / R ** 2) * 0.25 * (1 / h ** 4 - 1 / (h * h + R * R) ** 2) / (4 * math.pi) ** 2


def test_point_light_disk_jitter_is_uniform_over_the_disk():
    h, R, inten = 1.0, 1.0, 10.0
    W = Hh = 64
    AA = 16

    def check(mode):
        ps = _floor_scene([light("point", position=(P0[0], h, P0[2]), intensity=inten, radius=R)])
        img = mode.render(ps, W, Hh, AA)
        n = W * Hh * AA
        mean = inten * _disk_mean(h, R)
        sd = inten * math.sqrt(_disk_sq(h, R) - _disk_mean(h, R) ** 2)
        got = (img.mean(axis=(0, 1)) - A * AMB) / A
        assert np.all(np.abs(got - mean) <= 5 * sd / math.sqrt(n)), (got, mean, sd / math.sqrt(n))
        # every sample lies on the disk: a pixel's samples average values
        # between the disk's edge and its centre
        per = (img - A * AMB) / A
        lo, hi = inten * h / (4 * math.pi * (h * h + R * R) ** 1.5), inten / (4 * math.pi * h * h)
        assert per.min() >= lo * (1 - 1e-5) and per.max() <= hi * (1 + 1e-5)

    run_pin(check, ["DISK_LINEAR"])


def _area_scene(c, u, v, inten):
    return _floor_scene([light("area", position=c, u_axis=u, v_axis=v, intensity=inten)])


def test_small_area_light_is_a_point_light_times_cos_l():
    eps = 1e-5
    cases = [((0.3, 2.0, -0.2), 0.0), ((0.3, 2.0, -0.2), 0.7), ((1.8, 2.0, -0.2), -0.4), ((1.8, 2.0, -0.2), 1.2)]

    def check(mode):
        for c, beta in cases:
            u = (eps, 0.0, 0.0)
            v = (0.0, eps * math.sin(beta), -eps * math.cos(beta))
            n = unit(np.cross(f32(u), f32(v)))
            L = unit(f32(c) - P0)
            cos_l = float(n @ L)
            assert cos_l > 0.2
            img = mode.render(_area_scene(c, u, v, 5.0))
            want = A * AMB + (_point_value(c, 5.0) - A * AMB) * cos_l
            np.testing.assert_allclose(img, np.broadcast_to(want, img.shape), rtol=5e-5, err_msg=str((c, beta)))
            # the other side of the same quad sends nothing
            img = mode.render(_area_scene(c, v, u, 5.0))
            np.testing.assert_allclose(img, np.broadcast_to(A * AMB, img.shape), rtol=1e-6)

    run_pin(check, ["AREA_NO_COS", "AREA_SIDE", "POINT_DIST"])


def _area_moments(c, u, v, inten, n=160):
    """Mean and second moment over the quad's (ru, rv) in [-1, 1]^2 of
    I cos_l cos(theta) / (4 pi d^2), by Gauss-Legendre quadrature."""
    x, w = np.polynomial.legendre.leggauss(n)
    ru, rv = np.meshgrid(x, x, indexing="ij")
    ww = np.outer(w, w) / 4.0
    c, u, v = f32(c), f32(u), f32(v)
    nrm = unit(np.cross(u, v))
    sp = c[None, None, :] + ru[..., None] * u + rv[..., None] * v
    t = sp - P0
    d2 = np.sum(t * t, -1)
    L = t / np.sqrt(d2)[..., None]
    cos_l = L @ nrm
    f = np.where(cos_l > 0, inten * cos_l * np.maximum(L[..., 1], 0) / (4 * math.pi * d2), 0.0)
    return float(np.sum(ww * f)), float(np.sum(ww * f * f))


def test_large_area_light_mean_is_its_integral():
    c, u, v, inten = (1.3, 1.5, -0.2), (1.0, 0.0, 0.0), (0.0, 0.0, -0.8), 20.0
    W = Hh = 64
    AA = 16

    def check(mode):
        img = mode.render(_area_scene(c, u, v, inten), W, Hh, AA)
        m1, m2 = _area_moments(c, u, v, inten)
        se = math.sqrt(m2 - m1 * m1) / math.sqrt(W * Hh * AA)
        got = (img.mean(axis=(0, 1)) - A * AMB) / A
        assert np.all(np.abs(got - m1) <= 5 * se), (got, m1, se)

    run_pin(check, ["AREA_HALF", "AREA_NO_COS"])


SPOT_ANGLE = 0.5


def _spot_value(theta, h, inten, blend):
    """Floor point at angle theta off the axis of a spot aimed straight down
    from height h."""
    d = h / math.cos(theta)
    c = math.cos(theta)
    cos_half = math.cos(SPOT_ANGLE)
    cos_soft = math.cos(SPOT_ANGLE * (1 - blend))
    if c <= cos_half:
        w = 0.0
    elif cos_soft > cos_half + 1e-4:
        w = min((c - cos_half) / (cos_soft - cos_half), 1.0)
    else:
        w = 1.0
    return A * AMB + A * inten * w * c / (4 * math.pi * d * d), w


def test_spot_cone_and_edge_ramp():
    h, inten = 2.0, 9.0
    cases = [(0.1, 0.4), (0.33, 0.4), (0.40, 0.4), (0.46, 0.4), (0.55, 0.4), (0.45, 0.0), (0.52, 0.0), (0.2, 1.0)]

    def check(mode):
        seen = set()
        for theta, blend in cases:
            pos = (P0[0] + h * math.tan(theta), h, P0[2])
            ps = _floor_scene([light("spot", position=pos, direction=(0.0, -1.0, 0.0), angle=SPOT_ANGLE,
                                     blend=blend, intensity=inten)])
            img = mode.render(ps)
            want, w = _spot_value(theta, h, inten, blend)
            seen.add("out" if w == 0 else "full" if w == 1 else "ramp")
            np.testing.assert_allclose(img, np.broadcast_to(want, img.shape), rtol=3e-5, err_msg=str((theta, blend)))
        assert seen == {"out", "full", "ramp"}

    run_pin(check, ["SPOT_RAMP", "SPOT_HALF_ANGLE", "POINT_DIST"])


def _sun(L, inten=1.0):
    L = np.asarray(L, np.float32)
    return light("sun", direction=tuple(float(x) for x in -L), intensity=inten, angle=0.0)


def _blinn(Ldir, inten, spec, shin, base=A):
    L = f32(Ldir)
    V = np.array([0.0, 1.0, 0.0])
    diff = max(L[1], 0.0)
    out = base * AMB + base * inten * diff
    if spec > 0 and diff > 0:
        H = unit(L + V)
        out = out + inten * spec * max(H[1], 0.0) ** shin
    return out


def test_blinn_phong_lobe():
    spec, shin, inten = 0.5, 32.0, 1.3
    dirs = [unit((0, 1, 0)), unit((math.sin(0.6), math.cos(0.6), 0)), unit((0.2, 0.9, -0.4)),
            unit((math.sin(1.62), math.cos(1.62), 0.0))]

    def check(mode):
        for Ld in dirs:
            ps = _floor_scene([_sun(Ld, inten)], m=mat(color=A, amb=AMB, spec=spec, shin=shin))
            img = mode.render(ps)
            np.testing.assert_allclose(img, np.broadcast_to(_blinn(Ld, inten, spec, shin), img.shape), rtol=2e-5,
                                       err_msg=str(Ld))

        # a shading normal tilted away from the sun (vertex normals), the
        # geometric floor still facing it: the shadow ray is open, n.L < 0
        # and n.H > 0, so only the guard keeps the lobe out
        nrm = unit((-0.8, 0.6, 0.0))
        Ld = unit((0.7, 0.7, 0.0))
        ps = scene([mat(color=A, amb=AMB, spec=spec, shin=1.0)], hquad(0.0, -50, 50, -50, 50, 0, normal=nrm),
                   [_sun(Ld, inten)])
        camera(ps, P0, (0, -1, 0), dist=1.0)
        assert float(nrm @ Ld) < 0 < float(nrm @ unit(Ld + np.array([0, 1, 0])))
        img = mode.render(ps)
        np.testing.assert_allclose(img, np.broadcast_to(A * AMB, img.shape), rtol=1e-6)

    run_pin(check, ["SPEC_NL", "SPEC_NO_GUARD", "NO_COS"])


def _shininess(b):
    r = max(np.float32(b) / np.float32(255.0), 0.02)
    return min(max(2.0 / r ** 4 - 2.0, 1.0), 4096.0)


def test_mask_map_roughness_sets_the_shininess():
    Ld = unit((math.sin(0.1), math.cos(0.1), 0.0))

    def check(mode):
        for b in (255, 200, 128, 64):
            tex = np.zeros((4, 4, 3), np.uint8)
            tex[..., 2] = b
            ps = _floor_scene([_sun(Ld)], m=mat(color=A, amb=AMB, spec=1.0, shin=7.0, mask_id=0), uv=True)
            texture(ps, [tex])
            img = mode.render(ps)
            want = _blinn(Ld, 1.0, 1.0, _shininess(b))
            np.testing.assert_allclose(img, np.broadcast_to(want, img.shape), rtol=2e-4, err_msg=str(b))

    run_pin(check, ["ROUGH_R2"])


def test_emission_is_added_unlit():
    E, e = np.array([1.0, 0.8, 0.45]), 3.0

    def check(mode):
        m = mat(color=A, amb=AMB, emission_color=E, emission=e)
        ps = _floor_scene([], m=m)
        img = mode.render(ps)
        np.testing.assert_allclose(img, np.broadcast_to(A * AMB + E * e, img.shape), rtol=1e-6)

    run_pin(check, ["EMISSION_LIT"])


@pytest.mark.gpu
def test_light_pins_gpu_matches_oracle():
    """The HIP path on the pin scenes equals the oracle within 1e-4 relative."""
    from pin_util import Mode
    scenes = [_floor_scene([light("point", position=p, intensity=i)]) for p, i in POINT_CASES]
    scenes.append(_floor_scene([light("point", position=(P0[0], 1.0, P0[2]), intensity=10.0, radius=1.0)]))
    scenes.append(_area_scene((1.3, 1.5, -0.2), (1.0, 0.0, 0.0), (0.0, 0.0, -0.8), 20.0))
    scenes.append(_floor_scene([light("spot", position=(P0[0] + 2 * math.tan(0.4), 2.0, P0[2]),
                                      direction=(0.0, -1.0, 0.0), angle=SPOT_ANGLE, blend=0.4, intensity=9.0)]))
    scenes.append(_floor_scene([_sun(unit((0.2, 0.9, -0.4)), 1.3)], m=mat(color=A, amb=AMB, spec=0.5, shin=32.0)))
    tex = np.zeros((4, 4, 3), np.uint8)
    tex[..., 2] = 128
    ps = _floor_scene([_sun(unit((math.sin(0.1), math.cos(0.1), 0.0)))],
                      m=mat(color=A, amb=AMB, spec=1.0, shin=7.0, mask_id=0), uv=True)
    texture(ps, [tex])
    scenes.append(ps)
    for k, ps in enumerate(scenes):
        size = (32, 32, 4) if k in (4, 5) else (4, 4, 1)
        got = gpu_render(ps, *size)
        ref = Mode().render(ps, *size)
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-7, err_msg="scene %d" % k)
