"""The oracle's parameter substitution, which the gradient checks of the
differentiable entry points use as their FP64 reference."""
import ctypes as C

import numpy as np

from adjoint_util import scene_params


def test_unchanged_parameters_reproduce_the_render(oracle_mod, scenes_mod):
    sc = scenes_mod.generate(5, detail=300)
    p = scene_params(sc.scene)
    a = oracle_mod.Oracle(sc.scene, threads=8).render(48, 48, 2)["image"]
    o = oracle_mod.Oracle(sc.scene, threads=8)
    o.set_params(p["albedo"], p["emission"], p["intensity"], None)
    o.set_params(p["albedo"], p["emission"], p["intensity"], None)  # repeated calls start from the scene
    b = o.render(48, 48, 2)["image"]
    assert np.array_equal(a, b)


def test_emission_derivative_is_the_hit_fraction(oracle_mod, scenes_mod):
    """d pixel / d emission[m] = (samples of the pixel whose primary hit has
    material m) / spp, exactly (emission is added once per primary hit).  The
    pixel is linear in the parameter, so a central difference with a large
    step is exact up to the float32 rounding of the output image."""
    sc = scenes_mod.generate(5, detail=300)
    s = sc.scene.contents
    p = scene_params(sc.scene)
    W = H = 40
    aa = 2
    o = oracle_mod.Oracle(sc.scene, threads=8)
    hits = o.render(W, H, aa, hits=True)
    kind, idx = hits["kind"], hits["index"]
    # material of each primary hit: spheres (kind 0) and planes (kind 8)
    mat = np.full(kind.shape, -1)
    sph = kind == 0
    mat[sph] = [s.objects[i].material_id for i in idx[sph]]
    pl = kind == 8
    mat[pl] = [s.planes[i].material_id for i in idx[pl]]
    m = int(np.bincount(mat[mat >= 0]).argmax())
    h = 0.05
    imgs = []
    for sgn in (1, -1):
        em = p["emission"].copy()
        em[m, 1] += sgn * h
        o.set_params(p["albedo"], em, p["intensity"], None)
        imgs.append(o.render(W, H, aa, fp64=True)["image"].astype(np.float64))
    em_lo, em_hi = np.float32(p["emission"][m, 1] - h), np.float32(p["emission"][m, 1] + h)
    d = (imgs[0][..., 1] - imgs[1][..., 1]) / (float(em_hi) - float(em_lo))
    frac = (mat == m).sum(-1) / aa
    np.testing.assert_allclose(d, frac, atol=1e-5)
    assert np.allclose(imgs[0][..., 0], imgs[1][..., 0])


def _bilinear(lvl, ux, uy):
    """The bilinear tap (reference integrator/shading.h:26-58) in float32."""
    f = np.float32
    h, w = lvl.shape[:2]
    u = f(ux) - f(np.floor(f(ux)))
    v = f(1.0) - (f(uy) - f(np.floor(f(uy))))
    fx, fy = u * f(w - 1), v * f(h - 1)
    x0, y0 = int(fx), int(fy)
    x1, y1 = (x0 + 1) % w, (y0 + 1) % h
    tx, ty = fx - f(x0), fy - f(y0)
    a = lvl[y0, x0] * (f(1) - tx) + lvl[y0, x1] * tx
    b = lvl[y1, x0] * (f(1) - tx) + lvl[y1, x1] * tx
    return a * (f(1) - ty) + b * ty


def test_float_mips_are_two_by_two_means(oracle_mod, scenes_mod):
    sc = scenes_mod.generate(102)
    s = sc.scene.contents
    p = scene_params(sc.scene)
    rng = np.random.default_rng(3)
    tex = rng.random(p["texels"].shape).astype(np.float32)
    o = oracle_mod.Oracle(sc.scene, threads=1)
    o.set_params(None, None, None, tex)
    w, h = s.texture_widths[0], s.texture_heights[0]
    lvl = tex[: w * h * 3].reshape(h, w, 3).astype(np.float32)
    for level in range(0, 5):
        if level:
            lvl = ((lvl[0::2, 0::2] + lvl[0::2, 1::2]) + (lvl[1::2, 0::2] + lvl[1::2, 1::2])) * np.float32(0.25)
        for ux, uy in rng.random((6, 2)).astype(np.float32):
            out = np.zeros(3, np.float32)
            o._L.oracle_sample_texture(o._h, 0, C.c_float(ux), C.c_float(uy), C.c_float(level),
                                       out.ctypes.data_as(C.c_void_p))
            np.testing.assert_array_equal(out, _bilinear(lvl, ux, uy))
