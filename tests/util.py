"""Helpers for building small scenes from Python (ctypes records)."""
import ctypes as C

import numpy as np

from sheen_amd import contract as K


def f3(x, y, z):
    return K.Float3(float(x), float(y), float(z))


def material(color=(0.8, 0.8, 0.8), spec=0.0, shin=32.0, amb=0.1):
    m = K.Material()
    m.diffuse_color = f3(*color)
    m.specular, m.shininess, m.ambient = spec, shin, amb
    m.texture_id = m.mask_id = m.normal_id = -1
    m.hair_shift, m.ior, m.normal_strength = 0.1, 1.45, 1.0
    return m


class PyScene:
    """Keeps the ctypes arrays alive and exposes .ptr (POINTER(SheenScene))."""

    def __init__(self):
        self.s = K.SheenScene()
        self.s.csg_root = -1
        self.keep = []
        cam = self.s.camera
        cam.position, cam.forward, cam.right, cam.up = f3(0, 0, 5), f3(0, 0, -1), f3(1, 0, 0), f3(0, 1, 0)

    def _arr(self, typ, items):
        a = (typ * len(items))(*items)
        self.keep.append(a)
        return a

    def set(self, field, typ, items, count_field):
        if not items:
            return
        setattr(self.s, field, C.cast(self._arr(typ, items), C.POINTER(typ)))
        setattr(self.s, count_field, len(items))

    @property
    def ptr(self):
        return C.pointer(self.s)


def spheres(centers, radii, mats):
    out = []
    for c, r, m in zip(centers, radii, mats):
        o = K.Object()
        o.type, o.material_id, o.param1, o.param2 = 0, int(m), f3(*c), float(r)
        out.append(o)
    return out


def triangles(verts, mats=None):
    out = []
    for i, (a, b, c) in enumerate(verts):
        t = K.Triangle()
        t.v0, t.v1, t.v2 = f3(*a), f3(*b), f3(*c)
        t.material_id = 0 if mats is None else int(mats[i])
        out.append(t)
    return out


def rand_dirs(rng, n):
    d = rng.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
