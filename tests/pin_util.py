"""Scene builders and the self-check of closed-form pins of the oracle.

A pin compares oracle renders with a closed form.  `run_pin(check, plants)`
runs `check(mode)` three ways: with liboracle.so (it must pass), with
liboracle_plant.so and no mistake switched on (it must pass), and with each
listed planted mistake switched on (it must fail) — so each pin is shown to
catch the mistakes it is meant to catch.
"""
import ctypes as C
import os

import numpy as np

from sheen_amd import contract as K
from util import PyScene, f3, material

# planted mistakes of oracle/sheen_oracle.cpp (enum Plant)
P = dict(POINT_DIST=1, NO_COS=2, TWO_PI=3, DISK_LINEAR=4, AREA_NO_COS=5, AREA_SIDE=6, AREA_HALF=7, SPOT_RAMP=8,
         SPOT_HALF_ANGLE=9, SPEC_NL=10, SPEC_NO_GUARD=11, ROUGH_R2=12, IOR_RED=13, TRANS_NO_MASK=14, SKY_X=15,
         SKY_SWAP=16, REFL_NO_CLAMP=17, REFL_SWAP=18, ETA_FIXED=19, TIR_SIGN=20, W_ONCE=21, SHADOW_TMAX=22,
         SHADOW_4=23, CHORD_HALF=24, RADIUS_REV=25, TAU_NO_STOP=26, HF_HASH=27, HF_DETAIL=28, HF_LINEAR=29,
         HF_OCT_SEED=30, SIGMA=31, SHADOW_THRESH=32, EMISSION_LIT=33, BEAM_NO_SHADOW=34)

THREADS = min(8, os.cpu_count() or 1)
FOCAL, SENSOR = 1.0e7, 36.0  # near-orthographic: rays within 2e-6 rad of the axis


class Mode:
    """Which oracle build renders, and which planted mistake is on."""

    def __init__(self, plant=None):
        self.plant = plant

    def __repr__(self):
        return "oracle" if self.plant is None else "plant %d" % self.plant

    def oracle(self, ps):
        from oracle import Oracle, lib
        if self.plant is not None:
            lib(True).oracle_plant(int(self.plant))
        ptr = ps.ptr if hasattr(ps, "ptr") else ps.scene  # a PyScene or a loaded/generated scene
        return Oracle(ptr, threads=THREADS, plant=self.plant is not None)

    def render(self, ps, w=4, h=4, aa=1, hits=False, raw=False):
        """RGB as float64; raw: the oracle's float32 RGBA image; hits: the
        oracle's whole result."""
        from oracle import lib
        try:
            out = self.oracle(ps).render(w, h, aa, hits=hits)
        finally:
            if self.plant is not None:
                lib(True).oracle_plant(0)
        if hits:
            return out
        return out["image"] if raw else out["image"][..., :3].astype(np.float64)

    def lib(self):
        from oracle import lib
        L = lib(self.plant is not None)
        if self.plant is not None:
            L.oracle_plant(int(self.plant))
        return L

    def done(self):
        from oracle import lib
        if self.plant is not None:
            lib(True).oracle_plant(0)


def run_pin(check, plants):
    check(Mode())
    check(Mode(0))
    for k in plants:
        kid = P[k] if isinstance(k, str) else k
        try:
            check(Mode(kid))
        except AssertionError:
            continue
        finally:
            Mode(kid).done()
        raise AssertionError("the pin did not catch planted mistake %s" % k)


# --------------------------------------------------------------------------
# scene pieces
# --------------------------------------------------------------------------
def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def f32(v):
    return np.asarray(v, np.float32).astype(np.float64)


def mat(color=(0.8, 0.6, 0.4), amb=0.1, spec=0.0, shin=32.0, **kw):
    m = material(color=color, spec=spec, shin=shin, amb=amb)
    for k, v in kw.items():
        if k == "emission_color":
            m.emission_color = f3(*v)
        else:
            setattr(m, k, v)
    return m


def tri(a, b, c, m, uv=None, normal=None):
    """A triangle; normal (one vertex normal for all three corners) makes the
    shading normal differ from the face normal."""
    t = K.Triangle()
    t.v0, t.v1, t.v2 = f3(*a), f3(*b), f3(*c)
    t.n0 = t.n1 = t.n2 = f3(*(normal if normal is not None else (0, 0, 0)))
    if uv is not None:
        t.uv0.x, t.uv0.y = uv[0]
        t.uv1.x, t.uv1.y = uv[1]
        t.uv2.x, t.uv2.y = uv[2]
    t.material_id = int(m)
    return t


def quad(p0, p1, p2, p3, m, uv=False, normal=None):
    """Two triangles over the corners p0..p3 (in order around the quad); with
    uv, corners map to (0,0), (1,0), (1,1), (0,1)."""
    uvs = ((0, 0), (1, 0), (1, 1), (0, 1))
    return [tri(p0, p1, p2, m, (uvs[0], uvs[1], uvs[2]) if uv else None, normal),
            tri(p0, p2, p3, m, (uvs[0], uvs[2], uvs[3]) if uv else None, normal)]


def hquad(y, x0, x1, z0, z1, m, uv=False, normal=None):
    """Horizontal quad at height y over [x0, x1] x [z0, z1]."""
    return quad((x0, y, z0), (x1, y, z0), (x1, y, z1), (x0, y, z1), m, uv, normal)


def camera(ps, target, direction, dist=1.0, focal=FOCAL):
    """Camera at target - direction * dist looking along direction."""
    d = unit(direction)
    up0 = np.array([0.0, 0.0, -1.0]) if abs(d[1]) > 0.9 else np.array([0.0, 1.0, 0.0])
    right = unit(np.cross(d, up0))
    up = np.cross(right, d)
    cam = ps.s.camera
    pos = np.asarray(target, np.float64) - d * dist
    cam.position, cam.forward, cam.right, cam.up = f3(*pos), f3(*d), f3(*right), f3(*up)
    cam.focal_length, cam.sensor_width = focal, SENSOR


def light(kind, **kw):
    L = K.Light()
    L.type = {"sun": 0, "point": 1, "area": 2, "spot": 3}[kind]
    L.color = f3(*kw.pop("color", (1.0, 1.0, 1.0)))
    L.intensity = kw.pop("intensity", 1.0)
    L.flags = 1
    L.texture_id = -1
    for k, v in kw.items():
        setattr(L, k, f3(*v) if isinstance(v, (tuple, list, np.ndarray)) else v)
    return L


def texture(ps, img):
    """Adds RGB uint8 textures (list of (h, w, 3) arrays) to a scene."""
    ptrs, ws, hs = [], [], []
    for a in img:
        a = np.ascontiguousarray(a, np.uint8)
        buf = (C.c_ubyte * a.size).from_buffer_copy(a.tobytes())
        ps.keep.append(buf)
        ptrs.append(C.cast(buf, C.POINTER(C.c_ubyte)))
        hs.append(a.shape[0])
        ws.append(a.shape[1])
    ps.s.texture_pixels = ps._arr(C.POINTER(C.c_ubyte), ptrs)
    ps.s.texture_widths = ps._arr(C.c_int, ws)
    ps.s.texture_heights = ps._arr(C.c_int, hs)
    ps.s.num_textures = len(img)


def scene(mats, tris=(), lights=(), objects=(), strands=None):
    ps = PyScene()
    ps.set("materials", K.Material, list(mats), "num_materials")
    ps.set("triangles", K.Triangle, list(tris), "num_triangles")
    ps.set("lights", K.Light, list(lights), "num_lights")
    ps.set("objects", K.Object, list(objects), "num_objects")
    if strands:
        pts, st = strands
        ps.set("sweep_points", K.SweepPoint, pts, "num_sweep_points")
        ps.set("sweep_strands", K.SweepStrand, st, "num_sweep_strands")
    return ps


def gpu_render(ps, w=4, h=4, aa=1):
    from sheen_amd import sheen
    with sheen.Context(ps.ptr) as ctx:
        img, _ = ctx.render(w, h, aa)
    return img[..., :3].astype(np.float64)
