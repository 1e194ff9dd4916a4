"""The reference's scene files as stored in tests/golden/reference_scenes.npz
(written by tests/golden/make_reference_scenes.py): rebuilt into SheenScene
records for the oracle and the GPU on machines without the reference tree."""
import ctypes as C
import os

import numpy as np

from sheen_amd import contract as K

# (pointer field, count field, record type) of the stored arrays
ARRAYS = [("objects", "num_objects", K.Object), ("sweep_points", "num_sweep_points", K.SweepPoint),
          ("sweep_strands", "num_sweep_strands", K.SweepStrand), ("triangles", "num_triangles", K.Triangle),
          ("heightfields", "num_heightfields", K.HeightField),
          ("superellipsoids", "num_superellipsoids", K.Superellipsoid),
          ("lathe_points", "num_lathe_points", K.LathePoint), ("lathes", "num_lathes", K.Lathe),
          ("prism_points", "num_prism_points", K.PrismPoint), ("prisms", "num_prisms", K.Prism),
          ("csg_nodes", "num_csg_nodes", K.CSGNode), ("planes", "num_planes", K.Plane),
          ("materials", "num_materials", K.Material), ("lights", "num_lights", K.Light),
          ("media", "num_media", K.Medium)]

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_scenes.npz")
_DATA = None


def _data():
    global _DATA
    if _DATA is None:
        _DATA = np.load(PATH, allow_pickle=False)
    return _DATA


def names():
    return [str(n) for n in _data()["names"]]


class RefScene:
    """`.ptr` (POINTER(SheenScene)), `.width`, `.height`, `.aa`."""

    def __init__(self, name):
        d = _data()
        self.name = name
        self.keep = []
        s = K.SheenScene()
        C.memmove(C.byref(s.camera), d[name + "/camera"].tobytes(), C.sizeof(K.CameraData))
        w, h, aa, csg_root, photons = (int(v) for v in d[name + "/settings"])
        self.width, self.height, self.aa = w, h, aa
        s.csg_root, s.photons = csg_root, photons
        s.photon_cell, s.caustics = (float(v) for v in d[name + "/floats"])
        for field, count, typ in ARRAYS:
            raw = d["%s/%s" % (name, field)]
            n = raw.size // C.sizeof(typ)
            if n == 0:
                continue
            arr = (typ * n).from_buffer_copy(raw.tobytes())
            self.keep.append(arr)
            setattr(s, field, C.cast(arr, C.POINTER(typ)))
            setattr(s, count, n)
        for i in range(s.num_heightfields):
            key = "%s/hf_image%d" % (name, i)
            if key in d:
                img = (C.c_ubyte * d[key].size).from_buffer_copy(d[key].tobytes())
                self.keep.append(img)
                s.heightfields[i].image = C.cast(img, C.POINTER(C.c_ubyte))
            else:
                s.heightfields[i].image = None
        texs = []
        while "%s/tex%d" % (name, len(texs)) in d:
            texs.append(d["%s/tex%d" % (name, len(texs))])
        if texs:
            ptrs = []
            for t in texs:
                buf = (C.c_ubyte * t.size).from_buffer_copy(t.tobytes())
                self.keep.append(buf)
                ptrs.append(C.cast(buf, C.POINTER(C.c_ubyte)))
            for arr, field in (((C.POINTER(C.c_ubyte) * len(ptrs))(*ptrs), "texture_pixels"),
                               ((C.c_int * len(texs))(*[t.shape[1] for t in texs]), "texture_widths"),
                               ((C.c_int * len(texs))(*[t.shape[0] for t in texs]), "texture_heights")):
                self.keep.append(arr)
                setattr(s, field, C.cast(arr, C.POINTER(arr._type_)))
            s.num_textures = len(texs)
        self.s = s

    @property
    def ptr(self):
        return C.pointer(self.s)

    @property
    def scene(self):
        return self.ptr
