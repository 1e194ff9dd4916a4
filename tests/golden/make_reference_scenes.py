"""Writes tests/golden/reference_scenes.npz: the reference's own scene files
(reference scenes/*.sheen), parsed by this repository's loader and stored as
the SheenScene record arrays they produce (inputs only, no expected
outputs: the tests render them with the oracle and the GPU), so that the
GPU box, where the reference tree is absent, can render them.

    python tests/golden/make_reference_scenes.py

Scenes whose assets are missing from the reference tree (hf_image.sheen,
granite_vein.sheen name files under /tmp) are left out.
"""
import ctypes as C
import glob
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from ref_scenes import ARRAYS  # noqa: E402
from sheen_amd import scenes  # noqa: E402

REF_SCENES = "/root/reference/scenes"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_scenes.npz")

def _bytes(ptr, typ, n):
    if not ptr or n <= 0:
        return np.zeros(0, np.uint8)
    return np.frombuffer(C.string_at(C.cast(ptr, C.c_void_p), C.sizeof(typ) * n), np.uint8).copy()


def serialize(name, f, out):
    s = f.scene.contents
    out[name + "/camera"] = np.frombuffer(bytes(s.camera), np.uint8).copy()
    out[name + "/settings"] = np.array([f.width, f.height, f.aa, s.csg_root, s.photons], np.int64)
    out[name + "/floats"] = np.array([s.photon_cell, s.caustics], np.float32)
    for field, count, typ in ARRAYS:
        out["%s/%s" % (name, field)] = _bytes(getattr(s, field), typ, getattr(s, count))
    for i in range(s.num_heightfields):  # height-map images hang off the records
        hf = s.heightfields[i]
        if hf.image and hf.img_w > 0 and hf.img_h > 0:
            out["%s/hf_image%d" % (name, i)] = np.frombuffer(C.string_at(hf.image, hf.img_w * hf.img_h),
                                                              np.uint8).copy()
    for t in range(s.num_textures):
        w, h = s.texture_widths[t], s.texture_heights[t]
        out["%s/tex%d" % (name, t)] = np.frombuffer(C.string_at(s.texture_pixels[t], w * h * 3),
                                                     np.uint8).reshape(h, w, 3).copy()


def main():
    out = {}
    names = []
    for path in sorted(glob.glob(os.path.join(REF_SCENES, "*.sheen"))):
        name = os.path.basename(path)[:-6]
        try:
            f = scenes.load_sheen(path)
        except ValueError as e:
            print("skip %s: %s" % (name, e))
            continue
        serialize(name, f, out)
        names.append(name)
        print("%-14s %dx%d aa %d" % (name, f.width, f.height, f.aa))
    out["names"] = np.array(names)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
