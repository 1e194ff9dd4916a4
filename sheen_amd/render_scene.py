"""Render a .sheen scene file on the GPU and write a PNG.

    python -m sheen_amd.render_scene [scene.sheen] [--out file.png]
                                     [--width N] [--height N] [--aa N]
                                     [--repeat K]

Counterpart of the reference's scene_render tool (tools/render_scene.cu:14-31
for the arguments: command-line values override the file's `resolution`
line; the default scene is scenes/demo.sheen).  Prints the scene's element
counts, the photon map when the file asks for one, the time of one warm-up
frame and the mean of `--repeat` timed frames of the resident scene.
"""
import argparse
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m sheen_amd.render_scene")
    ap.add_argument("scene", nargs="?", default="scenes/demo.sheen")
    ap.add_argument("--out")
    ap.add_argument("--width", type=int, default=0)
    ap.add_argument("--height", type=int, default=0)
    ap.add_argument("--aa", type=int, default=0)
    ap.add_argument("--repeat", type=int, default=5)
    a = ap.parse_args(argv)

    from . import native, scenes, sheen
    try:
        sf = scenes.load_sheen(a.scene)
    except ValueError as e:
        print("%s: %s" % (a.scene, e), file=sys.stderr)
        return 1
    w = a.width if a.width > 0 else sf.width
    h = a.height if a.height > 0 else sf.height
    aa = a.aa if a.aa > 0 else sf.aa
    out = a.out or sf.output
    c = sf.scene.contents
    print("scene: %s" % a.scene)
    print("resolution: %dx%d, AA %d" % (w, h, aa))
    print("objects %d | triangles %d | strands %d | points %d | planes %d | lights %d" % (
        c.num_objects, c.num_triangles, c.num_sweep_strands, c.num_sweep_points, c.num_planes, c.num_lights))
    with sheen.Context(sf.scene) as ctx:
        if c.photons > 0 and c.caustics > 0:
            emitted, stored = ctx.photon_stats()
            print("photons: %d emitted, %d stored" % (emitted, stored))
        t0 = time.perf_counter()
        img, st = ctx.render(w, h, aa)
        first = time.perf_counter() - t0
        times = []
        for _ in range(max(0, a.repeat)):
            _, s = ctx.render(w, h, aa, download=False)
            times.append(s["frame_ms"])
    print("first frame %.2f ms (BVH upload and build included)" % (first * 1e3))
    if times:
        print("render: %.3f ms/frame over %d frames, %.1f Mrays/s" % (
            float(np.mean(times)), len(times), st["total_rays"] / (float(np.mean(times)) * 1e3)))
    if out:
        rgb = np.ascontiguousarray(sheen.encode_srgb(img))
        native.check(native.capi().sheen_write_png(out.encode(), w, h, native.as_ptr(rgb)))
        print("wrote %s" % out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
