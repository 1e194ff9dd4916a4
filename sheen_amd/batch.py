"""Batches of scenes rendered as one wavefront pipeline (sheen_batch_*)."""
import ctypes as C

import numpy as np

from . import native
from .sheen import _scene_ptr


class Batch:
    """`scenes`: SheenScene pointers (or GeneratedScene objects); `sizes`:
    (width, height, aa) per scene (default: each GeneratedScene's own)."""

    def __init__(self, scenes, sizes=None):
        lib = native.capi()
        self._lib = lib
        ptrs = [s.scene if hasattr(s, "scene") else _scene_ptr(s) for s in scenes]
        if sizes is None:
            sizes = [(s.width, s.height, s.aa) for s in scenes]
        self.sizes = [tuple(int(v) for v in z) for z in sizes]
        n = len(ptrs)
        arr = (C.c_void_p * n)(*[C.cast(p, C.c_void_p) for p in ptrs])
        w = (C.c_int * n)(*[z[0] for z in self.sizes])
        h = (C.c_int * n)(*[z[1] for z in self.sizes])
        a = (C.c_int * n)(*[z[2] for z in self.sizes])
        self._b = lib.sheen_batch_create(arr, w, h, a, n)
        if not self._b:
            raise RuntimeError(native.last_error())
        sc, smp, ch = C.c_int(), C.c_longlong(), C.c_int()
        native.check(lib.sheen_batch_info(self._b, C.byref(sc), C.byref(smp), C.byref(ch)))
        self.n, self.samples, self.chunks = sc.value, smp.value, ch.value

    def set_param(self, name, value):
        native.check(self._lib.sheen_batch_set_param(self._b, name.encode(), int(value)))
        ch = C.c_int()
        native.check(self._lib.sheen_batch_info(self._b, None, None, C.byref(ch)))
        self.chunks = ch.value

    def render(self):
        st = native.SheenRayStats()
        native.check(self._lib.sheen_batch_render(self._b, C.byref(st)))
        return st.as_dict()

    def render_serial(self):
        st = native.SheenRayStats()
        native.check(self._lib.sheen_batch_render_serial(self._b, C.byref(st)))
        return st.as_dict()

    def image(self, k, serial=False):
        w, h, _ = self.sizes[k]
        out = np.empty((h, w, 4), np.float32)
        native.check(self._lib.sheen_batch_image(self._b, k, int(bool(serial)), native.as_ptr(out)))
        return out

    def close(self):
        if self._b:
            self._lib.sheen_batch_destroy(self._b)
            self._b = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _main(argv=None):
    """`python -m sheen_amd.batch [batch|serial] [reps] [scenes]`: renders the
    generated uneven batch (256 scenes by default) `reps` times, as one
    pipeline or one render per scene, printing the GPU milliseconds of each
    (a workload to run under a profiler)."""
    import argparse
    from . import scenes as gen
    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("mode", nargs="?", default="batch", choices=("batch", "serial"))
    ap.add_argument("reps", nargs="?", type=int, default=3)
    ap.add_argument("scenes", nargs="?", type=int, default=256)
    a = ap.parse_args(argv)
    with Batch(gen.batch(a.scenes)) as b:
        print("batch of %d scenes, %d samples, %d chunks" % (b.n, b.samples, b.chunks), flush=True)
        for i in range(a.reps):
            s = b.render() if a.mode == "batch" else b.render_serial()
            print(a.mode, i, "%.2f ms" % s["trace_ms"], s["total_rays"], "rays", flush=True)


if __name__ == "__main__":
    _main()
