"""Mirror of the reference's ``sheen`` Python module.

Reference api/sheen_py_module.cpp:33-105: ``init(device_id, arch)``,
``cleanup()``, ``render_scene(scene_buffer, width, height, aa, path)`` with
RuntimeError(last error) on failure and TypeError for a non-buffer scene.
These three go through the CPython extension ``sheen_amd._sheen``; the
extensions below (float frames, primary hits, resident contexts, ray
counters) go through ctypes to the same libsheen_c_api.so.
"""
import ctypes as C

import numpy as np

from . import native

try:
    from . import _sheen as _ext
except ImportError as e:  # built by __graft_entry__.build()
    _ext = None
    _ext_error = e


def _need_ext():
    if _ext is None:
        raise ImportError("sheen_amd._sheen is not built: %s" % (_ext_error,))
    return _ext


def init(device_id=0, arch=0):
    _need_ext().init(int(device_id), int(arch))


def cleanup():
    _need_ext().cleanup()


def render_scene(scene, width, height, aa, output_path):
    _need_ext().render_scene(scene, int(width), int(height), int(aa), str(output_path))


def _scene_ptr(scene):
    if isinstance(scene, C._Pointer):
        return scene
    return C.pointer(scene)


class Context:
    """A scene resident on the GPU (BVHs built once) for repeated renders.
    The BVHs are built on the device; host_bvh=True builds the same trees
    with the host builder."""

    def __init__(self, scene, host_bvh=False):
        lib = native.capi()
        self._lib = lib
        self._c = lib.sheen_context_create_ex(_scene_ptr(scene), 1 if host_bvh else 0)
        if not self._c:
            raise RuntimeError(native.last_error())

    def bvh(self, category):
        """(nodes (N, 128) uint8, prims (P,) uint32, max_stack, depth) of the
        wide BVH of one primitive kind (0 spheres ... 6 prisms)."""
        info = (C.c_int * 4)()
        native.check(self._lib.sheen_context_bvh(self._c, category, None, 0, None, 0, info))
        nodes = np.zeros((info[0], 128), np.uint8)
        prims = np.zeros(info[1], np.uint32)
        native.check(self._lib.sheen_context_bvh(self._c, category, native.as_ptr(nodes), info[0], native.as_ptr(prims),
                                                 info[1], info))
        return nodes, prims, info[2], info[3]

    @property
    def build_ms(self):
        return self._lib.sheen_context_build_ms(self._c)

    def render(self, width, height, aa, download=True):
        st = native.SheenRayStats()
        out = np.empty((height, width, 4), np.float32) if download else None
        native.check(self._lib.sheen_context_render(self._c, width, height, aa, native.as_ptr(out), C.byref(st)))
        return out, st.as_dict()

    def primary_hits(self, width, height, aa):
        n = width * height * aa
        kind = np.empty(n, np.int32)
        idx = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        native.check(self._lib.sheen_context_primary_hits(self._c, width, height, aa, native.as_ptr(kind),
                                                          native.as_ptr(idx), native.as_ptr(t)))
        return kind.reshape(height, width, aa), idx.reshape(height, width, aa), t.reshape(height, width, aa)

    def set_profiling(self, on):
        self._lib.sheen_context_set_profiling(self._c, int(bool(on)))

    def set_counting(self, on):
        self._lib.sheen_context_set_counting(self._c, int(bool(on)))

    def set_chunk(self, samples):
        self._lib.sheen_context_set_chunk(self._c, int(samples))

    def set_param(self, name, value):
        native.check(self._lib.sheen_context_set_param(self._c, name.encode(), int(value)))

    def photon_stats(self):
        e, s = C.c_int(), C.c_int()
        self._lib.sheen_context_photon_stats(self._c, C.byref(e), C.byref(s))
        return e.value, s.value

    def photon_probe(self, stride, radius, max_samples=4096, brute=False):
        n = self.photon_stats()[1]
        counts = np.zeros(max(1, (n + stride - 1) // stride), np.int32)
        k = self._lib.sheen_context_photon_probe(self._c, stride, radius, max_samples, int(bool(brute)),
                                                 native.as_ptr(counts))
        if k < 0:
            raise RuntimeError(native.last_error())
        return counts[:k]

    def frame_device_ptr(self):
        return self._lib.sheen_context_frame_device(self._c)

    def close(self):
        if self._c:
            self._lib.sheen_context_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def render_float(scene, width, height, aa):
    """Linear float RGBA frame (height, width, 4) of one render."""
    out = np.empty((height, width, 4), np.float32)
    st = native.SheenRayStats()
    native.check(native.capi().sheen_render_float(_scene_ptr(scene), width, height, aa, native.as_ptr(out),
                                                  C.byref(st)))
    return out, st.as_dict()


def encode_srgb(rgba):
    h, w = rgba.shape[:2]
    rgba = np.ascontiguousarray(rgba, np.float32)
    out = np.empty((h, w, 3), np.uint8)
    native.capi().sheen_encode_srgb(native.as_ptr(rgba), w, h, native.as_ptr(out))
    return out
