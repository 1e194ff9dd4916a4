"""ctypes binding of oracle/liboracle.so (test infrastructure only)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
PLANT_LIB_PATH = os.path.join(HERE, "liboracle_plant.so")
_libs = {}


class OracleStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("primary_rays", "shadow_rays", "reflection_rays", "refraction_rays",
                                          "photon_rays", "total_rays", "samples")] + \
               [("photons_emitted", C.c_int), ("photons_stored", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(plant=False):
    """liboracle.so, or with plant=True liboracle_plant.so: the same oracle
    with the planted mistakes that oracle_plant(k) switches on (test
    self-checks only)."""
    if plant in _libs:
        return _libs[plant]
    path = PLANT_LIB_PATH if plant else LIB_PATH
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    P = C.c_void_p
    L.oracle_scene_create.restype = P
    L.oracle_scene_create.argtypes = [P, C.c_int]
    L.oracle_scene_destroy.argtypes = [P]
    L.oracle_render.restype = C.c_int
    L.oracle_render.argtypes = [P] + [C.c_int] * 7 + [P, P, P, P, C.POINTER(OracleStats)]
    L.oracle_trace.restype = C.c_int
    L.oracle_trace.argtypes = [P, C.c_int, P, P, C.c_float, C.c_float, C.c_int, P, P, P, P, P, P]
    L.oracle_trace_brute.restype = C.c_int
    L.oracle_trace_brute.argtypes = [P, C.c_int, P, P, C.c_float, C.c_float, C.c_int, P, P]
    L.oracle_aa_hash.restype = C.c_uint32
    L.oracle_aa_hash.argtypes = [C.c_uint32]
    L.oracle_aa_rand.restype = C.c_float
    L.oracle_aa_rand.argtypes = [C.c_uint32]
    L.oracle_sample_offset.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.oracle_hf_eval.restype = C.c_float
    L.oracle_hf_eval.argtypes = [P, C.c_float, C.c_float]
    L.oracle_hf_cell_hash.restype = C.c_float
    L.oracle_hf_cell_hash.argtypes = [C.c_int, C.c_int, C.c_int]
    L.oracle_hf_value_noise.restype = C.c_float
    L.oracle_hf_value_noise.argtypes = [C.c_float, C.c_float, C.c_int]
    L.oracle_math.restype = C.c_float
    L.oracle_math.argtypes = [C.c_int, C.c_float, C.c_float]
    L.oracle_hf_fbm.restype = C.c_float
    L.oracle_hf_fbm.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, C.c_int]
    if plant:
        L.oracle_plant.argtypes = [C.c_int]
    L.oracle_photon_cell_hash.restype = C.c_int
    L.oracle_photon_cell_hash.argtypes = [C.c_int] * 4
    L.oracle_build_mips.restype = C.c_int
    L.oracle_build_mips.argtypes = [P, C.c_int, C.c_int, P, C.POINTER(C.c_int)]
    L.oracle_sample_texture.restype = C.c_int
    L.oracle_sample_texture.argtypes = [P, C.c_int, C.c_float, C.c_float, C.c_float, P]
    L.oracle_photon_stats.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.oracle_photon_map.restype = C.c_int
    L.oracle_photon_map.argtypes = [P, P, P, P]
    L.oracle_photon_probe.restype = C.c_int
    L.oracle_photon_probe.argtypes = [P, C.c_int, C.c_float, C.c_int, P]
    L.oracle_photon_probe_brute.restype = C.c_int
    L.oracle_photon_probe_brute.argtypes = [P, C.c_int, C.c_float, C.c_int, P]
    L.oracle_set_params.argtypes = [P, P, P, P, P]
    L.oracle_render_reference.argtypes = [P] + [C.c_int] * 10 + [P]
    _libs[plant] = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Oracle:
    """CPU restatement of the render path over one scene (SheenScene pointer)."""

    def __init__(self, scene_ptr, threads=None, plant=False):
        self.threads = threads or os.cpu_count() or 1
        self._L = lib(plant)
        # the oracle's scene copy points into the caller's record arrays (it
        # copies the SheenScene struct only): keep the scene object alive
        self._scene_ref = scene_ptr
        self._h = self._L.oracle_scene_create(C.cast(scene_ptr, C.c_void_p), self.threads)
        if not self._h:
            raise RuntimeError("oracle scene creation failed")

    def render(self, width, height, aa, rows=None, fp64=False, hits=False):
        y0, y1 = rows if rows else (0, height)
        img = np.zeros((height, width, 4), np.float32)
        st = OracleStats()
        hk = hi = ht = None
        if hits:
            n = width * height * aa
            hk = np.full(n, -2, np.int32)
            hi = np.full(n, -2, np.int32)
            ht = np.zeros(n, np.float32)
        rc = self._L.oracle_render(self._h, width, height, aa, y0, y1, self.threads, int(fp64), _p(img), _p(hk),
                                   _p(hi), _p(ht), C.byref(st))
        if rc != 0:
            raise RuntimeError("oracle render failed")
        out = {"image": img, "stats": st.as_dict()}
        if hits:
            out["kind"] = hk.reshape(height, width, aa)
            out["index"] = hi.reshape(height, width, aa)
            out["t"] = ht.reshape(height, width, aa)
        return out

    def trace(self, orig, dirs, tmin=0.001, tmax=1e10, opaque=False):
        orig = np.ascontiguousarray(orig, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(orig)
        kind = np.empty(n, np.int32)
        idx = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        pt = np.empty((n, 3), np.float32)
        nr = np.empty((n, 3), np.float32)
        mat = np.empty(n, np.int32)
        self._L.oracle_trace(self._h, n, _p(orig), _p(dirs), tmin, tmax, int(opaque), _p(kind), _p(idx), _p(t),
                             _p(pt), _p(nr), _p(mat))
        return {"kind": kind, "index": idx, "t": t, "point": pt, "normal": nr, "material": mat}

    def trace_brute(self, orig, dirs, category, tmin=0.001, tmax=1e10):
        orig = np.ascontiguousarray(orig, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(orig)
        idx = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        self._L.oracle_trace_brute(self._h, n, _p(orig), _p(dirs), tmin, tmax, category, _p(idx), _p(t))
        return idx, t

    def photon_stats(self):
        e, s = C.c_int(), C.c_int()
        self._L.oracle_photon_stats(self._h, C.byref(e), C.byref(s))
        return e.value, s.value

    def photon_probe(self, stride, radius, max_samples=4096, brute=False):
        n = self.photon_stats()[1]
        out = np.zeros(max(1, (n + stride - 1) // stride), np.int32)
        f = self._L.oracle_photon_probe_brute if brute else self._L.oracle_photon_probe
        k = f(self._h, stride, radius, max_samples, _p(out))
        return out[:k]

    def render_reference(self, width, height, window, sub=8, aa=64, fp64=True):
        """Radiance of window (x0, y0, w, h) of the width x height image with
        sub*sub*aa samples per pixel over its footprint; (h, w, 3) float64."""
        x0, y0, cw, ch = window
        out = np.zeros((ch, cw, 3), np.float64)
        rc = self._L.oracle_render_reference(self._h, width, height, x0, y0, cw, ch, sub, aa, self.threads, int(fp64),
                                             _p(out))
        if rc != 0:
            raise RuntimeError("oracle reference render failed")
        return out

    def set_params(self, albedo=None, emission=None, intensity=None, texels=None):
        """Substitutes the differentiable parameters (float32 arrays: albedo and
        emission (materials, 3), intensity (lights,), texels (level 0 of every
        texture, RGB / 255))."""
        import numpy as np
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float32) for a in (albedo, emission, intensity, texels)]
        ptrs = [None if a is None else a.ctypes.data_as(C.c_void_p) for a in keep]
        self._L.oracle_set_params(self._h, *ptrs)

    def close(self):
        if self._h:
            self._L.oracle_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
