"""Seeded synthetic workloads (include/sheen_scenes.h) as Python objects."""
import ctypes as C

from . import native

CONFIGS = {1: "spheres-64-256x256", 2: "mesh-1M-1080p", 4: "batch-scene", 5: "shading-10k-512x512",
           100: "parity-lights", 101: "parity-mesh", 102: "parity-textures"}


class GeneratedScene:
    """Owns a generated scene; ``.scene`` is a ctypes SheenScene pointer."""

    def __init__(self, config, **overrides):
        lib = native.host()
        p = native.SheenSceneParams()
        if lib.sheen_scene_defaults(config, C.byref(p)) != 0:
            raise ValueError("unknown workload %r" % config)
        for k, v in overrides.items():
            setattr(p, k, v)
        self.params = p
        self._h = lib.sheen_scene_generate(C.byref(p))
        if not self._h:
            raise RuntimeError("scene generation failed for %r" % config)
        self.scene = lib.sheen_scene_get(self._h)
        self.width, self.height, self.aa = p.width, p.height, p.aa
        w, h, a = C.c_int(), C.c_int(), C.c_int()
        if lib.sheen_scene_view(self._h, C.byref(w), C.byref(h), C.byref(a)):
            self.width, self.height, self.aa = w.value, h.value, a.value

    def materials(self):
        n = C.c_int()
        ptr = native.host().sheen_scene_materials(self._h, C.byref(n))
        return [ptr[i] for i in range(n.value)]

    def lights(self):
        n = C.c_int()
        ptr = native.host().sheen_scene_lights(self._h, C.byref(n))
        return [ptr[i] for i in range(n.value)]

    def close(self):
        if self._h:
            native.host().sheen_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def generate(config, **overrides):
    return GeneratedScene(config, **overrides)


BATCH_SEED = 0x5EED0400


def batch(n=256, seed=BATCH_SEED, **overrides):
    """The uneven batch workload: n config-4 scenes with seeds seed..seed+n-1
    (triangle counts log-uniform in [1e3, 2e5], resolutions from
    {128, 256, 512, 1024}, 1-8 spp)."""
    return [GeneratedScene(4, seed=seed + k, **overrides) for k in range(n)]


class SheenFile:
    """A parsed .sheen scene file: ``.scene`` (SheenScene pointer), ``.width``,
    ``.height``, ``.aa``, ``.output``.  Raises ValueError with the loader's
    message (and line number) when the file does not parse."""

    def __init__(self, path):
        lib = native.host()
        self._h = lib.sheen_file_load(str(path).encode())
        err = lib.sheen_file_error(self._h).decode()
        if err:
            lib.sheen_file_free(self._h)
            self._h = None
            raise ValueError(err)
        self.scene = lib.sheen_file_scene(self._h)
        w, h, a, out = C.c_int(), C.c_int(), C.c_int(), C.c_char_p()
        lib.sheen_file_settings(self._h, C.byref(w), C.byref(h), C.byref(a), C.byref(out))
        self.width, self.height, self.aa, self.output = w.value, h.value, a.value, out.value.decode()

    def close(self):
        if self._h:
            native.host().sheen_file_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def load_sheen(path):
    return SheenFile(path)
