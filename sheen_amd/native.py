"""ctypes bindings of libsheen_c_api.so (HIP path) and libsheen_host.so.

Loading fails loudly when a library has not been built: there is no CPU
fallback behind these entry points.
"""
import ctypes as C
import os

import numpy as np

from . import contract

HERE = os.path.dirname(os.path.abspath(__file__))
CAPI_PATH = os.environ.get("SHEEN_CAPI_PATH") or os.path.join(HERE, "libsheen_c_api.so")
HOST_PATH = os.path.join(HERE, "libsheen_host.so")


class SheenRayStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("primary_rays", "shadow_rays", "reflection_rays", "refraction_rays",
                                          "photon_rays", "total_rays", "samples", "node_fetches", "prim_tests",
                                          "shadow_node_fetches", "shadow_prim_tests", "primary_node_fetches",
                                          "primary_prim_tests", "primary_stack_spills", "shadow_stack_spills")] + \
               [("frame_ms", C.c_float), ("trace_ms", C.c_float), ("primary_ms", C.c_float), ("shadow_ms", C.c_float), ("shade_ms", C.c_float),
                ("trace_launches", C.c_int), ("shadow_launches", C.c_int), ("shade_primary_ms", C.c_float),
                ("shadow_visit_lanes", C.c_uint64), ("cont_node_fetches", C.c_uint64), ("cont_prim_tests", C.c_uint64),
                ("shadow_cont_ms", C.c_float), ("cont_launches", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SheenSceneParams(C.Structure):
    _fields_ = [("config", C.c_int), ("seed", C.c_uint64), ("width", C.c_int), ("height", C.c_int), ("aa", C.c_int),
                ("detail", C.c_int), ("scale", C.c_float), ("reserved", C.c_int * 7)]


_capi = None
_host = None


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def capi():
    """libsheen_c_api.so with typed signatures (raises OSError if not built)."""
    global _capi
    if _capi is not None:
        return _capi
    if not os.path.exists(CAPI_PATH):
        raise OSError("libsheen_c_api.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(CAPI_PATH)
    P = C.c_void_p
    S = C.POINTER(contract.SheenScene)
    _sig(lib, "sheen_init", C.c_int, [C.c_int, C.c_int])
    _sig(lib, "sheen_cleanup", None, [])
    _sig(lib, "sheen_render_scene", C.c_int, [S, C.c_int, C.c_int, C.c_int, C.c_char_p])
    _sig(lib, "sheen_get_last_error", C.c_char_p, [])
    _sig(lib, "sheen_context_create", P, [S])
    _sig(lib, "sheen_context_create_ex", P, [S, C.c_int])
    _sig(lib, "sheen_context_bvh", C.c_int, [P, C.c_int, P, C.c_int, P, C.c_int, C.POINTER(C.c_int)])
    _sig(lib, "sheen_context_build_ms", C.c_double, [P])
    _sig(lib, "sheen_context_destroy", None, [P])
    _sig(lib, "sheen_context_render", C.c_int, [P, C.c_int, C.c_int, C.c_int, P, C.POINTER(SheenRayStats)])
    _sig(lib, "sheen_context_frame_device", P, [P])
    _sig(lib, "sheen_context_render_noalloc", C.c_int, [P, C.c_int, C.c_int, C.c_int, P, C.POINTER(SheenRayStats)])
    _sig(lib, "sheen_context_primary_hits", C.c_int, [P, C.c_int, C.c_int, C.c_int, P, P, P])
    _sig(lib, "sheen_context_set_profiling", C.c_int, [P, C.c_int])
    _sig(lib, "sheen_context_set_counting", C.c_int, [P, C.c_int])
    _sig(lib, "sheen_context_set_chunk", C.c_int, [P, C.c_int])
    _sig(lib, "sheen_context_set_param", C.c_int, [P, C.c_char_p, C.c_int])
    _sig(lib, "sheen_batch_create", P, [P, P, P, P, C.c_int])
    _sig(lib, "sheen_batch_destroy", None, [P])
    _sig(lib, "sheen_batch_set_param", C.c_int, [P, C.c_char_p, C.c_int])
    _sig(lib, "sheen_batch_render", C.c_int, [P, C.POINTER(SheenRayStats)])
    _sig(lib, "sheen_batch_render_serial", C.c_int, [P, C.POINTER(SheenRayStats)])
    _sig(lib, "sheen_batch_image", C.c_int, [P, C.c_int, C.c_int, P])
    _sig(lib, "sheen_batch_info", C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_int)])
    _sig(lib, "sheen_adjoint_sizes", C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_longlong)])
    _sig(lib, "sheen_context_stream_fallbacks", C.c_int, [P])
    _sig(lib, "sheen_adjoint_get_params", C.c_int, [P, P, P, P, P, P])
    _sig(lib, "sheen_adjoint_forward", C.c_int, [P, C.c_int, C.c_int, C.c_int, P, P, P, P, P, P])
    _sig(lib, "sheen_adjoint_backward", C.c_int, [P, C.c_int, C.c_int, C.c_int, P, P, P, P, P, P, P, P, P, P])
    _sig(lib, "sheen_context_photon_stats", C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int)])
    _sig(lib, "sheen_context_photon_map", C.c_int, [P, P, P, P])
    _sig(lib, "sheen_context_photon_probe", C.c_int, [P, C.c_int, C.c_float, C.c_int, C.c_int, P])
    _sig(lib, "sheen_render_float", C.c_int, [S, C.c_int, C.c_int, C.c_int, P, C.POINTER(SheenRayStats)])
    _sig(lib, "sheen_encode_srgb", None, [P, C.c_int, C.c_int, P])
    _sig(lib, "sheen_write_png", C.c_int, [C.c_char_p, C.c_int, C.c_int, P])
    _capi = lib
    return lib


def host():
    """libsheen_host.so (scene generators; CPU only)."""
    global _host
    if _host is not None:
        return _host
    if not os.path.exists(HOST_PATH):
        raise OSError("libsheen_host.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(HOST_PATH)
    P = C.c_void_p
    _sig(lib, "sheen_scene_defaults", C.c_int, [C.c_int, C.POINTER(SheenSceneParams)])
    _sig(lib, "sheen_scene_generate", P, [C.POINTER(SheenSceneParams)])
    _sig(lib, "sheen_scene_get", C.POINTER(contract.SheenScene), [P])
    _sig(lib, "sheen_scene_view", C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)])
    _sig(lib, "sheen_scene_materials", C.POINTER(contract.Material), [P, C.POINTER(C.c_int)])
    _sig(lib, "sheen_scene_lights", C.POINTER(contract.Light), [P, C.POINTER(C.c_int)])
    _sig(lib, "sheen_scene_texture", C.POINTER(C.c_ubyte), [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)])
    _sig(lib, "sheen_scene_free", None, [P])
    _sig(lib, "sheen_host_bvh_check", C.c_int, [C.POINTER(contract.SheenScene), C.c_int, C.POINTER(C.c_int)])
    _sig(lib, "sheen_host_bvh_nodes", C.c_int, [C.POINTER(contract.SheenScene), C.c_int, P, P, P, P, C.POINTER(C.c_int)])
    _sig(lib, "sheen_host_wbvh_check", C.c_int, [P, C.c_int, P, C.c_int, P, C.c_int, P, C.POINTER(C.c_int)])
    _sig(lib, "sheen_host_csg_boxes", C.c_int, [C.POINTER(contract.SheenScene), C.c_void_p])
    _sig(lib, "sheen_host_csg_spine", C.c_int, [C.POINTER(contract.SheenScene), C.POINTER(C.c_int), C.c_void_p, C.c_void_p])
    _sig(lib, "sheen_file_load", P, [C.c_char_p])
    _sig(lib, "sheen_file_error", C.c_char_p, [P])
    _sig(lib, "sheen_file_scene", C.POINTER(contract.SheenScene), [P])
    _sig(lib, "sheen_file_settings", C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                C.POINTER(C.c_char_p)])
    _sig(lib, "sheen_file_free", None, [P])
    _host = lib
    return lib


def last_error():
    return capi().sheen_get_last_error().decode()


def check(rc):
    if rc != 0:
        raise RuntimeError(last_error())
    return rc


def as_ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None
