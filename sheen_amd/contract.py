"""ctypes layout of the C ABI records (include/sheen_types.h, sheen_c_api.h).

Byte-compatible with the reference's ctypes contract
(addon_sheen/sheen_render/c_api_contract.py:40-274) and with
``SheenScene`` of reference sheen_c_api.h:18-87 (352 bytes here: the C
struct's own size; the reference's ctypes mirror stops at 344).
"""
import ctypes as C


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Float2(C.Structure):
    _pack_ = 8
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


def _al16(cls):
    """Pad a structure to a multiple of 16 bytes (the C records are alignas(16))."""
    size = C.sizeof(cls)
    pad = (-size) % 16
    if pad:
        cls = type(cls.__name__, (C.Structure,), {"_fields_": cls._fields_ + [("_tail", C.c_byte * pad)]})
    return cls


class CameraData(C.Structure):
    _fields_ = [("position", Float3), ("right", Float3), ("up", Float3), ("forward", Float3),
                ("focal_length", C.c_float), ("sensor_width", C.c_float), ("shift_x", C.c_float),
                ("shift_y", C.c_float)]


Material = _al16(type("Material", (C.Structure,), {"_fields_": [
    ("diffuse_color", Float3), ("specular", C.c_float), ("shininess", C.c_float), ("ambient", C.c_float),
    ("texture_id", C.c_int), ("_pad_tex", C.c_int), ("anisotropy", C.c_float), ("hair_shift", C.c_float),
    ("hair_spec_r", C.c_float), ("hair_exp_r", C.c_float), ("hair_spec_t", C.c_float), ("hair_exp_t", C.c_float),
    ("hair_backlight", C.c_float), ("reflect", C.c_float), ("transmission", C.c_float), ("ior", C.c_float),
    ("emission_color", Float3), ("emission", C.c_float), ("mask_id", C.c_int), ("normal_id", C.c_int),
    ("normal_strength", C.c_float)]}))

Triangle = _al16(type("Triangle", (C.Structure,), {"_fields_": [
    ("v0", Float3), ("v1", Float3), ("v2", Float3), ("n0", Float3), ("n1", Float3), ("n2", Float3),
    ("uv0", Float2), ("uv1", Float2), ("uv2", Float2), ("material_id", C.c_int), ("_pad0", C.c_int),
    ("_pad1", C.c_int)]}))

Object = _al16(type("Object", (C.Structure,), {"_fields_": [
    ("type", C.c_int), ("material_id", C.c_int), ("param1", Float3), ("param2", C.c_float),
    ("_pad0", C.c_float), ("_pad1", C.c_float), ("_pad2", C.c_float)]}))

Light = _al16(type("Light", (C.Structure,), {"_fields_": [
    ("direction", Float3), ("position", Float3), ("u_axis", Float3), ("v_axis", Float3), ("color", Float3),
    ("intensity", C.c_float), ("radius", C.c_float), ("angle", C.c_float), ("blend", C.c_float),
    ("beam", C.c_float), ("type", C.c_int), ("flags", C.c_int), ("texture_id", C.c_int), ("_pad1", C.c_int)]}))

Medium = _al16(type("Medium", (C.Structure,), {"_fields_": [
    ("center", Float3), ("half", Float3), ("sigma_s", Float3), ("sigma_t", C.c_float), ("samples", C.c_int),
    ("_pad0", C.c_int)]}))

Plane = _al16(type("Plane", (C.Structure,), {"_fields_": [
    ("normal", Float3), ("d", C.c_float), ("material_id", C.c_int), ("_pad0", C.c_int), ("_pad1", C.c_int),
    ("_pad2", C.c_int)]}))


class SweepPoint(C.Structure):
    _fields_ = [("position", Float3), ("radius", C.c_float)]


class SweepStrand(C.Structure):
    _fields_ = [("point_offset", C.c_int), ("point_count", C.c_int), ("material_id", C.c_int), ("_pad0", C.c_int)]


class HeightField(C.Structure):
    _fields_ = [("center", Float3), ("size_x", C.c_float), ("size_z", C.c_float), ("height", C.c_float),
                ("freq", C.c_float), ("octaves", C.c_int), ("seed", C.c_int), ("material_id", C.c_int),
                ("image", C.POINTER(C.c_ubyte)), ("img_w", C.c_int), ("img_h", C.c_int), ("_pad0", C.c_int),
                ("_tail", C.c_byte * 4)]


class LathePoint(C.Structure):
    _fields_ = [("y", C.c_float), ("radius", C.c_float), ("_pad0", C.c_int), ("_pad1", C.c_int)]


class Lathe(C.Structure):
    _fields_ = [("center", Float3), ("point_offset", C.c_int), ("point_count", C.c_int), ("material_id", C.c_int),
                ("_pad0", C.c_int), ("_tail", C.c_byte * 4)]


class PrismPoint(C.Structure):
    _fields_ = [("x", C.c_float), ("z", C.c_float), ("_pad0", C.c_int), ("_pad1", C.c_int)]


Prism = _al16(type("Prism", (C.Structure,), {"_fields_": [
    ("center", Float3), ("y_min", C.c_float), ("y_max", C.c_float), ("sweep", C.c_float),
    ("point_offset", C.c_int), ("point_count", C.c_int), ("material_id", C.c_int), ("_pad0", C.c_int)]}))


class CSGNode(C.Structure):
    _fields_ = [("op", C.c_int), ("leaf_kind", C.c_int), ("leaf_index", C.c_int), ("left", C.c_int),
                ("right", C.c_int), ("_tail", C.c_byte * 12)]


Superellipsoid = _al16(type("Superellipsoid", (C.Structure,), {"_fields_": [
    ("center", Float3), ("a", C.c_float), ("b", C.c_float), ("c", C.c_float), ("n", C.c_float),
    ("material_id", C.c_int), ("_pad0", C.c_int), ("_pad1", C.c_int)]}))


class Photon(C.Structure):
    _fields_ = [("position", Float3), ("direction", Float3), ("normal", Float3), ("flux", Float3)]


class SheenScene(C.Structure):
    _fields_ = [
        ("camera", CameraData),
        ("objects", C.POINTER(Object)), ("num_objects", C.c_int),
        ("sweep_points", C.POINTER(SweepPoint)), ("num_sweep_points", C.c_int),
        ("sweep_strands", C.POINTER(SweepStrand)), ("num_sweep_strands", C.c_int),
        ("triangles", C.POINTER(Triangle)), ("num_triangles", C.c_int),
        ("heightfields", C.POINTER(HeightField)), ("num_heightfields", C.c_int),
        ("superellipsoids", C.POINTER(Superellipsoid)), ("num_superellipsoids", C.c_int),
        ("lathe_points", C.POINTER(LathePoint)), ("num_lathe_points", C.c_int),
        ("lathes", C.POINTER(Lathe)), ("num_lathes", C.c_int),
        ("prism_points", C.POINTER(PrismPoint)), ("num_prism_points", C.c_int),
        ("prisms", C.POINTER(Prism)), ("num_prisms", C.c_int),
        ("csg_nodes", C.POINTER(CSGNode)), ("csg_root", C.c_int), ("num_csg_nodes", C.c_int),
        ("planes", C.POINTER(Plane)), ("num_planes", C.c_int),
        ("materials", C.POINTER(Material)), ("num_materials", C.c_int),
        ("texture_pixels", C.POINTER(C.POINTER(C.c_ubyte))), ("texture_widths", C.POINTER(C.c_int)),
        ("texture_heights", C.POINTER(C.c_int)), ("num_textures", C.c_int),
        ("lights", C.POINTER(Light)), ("num_lights", C.c_int),
        ("photons", C.c_int), ("photon_cell", C.c_float), ("caustics", C.c_float),
        ("media", C.POINTER(Medium)), ("num_media", C.c_int),
        ("_tail", C.c_byte * 12),
    ]


SIZES = {
    "CameraData": 64, "Material": 112, "Triangle": 112, "Object": 48, "Light": 96, "Medium": 48, "Plane": 32,
    "SweepPoint": 16, "SweepStrand": 16, "HeightField": 64, "LathePoint": 16, "Lathe": 32, "PrismPoint": 16,
    "Prism": 48, "CSGNode": 32, "Superellipsoid": 48, "Photon": 48, "SheenScene": 352,
}
