"""Build recipe of the native libraries (in-tree, gfx950).

libsheen_host.so   host C++: scene generators, wide-BVH builder, PNG/sRGB.
libsheen_c_api.so  HIP kernels + the C ABI of include/sheen_c_api.h.
_sheen*.so         CPython module `sheen_amd._sheen` (reference
                   api/sheen_py_module.cpp counterpart) over libsheen_c_api.so.
addon/sheen.so     the same module named `sheen`, the file the Blender
                   add-on imports from its folder.
"""
import os
import subprocess
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_obj")
ADDON = os.path.join(HERE, "addon")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SHEEN_OFFLOAD_ARCH", "gfx950")

HOST_SRCS = ["wbvh_build.cpp", "host_scene.cpp", "png.cpp"]
HIP_SRCS = ["render.hip", "wbvh_gpu.hip"]
CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", "-fwrapv", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "--offload-arch=" + ARCH,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _newer(out, deps):
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    return any(os.path.getmtime(d) > t for d in deps)


def _headers():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    hs += [os.path.join(INC, f) for f in os.listdir(INC) if f.endswith(".h")]
    return hs


def build(force=True):
    os.makedirs(OBJ, exist_ok=True)
    hdrs = _headers()
    objs = []
    for src in HOST_SRCS + ["scenes.cpp", "host_api.cpp", "sheen_file.cpp"]:
        o = os.path.join(OBJ, src.replace(".cpp", ".o"))
        if force or _newer(o, [os.path.join(CSRC, src)] + hdrs):
            _run(["g++"] + CXXFLAGS + ["-I", INC, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    host_lib = os.path.join(HERE, "libsheen_host.so")
    _run(["g++", "-shared", "-o", host_lib] + objs + ["-lz"])
    hip_objs = []
    for src in HIP_SRCS:
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        if force or _newer(o, [os.path.join(CSRC, src)] + hdrs):
            _run([HIPCC] + HIPFLAGS + ["-I", INC, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", o])
        hip_objs.append(o)
    capi = os.path.join(HERE, "libsheen_c_api.so")
    host_objs = [os.path.join(OBJ, s.replace(".cpp", ".o")) for s in HOST_SRCS]
    _run([HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", capi] + hip_objs + host_objs + ["-lz"])
    # CPython module, twice from one source: `sheen_amd._sheen` for the
    # package, and `sheen` as the file sheen.so in addon/, the name and file
    # the Blender add-on imports from its own folder (reference
    # CMakeLists.txt:170-177, engine.py:29,61); its rpath finds
    # libsheen_c_api.so beside it (an add-on install) or one folder up (here).
    pyinc = sysconfig.get_paths()["include"]
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    mod = os.path.join(HERE, "_sheen" + ext)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-DSHEEN_PY_MODULE=_sheen", "-I", INC, "-I", pyinc,
          os.path.join(CSRC, "py_sheen.cpp"), "-o", mod, "-L", HERE, "-lsheen_c_api", "-Wl,-rpath,$ORIGIN"])
    os.makedirs(ADDON, exist_ok=True)
    addon_mod = os.path.join(ADDON, "sheen.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-DSHEEN_PY_MODULE=sheen", "-I", INC, "-I", pyinc,
          os.path.join(CSRC, "py_sheen.cpp"), "-o", addon_mod, "-L", HERE, "-lsheen_c_api",
          "-Wl,-rpath,$ORIGIN:$ORIGIN/.."])
    return [host_lib, capi, mod, addon_mod]


if __name__ == "__main__":
    build()
