// py_sheen.cpp — the CPython module that the Blender add-on imports as
// `sheen` (reference contract: api/sheen_py_module.cpp, CMakeLists.txt:170-177,
// add-on engine.py:44-86).  Built twice from this file: as `sheen` (a file
// named sheen.so, what the add-on loads from its own folder) and as
// `sheen_amd._sheen` (the package's copy).  SHEEN_PY_MODULE names the module.
//
// The contract, as the add-on uses it:
//   init(device_id: int, arch: int) -> None        RuntimeError on failure
//   cleanup() -> None
//   render_scene(scene, width, height, aa, path) -> None
//       scene: any object exporting a buffer that holds a SheenScene
//       (normally a ctypes.Structure); TypeError when it is not a buffer or
//       is too short; RuntimeError(sheen_get_last_error()) when rendering
//       fails.
//
// Buffer length: the add-on's ctypes SheenScene is 344 bytes (its last
// field, num_media, ends at byte 340), the C struct is 352, and the
// reference module accepts 336 and up.  The bytes the caller has are copied
// into a zeroed SheenScene and the library reads only that copy, so a
// shorter buffer never lets it read past the caller's object: at 336 bytes
// num_media reads as 0 (no participating media).
#define PY_SSIZE_T_CLEAN
#include <Python.h>

#include <cstring>

#include "sheen_c_api.h"

#ifndef SHEEN_PY_MODULE
#define SHEEN_PY_MODULE sheen
#endif
#define SHEEN_STR2(x) #x
#define SHEEN_STR(x) SHEEN_STR2(x)
#define SHEEN_CAT2(a, b) a##b
#define SHEEN_CAT(a, b) SHEEN_CAT2(a, b)

namespace {

// Smallest scene buffer taken: everything up to the `media` pointer.
constexpr Py_ssize_t kMinSceneBytes = (Py_ssize_t)offsetof(SheenScene, num_media);
static_assert(kMinSceneBytes == 336, "scene prefix the reference module requires");

PyObject* library_error() {
    PyErr_SetString(PyExc_RuntimeError, sheen_get_last_error());
    return nullptr;
}

PyObject* mod_init(PyObject*, PyObject* args) {
    int device = 0, arch = 0;
    if (!PyArg_ParseTuple(args, "ii:init", &device, &arch)) return nullptr;
    if (sheen_init(device, arch) != 0) return library_error();
    Py_RETURN_NONE;
}

PyObject* mod_cleanup(PyObject*, PyObject*) {
    sheen_cleanup();
    Py_RETURN_NONE;
}

PyObject* mod_render_scene(PyObject*, PyObject* args) {
    PyObject* src = nullptr;
    int w = 0, h = 0, spp = 0;
    const char* path = nullptr;
    if (!PyArg_ParseTuple(args, "Oiiis:render_scene", &src, &w, &h, &spp, &path)) return nullptr;

    SheenScene local;
    std::memset(&local, 0, sizeof(local));
    {
        Py_buffer view;
        if (PyObject_GetBuffer(src, &view, PyBUF_SIMPLE) != 0) {
            PyErr_Clear();
            PyErr_SetString(PyExc_TypeError, "scene must be a buffer (ctypes SheenScene structure)");
            return nullptr;
        }
        const Py_ssize_t have = view.len;
        if (have < kMinSceneBytes) {
            PyBuffer_Release(&view);
            PyErr_Format(PyExc_TypeError, "scene buffer too small for SheenScene (%zd bytes, need at least %zd)", have,
                         kMinSceneBytes);
            return nullptr;
        }
        const size_t take = have < (Py_ssize_t)sizeof(local) ? (size_t)have : sizeof(local);
        std::memcpy(&local, view.buf, take);
        PyBuffer_Release(&view);
    }
    // The scene's arrays belong to the caller's Python objects, which keep
    // them alive for the duration of this call; the GIL is released while
    // the GPU works.
    int status = 0;
    Py_BEGIN_ALLOW_THREADS
    status = sheen_render_scene(&local, w, h, spp, path);
    Py_END_ALLOW_THREADS
    if (status != 0) return library_error();
    Py_RETURN_NONE;
}

PyMethodDef kMethods[] = {
    {"init", mod_init, METH_VARARGS, "init(device_id, arch): select the GPU (arch is ignored)."},
    {"cleanup", mod_cleanup, METH_NOARGS, "cleanup(): wait for the GPU to finish."},
    {"render_scene", mod_render_scene, METH_VARARGS,
     "render_scene(scene, width, height, aa, output_path): render a SheenScene buffer to an sRGB PNG."},
    {nullptr, nullptr, 0, nullptr},
};

PyModuleDef kModule = {
    PyModuleDef_HEAD_INIT, SHEEN_STR(SHEEN_PY_MODULE),
    "Sheen renderer on MI355X: the native module of the Blender add-on, over libsheen_c_api.so.", -1, kMethods,
    nullptr, nullptr, nullptr, nullptr,
};

}  // namespace

PyMODINIT_FUNC SHEEN_CAT(PyInit_, SHEEN_PY_MODULE)(void) { return PyModule_Create(&kModule); }
