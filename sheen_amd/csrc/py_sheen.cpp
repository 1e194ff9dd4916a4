// py_sheen.cpp — CPython module `sheen_amd._sheen`, the counterpart of the
// reference's `sheen` module (api/sheen_py_module.cpp:33-105): init, cleanup
// and render_scene over the C ABI, the scene passed as any buffer of at least
// sizeof(SheenScene) bytes (a ctypes SheenScene), zero copy.
#define PY_SSIZE_T_CLEAN
#include <Python.h>

#include "sheen_c_api.h"

static PyObject* py_init(PyObject*, PyObject* args) {
    int dev = 0, arch = 0;
    if (!PyArg_ParseTuple(args, "ii", &dev, &arch)) return nullptr;
    if (sheen_init(dev, arch) != 0) {
        PyErr_SetString(PyExc_RuntimeError, sheen_get_last_error());
        return nullptr;
    }
    Py_RETURN_NONE;
}

static PyObject* py_cleanup(PyObject*, PyObject*) {
    sheen_cleanup();
    Py_RETURN_NONE;
}

static PyObject* py_render_scene(PyObject*, PyObject* args) {
    PyObject* scene_obj = nullptr;
    int width = 0, height = 0, aa = 4;
    const char* out_path = nullptr;
    if (!PyArg_ParseTuple(args, "Oiiis", &scene_obj, &width, &height, &aa, &out_path)) return nullptr;
    Py_buffer buf;
    if (PyObject_GetBuffer(scene_obj, &buf, PyBUF_SIMPLE) != 0) {
        PyErr_Clear();
        PyErr_SetString(PyExc_TypeError, "scene must be a buffer (ctypes SheenScene structure)");
        return nullptr;
    }
    if (buf.len < (Py_ssize_t)sizeof(SheenScene)) {
        PyBuffer_Release(&buf);
        PyErr_SetString(PyExc_TypeError, "scene buffer too small for SheenScene");
        return nullptr;
    }
    const SheenScene* scene = static_cast<const SheenScene*>(buf.buf);
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = sheen_render_scene(scene, width, height, aa, out_path);
    Py_END_ALLOW_THREADS
    PyBuffer_Release(&buf);
    if (rc != 0) {
        PyErr_SetString(PyExc_RuntimeError, sheen_get_last_error());
        return nullptr;
    }
    Py_RETURN_NONE;
}

static PyMethodDef Methods[] = {
    {"init", py_init, METH_VARARGS, "init(device_id, arch): select the HIP device."},
    {"cleanup", py_cleanup, METH_NOARGS, "cleanup(): synchronise the device."},
    {"render_scene", py_render_scene, METH_VARARGS,
     "render_scene(scene, width, height, aa, output_path): render a SheenScene buffer to PNG."},
    {nullptr, nullptr, 0, nullptr},
};

static struct PyModuleDef Module = {PyModuleDef_HEAD_INIT, "_sheen",
                                    "Sheen render path on MI355X (HIP) behind the reference's module API.", -1, Methods};

PyMODINIT_FUNC PyInit__sheen(void) { return PyModule_Create(&Module); }
