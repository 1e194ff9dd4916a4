"""The `sheen` module the reference's Blender add-on imports (reference
engine.py:29,44-86: `import sheen` from the add-on's folder, which holds
sheen.so and libsheen_c_api.so; api/sheen_py_module.cpp is the module it
replaces).  Runs in a subprocess with only the add-on folder on sys.path, the
way the add-on inserts it, and checks the contract on the CPU: the module
name, the three functions, a 344-byte ctypes SheenScene (the add-on's
layout, c_api_contract.py:233-274) reaching the library's own validation,
and the error types.  No device call is made.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "sheen_amd", "addon")

SCRIPT = r'''
import ctypes as C, sys
sys.path.insert(0, sys.argv[1])
sys.path.insert(1, sys.argv[2])
import sheen
from sheen_amd import contract as K
assert sheen.__name__ == "sheen", sheen.__name__
assert sheen.__file__.endswith("addon/sheen.so"), sheen.__file__
assert {"init", "cleanup", "render_scene"} <= set(dir(sheen))

# the add-on's ctypes SheenScene: the C fields without the alignas(16) tail
class AddonScene(C.Structure):
    _fields_ = [f for f in K.SheenScene._fields_ if f[0] != "_tail"]
assert C.sizeof(AddonScene) == 344, C.sizeof(AddonScene)
s = AddonScene()
s.csg_root = -1
# validation order of the reference (api/sheen_c_api.cu:189-193): the
# library is reached and answers with its own messages
for args, msg in (((s, 0, 8, 1, "/tmp/x.png"), "bad render params"),
                  ((s, 8, 8, 65, "/tmp/x.png"), "bad render params"),
                  ((s, 8, 8, 1, "/tmp/x.png"), "no materials")):
    try:
        sheen.render_scene(*args)
    except RuntimeError as e:
        assert str(e) == msg, (str(e), msg)
    else:
        raise AssertionError("no error for %r" % (args[1:],))
# a 336-byte prefix is the smallest buffer taken (the reference module's
# minimum); 335 bytes and non-buffers are TypeErrors
buf = bytearray(bytes(s)[:336])
try:
    sheen.render_scene(buf, 8, 8, 1, "/tmp/x.png")
except RuntimeError as e:
    assert str(e) == "no materials", str(e)
for bad in (bytearray(335), 12345, "scene"):
    try:
        sheen.render_scene(bad, 8, 8, 1, "/tmp/x.png")
    except TypeError:
        pass
    else:
        raise AssertionError("no TypeError for %r" % type(bad))
try:
    sheen.init("x", 61)
except TypeError:
    pass
else:
    raise AssertionError("init accepted a string")
print("ok")
'''


def test_addon_imports_sheen_and_passes_its_344_byte_scene():
    assert os.path.exists(os.path.join(ADDON, "sheen.so")), "build the library first"
    env = dict(os.environ)
    env.pop("PYTHONPATH", None)
    r = subprocess.run([sys.executable, "-c", SCRIPT, ADDON, ROOT], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
