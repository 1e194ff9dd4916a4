"""MI355X render path of Sheen: HIP kernels behind the reference's C ABI.

Modules:
  sheen      -- mirror of the reference's ``sheen`` Python module
                (init / cleanup / render_scene) plus float-output extensions.
  contract   -- ctypes layout of the scene records (SheenScene and friends).
  scenes     -- seeded synthetic workloads.
  native     -- ctypes access to libsheen_c_api.so / libsheen_host.so.
"""
__all__ = ["sheen", "contract", "scenes", "native"]
