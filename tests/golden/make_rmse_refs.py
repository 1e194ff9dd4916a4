"""Writes tests/golden/rmse_refs.npz: for each workload, a centred crop's
radiance from the FP64 oracle with 4096 samples per pixel (8x8 sub-pixels x
64 samples over the pixel footprint), and the oracle's FP32 (parity path)
and FP64 estimates at the workload's own spp on the same crop.

    python tests/golden/make_rmse_refs.py      (about 3 minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import Oracle  # noqa: E402
from sheen_amd import scenes  # noqa: E402

# name: (config, generator overrides, crop size)
WORKLOADS = {
    "cfg1": (1, {}, 32),
    "cfg2": (2, {}, 32),
    "cfg3": (3, {}, 16),
    "cfg4": (4, {"seed": scenes.BATCH_SEED}, 32),
    "cfg5": (5, {}, 32),
}


def main():
    out = {}
    threads = os.cpu_count() or 1
    for name, (cfg, kw, c) in WORKLOADS.items():
        t = time.time()
        sc = scenes.generate(cfg, **kw)
        W, H, A = sc.width, sc.height, sc.aa
        win = (W // 2 - c // 2, H // 2 - c // 2, c, c)
        o = Oracle(sc.scene, threads=threads)
        ref = o.render_reference(W, H, win, sub=8, aa=64)
        est = o.render_reference(W, H, win, sub=1, aa=A, fp64=False)
        est64 = o.render_reference(W, H, win, sub=1, aa=A)
        o.close()
        out[name + "_ref"] = ref
        out[name + "_est"] = est
        out[name + "_est64"] = est64
        out[name + "_window"] = np.array(win, np.int32)
        out[name + "_view"] = np.array([W, H, A], np.int32)
        rmse = float(np.sqrt(np.mean((est - ref) ** 2)))
        rmse64 = float(np.sqrt(np.mean((est64 - ref) ** 2)))
        print("%s %dx%d x%d spp, crop %s: RMSE %.5f (FP64 estimate %.5f) (%.0f s)"
              % (name, W, H, A, win, rmse, rmse64, time.time() - t), flush=True)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rmse_refs.npz"), **out)


if __name__ == "__main__":
    main()
