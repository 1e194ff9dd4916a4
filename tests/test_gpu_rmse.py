"""RMSE of each workload's GPU render (its own spp) against the FP64 oracle
with 4096 samples per pixel, on the crops of tests/golden/rmse_refs.npz
(made by tests/golden/make_rmse_refs.py).  The GPU takes the same samples as
the oracle's FP32 estimate at that spp, so its RMSE must equal that
estimator's.  (An FP64 estimate at the same spp can differ by a few percent
where primitives are found iteratively: superellipsoid bisection and height
field marching take other steps in double.)"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REFS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rmse_refs.npz")
WORKLOADS = {"cfg1": (1, {}), "cfg2": (2, {}), "cfg3": (3, {}), "cfg4": (4, {"seed": 0x5EED0400}), "cfg5": (5, {})}


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_rmse_against_fp64_4096spp(name):
    from sheen_amd import scenes, sheen
    refs = np.load(REFS)
    cfg, kw = WORKLOADS[name]
    sc = scenes.generate(cfg, **kw)
    W, H, A = (int(v) for v in refs[name + "_view"])
    assert (sc.width, sc.height, sc.aa) == (W, H, A)
    x0, y0, cw, ch = (int(v) for v in refs[name + "_window"])
    with sheen.Context(sc.scene) as ctx:
        img, _ = ctx.render(W, H, A)
    crop = img[y0:y0 + ch, x0:x0 + cw, :3].astype(np.float64)
    ref, est = refs[name + "_ref"], refs[name + "_est"]
    rmse = float(np.sqrt(np.mean((crop - ref) ** 2)))
    rmse_est = float(np.sqrt(np.mean((est - ref) ** 2)))
    rmse64 = float(np.sqrt(np.mean((refs[name + "_est64"] - ref) ** 2)))
    print("%s: RMSE vs 4096-spp FP64 %.5f (oracle FP32 estimator %.5f, FP64 estimator %.5f, mean radiance %.4f)"
          % (name, rmse, rmse_est, rmse64, ref.mean()))
    assert abs(rmse - rmse_est) <= 0.01 * rmse_est + 1e-5
    rel = np.abs(crop - est) / np.maximum(np.abs(est), 1e-3)
    assert np.mean(rel.max(-1) > 1e-3) <= 0.02
