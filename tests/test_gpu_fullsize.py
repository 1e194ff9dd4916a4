"""The workloads at their full sizes on the GPU against the oracle:
* the 1M-triangle mesh at 1920x1080 x 4 spp (1,048,820 triangles, 2 point
  lights + 1 area light, 30 % mirror / glass);
* the 10,000-sphere textured shading workload at 512x512 x 8 spp;
* the analytic workload (10,000 primitives, CSG, caustics from 1.65 M
  emitted photons) at 1280x720 x 16 spp, on a band of 16 rows through the
  middle of the frame (the oracle traces the full frame in minutes);
* the 256-scene uneven batch: every scene's batched image equals its own
  render bit for bit, and the batched images of scenes at each of the four
  resolutions equal the oracle.
Per case the fraction of pixels beyond 1e-4 relative and the largest
relative error are printed and bounded (relative error
|gpu - oracle| / max(|oracle|, 1e-3), worst channel); primary hits
(category, index) must match bit for bit.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TH = min(16, os.cpu_count() or 1)
# (name, config, overrides, rows (None: all), max fraction beyond 1e-4, max relative error)
CASES = [
    ("mesh-1M-1080p", 2, {}, None, 0.0, 1e-4),
    ("shading-10k-512", 5, {}, None, 0.0, 1e-4),
    ("analytic-10k-720p", 3, {}, 16, 0.0, 1e-4),
]


def _rel(a, b):
    return (np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(b.astype(np.float64)), 1e-3)).max(-1)


def _check(name, img, ref, max_frac, max_rel):
    rel = _rel(img[..., :3], ref[..., :3])
    frac = float(np.mean(rel > 1e-4))
    print("%s: %d pixels, beyond 1e-4: %.6f, max rel %.3g" % (name, rel.size, frac, rel.max()), flush=True)
    assert frac <= max_frac, frac
    assert rel.max() <= max_rel, rel.max()


@pytest.mark.parametrize("name,cfg,kw,band,max_frac,max_rel", CASES, ids=[c[0] for c in CASES])
def test_full_size_workload_matches_oracle(name, cfg, kw, band, max_frac, max_rel):
    from oracle import Oracle
    from sheen_amd import scenes, sheen
    sc = scenes.generate(cfg, **kw)
    W, H, A = sc.width, sc.height, sc.aa
    print("case", name, W, H, A, flush=True)
    with sheen.Context(sc.scene) as ctx:
        img, st = ctx.render(W, H, A)
        kind, idx, _ = ctx.primary_hits(W, H, A)
    rows = (H // 2 - band // 2, H // 2 + band // 2) if band else (0, H)
    ref = Oracle(sc.scene, threads=TH).render(W, H, A, rows=rows, hits=True)
    y0, y1 = rows
    assert np.array_equal(kind[y0:y1], ref["kind"][y0:y1])
    assert np.array_equal(idx[y0:y1], ref["index"][y0:y1])
    _check(name, img[y0:y1], ref["image"][y0:y1], max_frac, max_rel)
    if band is None:
        for k in ("primary_rays", "shadow_rays", "reflection_rays", "refraction_rays"):
            assert st[k] == ref["stats"][k], k


def test_batch_256_every_scene_equals_its_own_render_and_the_oracle():
    from oracle import Oracle
    from sheen_amd import scenes
    from sheen_amd.batch import Batch
    sc = scenes.batch(256)
    b = Batch(sc)
    print("batch of %d scenes, %d samples, %d chunks" % (b.n, b.samples, b.chunks), flush=True)
    sb = b.render()
    ss = b.render_serial()
    for key in ("samples", "primary_rays", "shadow_rays", "reflection_rays", "refraction_rays", "total_rays"):
        assert sb[key] == ss[key], key
    by_res = {}
    for k in range(b.n):
        img = b.image(k)
        assert np.array_equal(img, b.image(k, serial=True)), k
        by_res.setdefault(b.sizes[k][0], k)
    assert sorted(by_res) == [128, 256, 512, 1024]
    for res, k in sorted(by_res.items()):
        W, H, A = b.sizes[k]
        ref = Oracle(sc[k].scene, threads=TH).render(W, H, A)["image"]
        _check("batch scene %d (%dx%d x %d spp)" % (k, W, H, A), b.image(k), ref, 0.0, 1e-4)
    b.close()
