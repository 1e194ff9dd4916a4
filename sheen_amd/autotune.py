"""Launch-shape autotuner for the wavefront render path.

The knobs the reference's build fixes by hand (workgroup size, BVH width,
queue depth, rays per thread; `render_kernel.cu:1240-1262` launches one
16x16 block per tile) map here to:

* rays per thread -> blocks per CU of each grid-stride kernel
  (`primary_grid`, `sec_grid`, `shadow_grid`, `shade_grid`,
  `shade_sec_grid`; fewer blocks = more rays per thread);
* queue depth     -> `chunk`, the samples of one wavefront chunk (every
  device queue is sized from it);
* workgroup size -> `shade_block` for the shading kernels (up to the bound
  they are compiled for); the traversal kernels' 256 threads are compile-time
  (the LDS stack columns are laid out per 256-lane block), as is the BVH
  width, 8 children (one 128-byte node);
* queue order     -> `shadow_order`, `sec_order`, `light_runs` (which rays
  share a wave of the trace kernels) and `sec_sort` (secondary rays sorted
  by kind and direction before each trace round).

`tune()` runs coordinate descent: each knob is swept with the others held,
scored by the GPU time of the kernel it shapes (hipEvents on the render
stream), and the best value kept.  Results do not depend on the knobs (every
queue entry carries its sample slot), so tuning changes time only.
"""
import json

SPACE = {
    "chunk": (1 << 21, 1 << 22, 1 << 23, 0),
    "shadow_grid": (16, 32, 64, 96, 128, 0),
    "primary_grid": (4, 8, 16, 32, 64, 0),
    "sec_grid": (4, 6, 8, 12, 16, 24),
    "shade_grid": (1, 2, 3, 4, 6, 8, 16),
    "shade_sec_grid": (1, 2, 4, 8, 16),
    "shade_block": (1024, 512, 256),
    "overlap": (0, 1),
    "light_runs": (1, 0),
    "shadow_order": (1, 0),
    "sec_order": (1, 0),
    "sec_sort": (0, 1, 2),
}

# the library's built-in launch shapes (tuned on the 1M-triangle 1080p frame)
DEFAULTS = {"primary_grid": 64, "sec_grid": 6, "shadow_grid": 64, "shade_grid": 1, "shade_sec_grid": 1, "chunk": 0,
            "shade_block": 0, "overlap": 0, "light_runs": 1, "shadow_order": 1, "sec_order": 1,
            "sec_sort": 2}


def _score(name, st):
    """GPU milliseconds of the kernel a knob shapes."""
    if name == "primary_grid":
        return st["primary_ms"]
    if name == "sec_grid":
        return st["trace_ms"] - st["primary_ms"]
    if name == "shadow_grid":  # also the grid of the glass continuation kernels
        return st["shadow_ms"] + st["shadow_cont_ms"]
    if name == "shade_grid":
        return st["shade_primary_ms"]
    if name == "shade_sec_grid":
        return st["shade_ms"] - st["shade_primary_ms"]
    # "overlap" runs kernels concurrently, and the workgroup size and the
    # queue orders change which rays share a wave of the trace kernels: only
    # the frame time tells
    return st["frame_ms"]


def apply(ctx, params):
    for k, v in params.items():
        ctx.set_param(k, v)


def tune(ctx, width, height, aa, space=None, reps=3, passes=1, log=None):
    """Pick launch parameters for `ctx` rendering width x height x aa.

    Returns {"params": {...}, "frame_ms": best frame time, "trials": [...]}.
    The context is left configured with the chosen parameters.
    """
    space = dict(SPACE if space is None else space)
    params = dict(DEFAULTS)
    apply(ctx, params)
    ctx.set_profiling(True)
    trials = []

    def measure():
        ctx.render(width, height, aa, download=False)  # first launch at a new shape
        best = None
        for _ in range(reps):
            _, st = ctx.render(width, height, aa, download=False)
            if best is None or st["frame_ms"] < best["frame_ms"]:
                best = dict(st)
        return best

    try:
        for _ in range(passes):
            for name, values in space.items():
                scores = {}
                for v in values:
                    ctx.set_param(name, v)
                    st = measure()
                    scores[v] = _score(name, st)
                    trials.append({"param": name, "value": v, "ms": round(scores[v], 4),
                                   "frame_ms": round(st["frame_ms"], 4)})
                    if log:
                        log(trials[-1])
                params[name] = min(scores, key=scores.get)
                ctx.set_param(name, params[name])
        final = measure()
    finally:
        ctx.set_profiling(False)
    return {"params": params, "frame_ms": final["frame_ms"], "trials": trials}


def save(result, path):
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load(path):
    with open(path) as f:
        return json.load(f)["params"]


# Batches: the knobs of one frame of every scene, scored by the batch's GPU
# frame time (the kernels are replayed from one captured graph, so per-kernel
# timers do not apply).
BATCH_SPACE = {
    "chunk": (1 << 24, 1 << 25, 1 << 26),
    "shadow_grid": (16, 32, 64, 128),
    "primary_grid": (8, 16, 64, 0),
    "sec_grid": (4, 8, 16, 32),
    "shade_grid": (1, 2, 4, 8),
    "shade_sec_grid": (1, 2, 4, 8),
    "shade_block": (1024, 512, 256),
    "shadow_order": (1, 0),
    "sec_order": (1, 0),
    "sec_sort": (0, 1, 2),
}


def tune_batch(batch, space=None, reps=2, log=None):
    """Coordinate descent over the batch knobs; leaves the batch configured
    with the best values.  Returns {"params", "frame_ms", "trials"}."""
    space = dict(BATCH_SPACE if space is None else space)
    params = {}
    trials = []

    def measure():
        batch.render()  # first frame at a new shape recaptures the graph
        return min(batch.render()["frame_ms"] for _ in range(reps))

    for name, values in space.items():
        scores = {}
        for v in values:
            batch.set_param(name, v)
            scores[v] = measure()
            trials.append({"param": name, "value": v, "frame_ms": round(scores[v], 3)})
            if log:
                log(trials[-1])
        params[name] = min(scores, key=scores.get)
        batch.set_param(name, params[name])
    return {"params": params, "frame_ms": measure(), "trials": trials}
