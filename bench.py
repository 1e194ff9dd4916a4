#!/usr/bin/env python3
"""Throughput of the MI355X render path on the 1M-triangle 1080p workload.

A step is one frame: 1920x1080 pixels x 4 jittered samples of the
1,048,820-triangle displaced geodesic sphere (70% diffuse / 15% mirror / 15%
glass, 2 point lights + 1 area light), traced through the full reference path
(primary, shadow passes, reflection, up to three refraction faces).  `value`
is every traversal call of the timed frames divided by their wall time, with
the scene already resident in HBM.

Also measured (secondary, in the same JSON line): "scene_setup", the time
to validate, build the BVHs of and upload the 1M-triangle scene with the
device and with the host builder; "analytic", Mrays/s of the 10k
analytic-primitive + CSG + caustics workload at 1280x720 x 16 spp; "batch",
the 256-scene uneven batch as one pipeline against one render per scene
(speedup); and "adjoint", forward and backward time of the
differentiable-shading workload (adjoint / forward).

Prints one JSON line (see --help).  --gpus above 1 is refused: the path runs
on one GPU.
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s, 1M-tri 1080p; batched-vs-serial scene speedup; adjoint/forward time"
HBM_PEAK_GBS = 8000.0
PARAMS = {}  # --param knobs
NODE_BYTES = 128   # one wide BVH node per fetch
TRI_LEAF_BYTES = 48  # vertices + index of one triangle leaf slot
# the device code the committed counter profiles were measured on
KERNEL_SOURCES = ["render.hip", "kernels.h", "dev_trace.h", "dev_prims.h", "dev_shade.h", "dev_common.h",
                  "sheen_math.h", "wbvh.h", "wbvh_rules.h", "wbvh_gpu.hip"]


def kernel_sources_sha():
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "sheen_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def counter_profile(kernel, avg_launch_ms):
    """The committed rocprofv3 counter summary of `kernel` on this workload
    (profiles/counters_<kernel>_mesh1m.json): HBM bytes per launch from
    FETCH_SIZE / WRITE_SIZE, L2 hit rate, VALU issue share and lane
    utilisation, with the hash of the device sources it was measured on and
    whether that is the code running now."""
    path = os.path.join(ROOT, "profiles", "counters_%s_mesh1m.json" % kernel)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        c = json.load(f)
    traffic = c.get("hbm_bytes_per_launch")
    out = {
        "source": os.path.relpath(path, ROOT),
        "measured_on_sources": c.get("sources_sha"),
        "sources_current": c.get("sources_sha") == kernel_sources_sha(),
        "hbm_bytes_per_launch": traffic,
        "hbm_frac": round(traffic / (avg_launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
        "l2_hit": c.get("l2_hit"),
        "valu_issue_frac": c.get("valu_issue_frac"),
        "lane_utilisation": c.get("lane_utilisation"),
        "limiter": c.get("limiter"),
    }
    return traffic, out


def cpu_baseline(scene, width, height, aa, budget_s=20.0):
    """The oracle (scalar C++ restatement, OpenMP over rows) on the same frame,
    on the host cores of this box: the whole frame when it fits the time
    budget, else a band of rows centred on the mesh."""
    from oracle import Oracle
    threads = min(16, os.cpu_count() or 1)
    o = Oracle(scene, threads=threads)
    probe = 8
    y0 = height // 2 - probe // 2
    t = time.perf_counter()
    o.render(width, height, aa, rows=(y0, y0 + probe))
    dt = time.perf_counter() - t
    rows = min(height, max(probe, int(probe * budget_s / max(dt, 1e-3))))
    y0 = (height - rows) // 2
    t = time.perf_counter()
    st = o.render(width, height, aa, rows=(y0, y0 + rows))["stats"]
    dt = time.perf_counter() - t
    o.close()
    what = "the whole %dx%d x%d-spp frame" % (width, height, aa) if rows == height else \
        "rows %d..%d of the %dx%d x%d-spp frame" % (y0, y0 + rows - 1, width, height, aa)
    return {
        "value": st["total_rays"] / dt / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": "%s (%d samples, %d rays, %.1f s)" % (what, st["samples"], st["total_rays"], dt),
    }


def adjoint_leg(steps):
    """Forward and backward (path replay) of the differentiable-shading
    workload: 512x512, 8 spp, 10,000 spheres with their own albedo, a
    1024x1024 value-noise texture on the ground plane, 3 lights; gradients of
    an L2 image loss for albedo, emission, light intensity and texels."""
    import torch
    from sheen_amd import adjoint, scenes
    keys = ("albedo", "emission", "intensity", "texels")
    sc = scenes.generate(5)
    r = adjoint.DiffRenderer(sc.scene, sc.width, sc.height, sc.aa)
    p = {k: v.clone().requires_grad_(True) for k, v in r.params().items()}
    target = None
    with torch.no_grad():
        target = r.render(*(p[k] * 1.1 for k in keys))
    for _ in range(2):
        img = r.render(*(p[k] for k in keys))
        ((img - target) ** 2).sum().backward()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fw = bw = 0.0
    for _ in range(steps):
        ev[0].record()
        img = r.render(*(p[k] for k in keys))
        ev[1].record()
        loss = ((img - target) ** 2).sum()
        torch.autograd.grad(loss, [p[k] for k in keys])
        ev[2].record()
        torch.cuda.synchronize()
        fw += ev[0].elapsed_time(ev[1]) / steps
        bw += ev[1].elapsed_time(ev[2]) / steps
    r.close()
    return {"workload": "diff-shading-512: %d spheres, 1024x1024 texture, %dx%d, %d spp, 3 lights"
                        % (sc.scene.contents.num_objects, sc.width, sc.height, sc.aa),
            "forward_ms": round(fw, 3), "backward_ms": round(bw, 3), "adjoint_over_forward": round(bw / fw, 3),
            "parameters": {k: int(v.numel()) for k, v in p.items()}}


def analytic_leg(steps):
    """The analytic-primitive workload: 10,000 seeded spheres, cones,
    superellipsoids, Bezier lathes, 12-point prisms and sweeps, one
    1024x1024 fBm height field, CSG trees over half of the spheres and
    superellipsoids, glass spheres casting caustics from a photon map;
    1280x720 at 16 spp.  Mrays/s over the frame; the photon pass runs once
    when the scene is created."""
    from sheen_amd import scenes, sheen
    sc = scenes.generate(3)
    t = time.perf_counter()
    ctx = sheen.Context(sc.scene)
    setup_s = time.perf_counter() - t
    # the primary shading's shadow rays on a second stream during the
    # secondary rays: faster on this workload (autotuner knob "overlap")
    ctx.set_param("overlap", 1)
    for k, v in PARAMS.items():
        ctx.set_param(k, v)
    W, H, A = sc.width, sc.height, sc.aa
    ctx.render(W, H, A, download=False)
    rays = 0
    t = time.perf_counter()
    for _ in range(steps):
        _, st = ctx.render(W, H, A, download=False)
        rays += st["total_rays"]
    dt = time.perf_counter() - t
    emitted, stored = ctx.photon_stats()
    ctx.close()
    c = sc.scene.contents
    return {"workload": "analytic-10k-720p: %d analytic primitives, %d CSG nodes, %dx%d, %d spp, caustics"
                        % (c.num_objects + c.num_superellipsoids + c.num_lathes + c.num_prisms + c.num_sweep_strands,
                           c.num_csg_nodes, W, H, A),
            "mrays_s": round(rays / dt / 1e6, 1), "ms_per_frame": round(dt / steps * 1e3, 2),
            "rays_per_frame": rays // steps, "photons_emitted": emitted, "photons_stored": stored,
            "scene_setup_s": round(setup_s, 3), "side_stream_shadows": True}


def batch_leg(steps):
    """The uneven batch: 256 scenes (triangle counts log-uniform in
    [1e3, 2e5], square images of 128-1024 pixels, 1-8 spp) rendered as one
    wavefront pipeline replayed from a hipGraph, against one render per
    scene through the same resident contexts (wall clock of each call)."""
    from sheen_amd import scenes
    from sheen_amd.batch import Batch
    sc = scenes.batch(256)
    b = Batch(sc)
    for k, v in PARAMS.items():
        b.set_param(k, v)
    b.render()
    b.render_serial()
    bt = st = 0.0
    rays = 0
    for _ in range(steps):
        s = b.render()
        bt += s["trace_ms"] / steps
        rays = s["total_rays"]
    for _ in range(steps):
        st += b.render_serial()["trace_ms"] / steps
    out = {"workload": "batch-256: %d scenes, %.1fM triangles, %.0fM samples" %
                       (b.n, sum(s.scene.contents.num_triangles for s in sc) / 1e6, b.samples / 1e6),
           "batched_ms": round(bt, 2), "serial_ms": round(st, 2), "speedup": round(st / bt, 3),
           "batched_mrays_s": round(rays / bt / 1e3, 1), "chunks": b.chunks}
    b.close()
    return out


def batch_cap_leg(reps=3):
    """Where the batch speedup stops: batched against one-render-per-scene
    wall time for the whole 256-scene batch, for its 1024x1024 scenes alone
    and for the other scenes alone (medians of `reps` frames each).  The
    large scenes fill the GPU on their own, so serial(all) / batched(large)
    bounds the speedup even if the small scenes cost nothing."""
    import statistics
    from sheen_amd import scenes
    from sheen_amd.batch import Batch
    sc = scenes.batch(256)
    out = {}
    for name, keep in (("all", lambda s: True), ("large", lambda s: s.width == 1024),
                       ("small", lambda s: s.width != 1024)):
        sub = [s for s in sc if keep(s)]
        b = Batch(sub)
        b.render()
        b.render_serial()
        bt = statistics.median(b.render()["trace_ms"] for _ in range(reps))
        st = statistics.median(b.render_serial()["trace_ms"] for _ in range(reps))
        out[name] = {"scenes": len(sub), "samples": b.samples, "batched_ms": round(bt, 2), "serial_ms": round(st, 2),
                     "speedup": round(st / bt, 3), "batched_ns_per_sample": round(bt * 1e6 / b.samples, 4),
                     "serial_ns_per_sample": round(st * 1e6 / b.samples, 4)}
        b.close()
    out["speedup_bound_if_small_scenes_were_free"] = round(out["all"]["serial_ms"] / out["large"]["batched_ms"], 3)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-shading leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-scene batch leg")
    ap.add_argument("--no-analytic", action="store_true", help="skip the analytic-primitive (10k + CSG + caustics) leg")
    ap.add_argument("--detail", type=int, default=0, help="geodesic frequency override (default 229)")
    ap.add_argument("--param", action="append", default=[], metavar="NAME=VALUE",
                    help="launch or queue-order knob (sheen_context_set_param / sheen_batch_set_param) for the "
                         "mesh, analytic and batch workloads; repeatable")
    ap.add_argument("--leg", choices=("analytic", "batch", "adjoint"),
                    help="run only this secondary leg and print its JSON object (a workload to profile on its own)")
    ap.add_argument("--batch-cap", action="store_true",
                    help="only measure the batch speedup bound (whole batch, large scenes, small scenes)")
    args = ap.parse_args()
    global PARAMS
    PARAMS = {}
    for kv in args.param:
        k, _, v = kv.partition("=")
        PARAMS[k] = int(v)
    if args.gpus != 1:
        print("bench.py: this render path runs on one GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2
    if args.batch_cap:
        print(json.dumps({"batch_cap": batch_cap_leg()}), flush=True)
        return 0
    if args.leg:
        if args.leg == "adjoint":
            import torch
            torch.cuda.init()
        leg = {"analytic": lambda: analytic_leg(max(2, min(args.steps, 5))),
               "batch": lambda: batch_leg(max(2, min(args.steps, 3))),
               "adjoint": lambda: adjoint_leg(max(3, min(args.steps, 10)))}[args.leg]
        print(json.dumps({args.leg: leg()}), flush=True)
        return 0

    if not args.no_adjoint:
        # torch ships its own HIP runtime; it must initialise before this
        # library's (ROCm's) when both are used in one process
        import torch
        torch.cuda.init()
    from sheen_amd import scenes, sheen

    kw = {"detail": args.detail} if args.detail else {}
    sc = scenes.generate(2, **kw)
    W, H, A = sc.width, sc.height, sc.aa
    ntri = sc.scene.contents.num_triangles
    # scene setup (validation, bounds, BVH build, upload): device build, then
    # the host builder for comparison
    with sheen.Context(sc.scene, host_bvh=True) as hctx:
        host_build_ms = hctx.build_ms
    ctx = sheen.Context(sc.scene)
    setup = {"device_bvh_ms": round(ctx.build_ms, 1), "host_bvh_ms": round(host_build_ms, 1)}
    for k, v in PARAMS.items():
        ctx.set_param(k, v)

    for _ in range(args.warmup):
        ctx.render(W, H, A, download=False)
    # untimed counting pass: nodes fetched / primitives tested per kernel
    ctx.set_counting(True)
    _, cst = ctx.render(W, H, A, download=False)
    ctx.set_counting(False)

    ctx.set_profiling(True)
    tot = {"rays": 0, "frame_ms": 0.0, "primary_ms": 0.0, "trace_ms": 0.0, "shadow_ms": 0.0, "shade_ms": 0.0,
           "shadow_cont_ms": 0.0,
           "shadow_launches": 0, "trace_launches": 0}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, st = ctx.render(W, H, A, download=False)
        tot["rays"] += st["total_rays"]
        for k in ("frame_ms", "primary_ms", "trace_ms", "shadow_ms", "shade_ms", "shadow_cont_ms", "shadow_launches",
                  "trace_launches"):
            tot[k] += st[k]
    t1 = time.perf_counter()
    ctx.set_profiling(False)
    wall = t1 - t0
    value = tot["rays"] / wall / 1e6

    # roofline of the dominant kernel (largest summed time over the timed frames);
    # one shadow and one primary launch per frame chunk
    chunks = max(1, tot["shadow_launches"] // args.steps)
    kernels = {
        "k_shadow": (tot["shadow_ms"], cst["shadow_node_fetches"], cst["shadow_prim_tests"]),
        "k_primary": (tot["primary_ms"], cst["primary_node_fetches"], cst["primary_prim_tests"]),
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    ms, frame_nodes, frame_prims = kernels[dom]
    frame_bytes = frame_nodes * NODE_BYTES + frame_prims * TRI_LEAF_BYTES
    bytes_per_launch = frame_bytes / chunks
    avg_launch_ms = ms / (chunks * args.steps)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    traffic, counters = counter_profile(dom, avg_launch_ms)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded geodesic sphere, fBm displacement)",
        "config": {
            "workload": "mesh-1M-1080p: %d triangles, %dx%d, %d spp, 2 point + 1 area light, 30%% mirror/glass"
                        % (ntri, W, H, A),
            "triangles": ntri, "width": W, "height": H, "spp": A,
            "rays_per_frame": tot["rays"] // args.steps,
            "gpu_frame_ms": round(tot["frame_ms"] / args.steps, 3),
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            # the unit the counters show limiting the kernel (counters.limiter)
            "limiter": (counters or {}).get("limiter"),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": int(bytes_per_launch),
            "node_fetches_per_launch": frame_nodes // chunks,
            "prim_tests_per_launch": frame_prims // chunks,
            # counting frame: share of a shadow-kernel wave's lanes that visit a
            # node per traversal iteration (the rest have finished their ray)
            "shadow_traversal_lane_use": round(cst["shadow_node_fetches"] / max(1, cst["shadow_visit_lanes"]), 4),
            "avg_launch_ms": round(avg_launch_ms, 4),
            # what the counters say limits the kernel: "achieved" counts the
            # algorithmic node and triangle bytes, most of which the caches serve
            "counters": counters,
            "note": "achieved = algorithmic bytes (128 B per node fetch, 48 B per triangle test, counted by the "
                    "counting frame) per launch / launch time; frac above 1 means the caches serve most of them: "
                    "counters.hbm_frac is the HBM traffic measured by FETCH_SIZE/WRITE_SIZE and counters.limiter "
                    "the unit that bounds the kernel",
        },
        "kernel_ms_per_step": {k: round(tot[k] / args.steps, 3) for k in ("primary_ms", "trace_ms", "shadow_ms",
                                                                        "shadow_cont_ms", "shade_ms")},
    }
    if PARAMS:
        line["params"] = PARAMS
    line["scene_setup"] = setup
    # Not part of `value`: the same frames with the shadow requests whose
    # unshadowed contribution is exactly zero left untraced (knob skip_dark;
    # identical images, fewer rays than the reference traces)
    if "skip_dark" not in PARAMS:
        ctx.set_param("skip_dark", 1)
        ctx.render(W, H, A, download=False)
        rays = 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, st = ctx.render(W, H, A, download=False)
            rays += st["total_rays"]
        dt = time.perf_counter() - t0
        ctx.set_param("skip_dark", 0)
        line["skip_dark"] = {"ms_per_frame": round(dt / args.steps * 1e3, 3), "rays_per_frame": rays // args.steps,
                             "note": "same images; shadow rays of lights behind the shading normal not traced"}
    ctx.close()
    if not args.no_analytic:
        line["analytic"] = analytic_leg(max(2, min(args.steps, 5)))
    if not args.no_batch:
        line["batch"] = batch_leg(max(2, min(args.steps, 3)))
    if not args.no_adjoint:
        line["adjoint"] = adjoint_leg(max(3, min(args.steps, 10)))
    if not args.no_cpu_baseline:
        cb = cpu_baseline(sc.scene, W, H, A)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = round(value / cb["value"], 1)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
