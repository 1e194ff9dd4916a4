"""Differentiable shading (sheen_adjoint_* through torch.autograd): forward
parity with the oracle, gradients against central finite differences of the
GPU forward (h = 1e-3) and of the oracle's FP64 forward.

The image is linear in every parameter group for fixed visibility, and the
parameters do not move geometry or shadows, so the L2 loss is quadratic in
each parameter and a central difference equals the derivative up to the
rounding of the rendered images.
"""
import os

import numpy as np
import pytest

from adjoint_util import loss_delta, scene_params

pytestmark = pytest.mark.gpu

W = H = 96
AA = 8
KEYS = ("albedo", "emission", "intensity", "texels")


@pytest.fixture(scope="module")
def setup():
    import torch
    from sheen_amd import adjoint, scenes
    sc = scenes.generate(5, detail=2000)
    r = adjoint.DiffRenderer(sc.scene, W, H, AA)
    p = r.params()
    rng = np.random.default_rng(5)
    tp = {k: (v * torch.tensor(1.0 + 0.2 * rng.uniform(-1, 1, tuple(v.shape)), dtype=torch.float32, device=v.device))
          for k, v in p.items()}
    with torch.no_grad():
        target = r.render(*(tp[k] for k in KEYS)).cpu().numpy()
    yield {"sc": sc, "r": r, "p": p, "target": target}
    r.close()


def _np(p):
    return {k: v.detach().cpu().numpy() for k, v in p.items()}


def _render(r, p):
    import torch
    with torch.no_grad():
        return r.render(*(p[k] for k in KEYS)).cpu().numpy()


def test_params_are_the_scene_values(setup):
    ref = scene_params(setup["sc"].scene)
    got = _np(setup["p"])
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k


def test_forward_matches_oracle_with_float_texels(setup):
    from oracle import Oracle
    img = _render(setup["r"], setup["p"])
    o = Oracle(setup["sc"].scene, threads=min(16, os.cpu_count() or 1))
    q = _np(setup["p"])
    o.set_params(q["albedo"], q["emission"], q["intensity"], q["texels"])
    ref = o.render(W, H, AA)["image"]
    rel = (np.abs(img[..., :3] - ref[..., :3]) / np.maximum(np.abs(ref[..., :3]), 1e-3)).max(-1)
    assert np.mean(rel > 1e-4) <= 0.001, rel.max()


def _loss_and_grads(setup, mask=None):
    """Gradients of sum(mask * (image - target)^2) over RGB."""
    import torch
    p = {k: v.clone().requires_grad_(True) for k, v in setup["p"].items()}
    img = setup["r"].render(*(p[k] for k in KEYS))
    t = torch.tensor(setup["target"], device=img.device)
    d2 = (img[..., :3] - t[..., :3]) ** 2
    if mask is not None:
        d2 = d2 * torch.tensor(mask[..., None].astype(np.float32), device=img.device)
    d2.sum().backward()
    return {k: p[k].grad.detach().cpu().numpy() for k in KEYS}


def _pick(g, rng):
    """Single entries to check: the largest gradients of each group plus
    random entries whose gradient is at least 1% of the group's largest.
    Texels are checked along directions only (below): one level-0 texel of a
    minified texture moves the loss by less than the float32 rounding of the
    images at h = 1e-3."""
    out = []
    for k, n_top, n_rand in (("albedo", 4, 2), ("emission", 3, 1), ("intensity", 3, 0)):
        flat = np.abs(g[k].ravel())
        assert flat.max() > 0, k
        nz = np.flatnonzero(flat >= 0.01 * flat.max())
        top = list(np.argsort(-flat)[:min(n_top, nz.size)])
        rest = [i for i in nz if i not in top]
        if rest and n_rand:
            top += list(rng.choice(rest, size=min(n_rand, len(rest)), replace=False))
        out += [(k, int(i)) for i in top]
    return out


def _directions(g, rng):
    """Per group: v = sign(g) on a random half of the entries, 0 elsewhere, so
    the directional derivative g.v sums |g| over a random subset."""
    out = []
    for k in KEYS:
        v = np.sign(g[k]) * (rng.random(g[k].shape) < 0.5)
        out.append((k, v.astype(np.float32)))
    return out


def _fd(render, base, target, key, idx, h):
    plus = {k: v.copy() for k, v in base.items()}
    minus = {k: v.copy() for k, v in base.items()}
    x = base[key].ravel()[idx]
    xp, xm = np.float32(x + h), np.float32(x - h)
    plus[key].reshape(-1)[idx] = xp
    minus[key].reshape(-1)[idx] = xm
    ip, im = render(plus), render(minus)
    d = (ip[..., :3].astype(np.float64) - im[..., :3]) / (float(xp) - float(xm))
    return loss_delta(ip, im, target) / (float(xp) - float(xm)), d


def _fd_dir(render, base, target, key, v, h):
    """Central difference along v; the derivative it estimates is g.v with
    the float32-rounded step actually taken per entry."""
    plus = {k: x.copy() for k, x in base.items()}
    minus = {k: x.copy() for k, x in base.items()}
    plus[key] = (base[key] + np.float32(h) * v).astype(np.float32)
    minus[key] = (base[key] - np.float32(h) * v).astype(np.float32)
    step = (plus[key].astype(np.float64) - minus[key].astype(np.float64)) / (2.0 * h)
    ip, im = render(plus), render(minus)
    d = (ip[..., :3].astype(np.float64) - im[..., :3]) / (2.0 * h)
    return loss_delta(ip, im, target) / (2.0 * h), step, d


def _check(render, g, base, target, h, seed, image_gap=None):
    """Adjoint gradient entries and directional derivatives against central
    differences of `render`.  image_gap: per-pixel |difference| between the
    image the adjoint differentiates and the image `render` produces at the
    base point; the loss gradients of the two then differ by up to
    2 * sum(image_gap * |d pixel / d parameter|), which widens the bound."""
    rng = np.random.default_rng(seed)

    def bound(a, d):
        extra = 0.0 if image_gap is None else 2.0 * float(np.sum(image_gap * np.abs(d)))
        return 2e-3 * abs(a) + 1e-4 + extra

    for key, idx in _pick(g, rng):
        fd, d = _fd(render, base, target, key, idx, h)
        a = float(g[key].ravel()[idx])
        assert abs(fd - a) <= bound(a, d), (key, idx, a, fd)
    for key, v in _directions(g, rng):
        fd, step, d = _fd_dir(render, base, target, key, v, h)
        a = float(np.sum(g[key].astype(np.float64) * step))
        assert abs(fd - a) <= bound(a, d), (key, a, fd)


def test_gradients_match_gpu_central_differences(setup):
    import torch
    g = _loss_and_grads(setup)
    base = _np(setup["p"])
    dev = setup["p"]["albedo"].device

    def render(p):
        return _render(setup["r"], {k: torch.tensor(v, device=dev) for k, v in p.items()})

    _check(render, g, base, setup["target"], 1e-3, 1)


def test_gradients_match_fp64_oracle(setup):
    """FP64 reference gradient with visibility held fixed.  The oracle's FP64
    forward also intersects and looks up textures in double, which moves a
    few samples at silhouettes, shadow edges and mip-level boundaries (the
    pixel changes by percents) and shifts bilinear weights slightly
    elsewhere.  The loss is restricted, on both sides, to pixels whose FP32
    and FP64 oracle renders agree to 5e-3; the remaining image difference
    enters the bound (see _check)."""
    from oracle import Oracle
    base = _np(setup["p"])
    o = Oracle(setup["sc"].scene, threads=min(16, os.cpu_count() or 1))
    o.set_params(base["albedo"], base["emission"], base["intensity"], base["texels"])
    i32 = o.render(W, H, AA)["image"][..., :3].astype(np.float64)
    i64 = o.render(W, H, AA, fp64=True)["image"][..., :3].astype(np.float64)
    mask = np.all(np.abs(i64 - i32) <= 5e-3 * np.maximum(np.abs(i32), 1e-2), -1)
    assert mask.mean() > 0.95, mask.mean()
    g = _loss_and_grads(setup, mask)
    gap = np.abs(i64 - i32) * mask[..., None]

    def render(p):
        o.set_params(p["albedo"], p["emission"], p["intensity"], p["texels"])
        img = o.render(W, H, AA, fp64=True)["image"].copy()
        img[~mask] = setup["target"][~mask]  # excluded pixels add nothing to the loss
        return img

    _check(render, g, base, setup["target"], 0.05, 2, image_gap=gap)


def test_context_keeps_its_own_scene(setup):
    import ctypes as C
    from sheen_amd import sheen
    r = setup["r"]
    img = np.empty((H, W, 4), np.float32)
    assert r._lib.sheen_context_render(r._c, W, H, AA, img.ctypes.data_as(C.c_void_p), None) == 0
    with sheen.Context(setup["sc"].scene) as ctx:
        ref, _ = ctx.render(W, H, AA)
    assert np.array_equal(img, ref)


def test_non_differentiable_scene_is_refused():
    import torch
    from sheen_amd import adjoint, scenes
    sc = scenes.generate(2, detail=10)
    with adjoint.DiffRenderer(sc.scene, 32, 32, 1) as r:
        p = r.params()
        with pytest.raises(RuntimeError, match="not differentiable: material with reflection"):
            r.render(*(p[k] for k in KEYS))
        torch.cuda.synchronize()


def _ordering_run(setup, stream=None):
    """Parameters written by torch kernels immediately before the forward and
    backward calls, outputs read by torch kernels immediately after, with no
    synchronisation in between, on torch's current stream (the null stream)
    or on a side stream."""
    import torch
    r, p = setup["r"], setup["p"]
    torch.cuda.synchronize()
    ctx = torch.cuda.stream(stream) if stream is not None else torch.cuda.stream(torch.cuda.current_stream())
    with ctx:
        # a long chain of small kernels writes the parameters: a render that
        # started before them would read stale or partial values
        q = {k: torch.zeros_like(v) for k, v in p.items()}
        for k in KEYS:
            for _ in range(20):
                q[k].add_(p[k], alpha=0.05)
            q[k].requires_grad_(True)
        img = r.render(*(q[k] for k in KEYS))
        out = (img * 2.0).sum()  # read by a torch kernel right away
        t = torch.tensor(setup["target"], device=img.device)
        ((img[..., :3] - t[..., :3]) ** 2).sum().backward()
        grads = {k: (q[k].grad * 1.0) for k in KEYS}
    torch.cuda.synchronize()
    return float(out), {k: v.cpu().numpy() for k, v in grads.items()}


def test_caller_stream_ordering_without_synchronisation(setup):
    """The library orders its stream after the caller's on entry and the
    caller's after its own on return (null stream and a side stream), so the
    unsynchronised runs equal a run with full synchronisation (the image bit
    for bit), and no ordering fell back to host synchronisation."""
    import torch
    r, p = setup["r"], setup["p"]
    before = r.stream_fallbacks()
    ref_q = {k: torch.zeros_like(v) for k, v in p.items()}
    for k in KEYS:
        for _ in range(20):
            ref_q[k].add_(p[k], alpha=0.05)
        ref_q[k].requires_grad_(True)
    torch.cuda.synchronize()
    img = r.render(*(ref_q[k] for k in KEYS))
    torch.cuda.synchronize()
    ref_out = float((img * 2.0).sum())
    t = torch.tensor(setup["target"], device=img.device)
    ((img[..., :3] - t[..., :3]) ** 2).sum().backward()
    torch.cuda.synchronize()
    ref_g = {k: ref_q[k].grad.cpu().numpy() for k in KEYS}
    for stream in (None, torch.cuda.Stream()):
        out, g = _ordering_run(setup, stream)
        assert out == ref_out
        for k in KEYS:
            # gradient sums use float atomics: equal up to summation order
            scale = float(np.abs(ref_g[k]).max()) if ref_g[k].size else 0.0
            np.testing.assert_allclose(g[k], ref_g[k], rtol=1e-5, atol=1e-6 * scale, err_msg=k)
    assert r.stream_fallbacks() == before == 0
