"""The distance-space child-box test of the wide-BVH traversal must contain
the exact-path interval of every child (so the node test stays a superset
of the reference slab test on primitive bounds), with the per-ray margin
E_a = 2^-20 |inv_a| (K_a + |ray_a|) + 2^-100 of trav_begin, K_a bounding
2|o_a| + 2 Q s_a over the BVH's nodes.

Child planes are integers q in 0..2048 held as fp16 patterns; the
traversal's mixed-precision FMA takes q exactly.  Float32 emulation: + - *
are IEEE-rounded by numpy; fmaf(q, B, A) is emulated as an exact float64
product plus a float64 sum rounded to float32 (double rounding can move a
result by one ulp in rare halfway cases, far inside the 16x margin of the
bound).
"""
import numpy as np

F = np.float32
QMAX = 2048


def fma32(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F)


def exact_path(q_lo, q_hi, o, s, ro, inv):
    lo = fma32(q_lo, s, o)
    hi = fma32(q_hi, s, o)
    a = ((lo - ro).astype(F) * inv).astype(F)
    b = ((hi - ro).astype(F) * inv).astype(F)
    return np.minimum(a, b), np.maximum(a, b)


def box_k(o, s):
    """The per-BVH bound of k_box_k (render.hip), float32: 2|o| + 2 Q s.
    The traversal takes the largest over the nodes; the test gives every
    node the smallest value that bound can have, its own."""
    return (F(2.0) * np.abs(o) + (F(2.0 * QMAX) * s).astype(F)).astype(F)


def distance_space(q_lo, q_hi, o, s, ro, inv):
    A = ((o - ro).astype(F) * inv).astype(F)
    B = (s * inv).astype(F)
    # per-ray margin of trav_begin (dev_trace.h)
    xa = (np.abs(inv) * (box_k(o, s) + np.abs(ro)).astype(F)).astype(F)
    E = fma32(xa, F(2.0 ** -20) * np.ones_like(xa), F(2.0 ** -100) * np.ones_like(xa))
    pos = inv >= 0
    qn = np.where(pos, q_lo, q_hi)
    qf = np.where(pos, q_hi, q_lo)
    near = fma32(qn, B, (A - E).astype(F))
    far = fma32(qf, B, (A + E).astype(F))
    return near, far


def _case(rng, n, scale):
    o = (rng.uniform(-1, 1, n) * scale).astype(F)
    exp = rng.integers(-30, 12, n)
    s = np.ldexp(np.ones(n), exp).astype(F)
    q_lo = rng.integers(0, QMAX + 1, n).astype(F)
    q_hi = np.minimum(q_lo + rng.integers(0, QMAX + 1, n), QMAX).astype(F)
    d = rng.normal(size=n).astype(F)
    d = np.where(rng.random(n) < 0.1, (d * 1e-6).astype(F), d)
    inv = (F(1.0) / d).astype(F)
    # ray origins: anywhere, or within a few ulps of a decoded box plane
    ro = (o + rng.normal(size=n).astype(F) * F(scale)).astype(F)
    plane = fma32(q_lo, s, o)
    near_plane = rng.random(n) < 0.3
    ro = np.where(near_plane, (plane + (rng.integers(-3, 4, n) * np.spacing(plane))).astype(F), ro)
    return q_lo, q_hi, o, s, ro, inv


def test_distance_space_interval_contains_exact_interval():
    rng = np.random.default_rng(7)
    for scale in (1e-3, 1.0, 30.0, 1e4, 1e7):
        args = _case(rng, 200000, scale)
        en, ef = exact_path(*args)
        fn, ff = distance_space(*args)
        ok = np.isfinite(en) & np.isfinite(ef) & np.isfinite(fn) & np.isfinite(ff)
        assert ok.mean() > 0.99
        assert np.all(fn[ok] <= en[ok]), scale
        assert np.all(ff[ok] >= ef[ok]), scale


def test_margin_is_needed():
    """Without the E offsets the distance-space planes fall on both sides of
    the exact-path planes, so the margin is load-bearing."""
    rng = np.random.default_rng(11)
    q_lo, q_hi, o, s, ro, inv = _case(rng, 200000, 30.0)
    en, _ = exact_path(q_lo, q_hi, o, s, ro, inv)
    A = ((o - ro).astype(F) * inv).astype(F)
    B = (s * inv).astype(F)
    near = fma32(np.where(inv >= 0, q_lo, q_hi), B, A)
    assert np.any(near > en) and np.any(near < en)


def test_plane_integers_are_exact_in_fp16():
    """Every plane value 0..2048 survives the fp16 pattern the nodes store."""
    q = np.arange(QMAX + 1)
    h = q.astype(np.float16)
    assert np.array_equal(h.astype(np.float64), q.astype(np.float64))
