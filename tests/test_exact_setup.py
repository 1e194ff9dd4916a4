"""Float identities the exact set-ups of the HIP path rely on.

hit_cone_y (sheen_amd/csrc/dev_prims.h) replaces the frustum set-up of the
reference's intersect_cone_segment (primitives/cone.h:11-30) for lathe
frusta on the y axis, L = sqrtf(ay*ay) and day = ay / L, by L = |ay| and
day = +-1.  That is exact when sqrtf of a rounded square is the magnitude
and a / |a| is +-1; both are checked here for every float32 of [1, 4),
which covers every exponent (a power-of-two scale passes exactly through
the square and the square root while the square stays a normal float).
"""
import numpy as np


def test_sqrt_of_rounded_square_is_the_magnitude():
    lo, hi = np.float32(1.0).view(np.uint32), np.float32(4.0).view(np.uint32)
    x = np.arange(lo, hi, dtype=np.uint32).view(np.float32)
    for a in (x, -x):
        sq = np.sqrt(a * a)
        assert sq.dtype == np.float32
        assert np.array_equal(sq, np.abs(a))
        assert np.all(a / sq == np.sign(a))


def test_the_identity_holds_at_scaled_binades():
    rng = np.random.default_rng(7)
    m = rng.uniform(1.0, 2.0, 200_000).astype(np.float32)
    for e in (-20, -10, -1, 3, 17, 40, 63):
        a = np.ldexp(m, e).astype(np.float32)
        assert np.array_equal(np.sqrt(a * a), a)


def test_the_check_catches_a_non_identity():
    # the same check fails where the identity does not hold: the square root
    # of a rounded cube is not a power of the input in general
    x = np.linspace(1, 4, 10_001, dtype=np.float32)
    assert not np.array_equal(np.sqrt(x * x * x) * np.sqrt(x * x * x), x * x * x)
