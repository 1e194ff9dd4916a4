"""The float transcendentals shared by the HIP path and the oracle
(sheen_amd/csrc/sheen_math.h) against the host's double-precision libm
(numpy float64), rounded to float: the series are accurate to about 1e-16,
so the float result must equal the correctly rounded value except where the
true value lies within that distance of a rounding boundary (none expected
in these samples), and never be more than 1 ulp away.  Domains as the
render path uses them: angles and 2 pi u for sin, cos and tan, any exponent
for exp (exp(-tau) of optical depths), x >= 1 for log2 (texture LOD), and
x in [0, 1] with y in [1, 4096] for pow (Blinn-Phong and hair lobes), plus
spot checks of the special values.  Measured: every sample equals the
rounded libm value.
"""
import numpy as np
import pytest

N = 200000


def _f(fn, x, y=None):
    from oracle import lib
    L = lib()
    y = np.zeros_like(x) if y is None else y
    return np.array([L.oracle_math(fn, float(a), float(b)) for a, b in zip(x, y)], np.float32)


def _ulps(a, b):
    a = a.astype(np.float32).view(np.int32).astype(np.int64)
    b = b.astype(np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


CASES = {
    "sin": (0, lambda r: np.concatenate([r.uniform(0, 2 * np.pi, N // 2), r.uniform(-40, 40, N // 2)]), np.sin),
    "cos": (1, lambda r: np.concatenate([r.uniform(0, 2 * np.pi, N // 2), r.uniform(-40, 40, N // 2)]), np.cos),
    "tan": (2, lambda r: r.uniform(-1.5, 1.5, N), np.tan),
    "exp": (3, lambda r: np.concatenate([-r.exponential(3.0, N // 2), r.uniform(-100, 80, N // 2)]), np.exp),
    "log2": (4, lambda r: np.exp(r.uniform(0, 20, N)), np.log2),
}


@pytest.mark.parametrize("name", list(CASES))
def test_against_libm(name):
    fn, gen, ref = CASES[name]
    x = gen(np.random.default_rng(fn + 1)).astype(np.float32)
    got = _f(fn, x)
    want = ref(x.astype(np.float64)).astype(np.float32)
    u = _ulps(got, want)
    assert u.max() <= 1, (name, x[np.argmax(u)])
    assert np.mean(u == 0) >= 0.9999, np.mean(u == 0)


def test_pow_against_libm():
    r = np.random.default_rng(9)
    x = r.uniform(0, 1, N).astype(np.float32)
    y = np.exp(r.uniform(0, np.log(4096), N)).astype(np.float32)
    got = _f(5, x, y)
    want = np.power(x.astype(np.float64), y.astype(np.float64)).astype(np.float32)
    u = _ulps(got, want)
    assert u.max() <= 1
    assert np.mean(u == 0) >= 0.9999


def test_special_values():
    assert _f(0, np.array([0.0], np.float32))[0] == 0.0
    assert _f(1, np.array([0.0], np.float32))[0] == 1.0
    assert _f(3, np.array([0.0], np.float32))[0] == 1.0
    assert _f(3, np.array([-200.0], np.float32))[0] == 0.0
    assert np.isinf(_f(3, np.array([100.0], np.float32))[0])
    assert _f(4, np.array([1.0], np.float32))[0] == 0.0
    assert _f(4, np.array([8.0], np.float32))[0] == 3.0
    assert _f(5, np.array([0.0], np.float32), np.array([2.0], np.float32))[0] == 0.0
    assert _f(5, np.array([0.5], np.float32), np.array([0.0], np.float32))[0] == 1.0
    assert _f(5, np.array([0.5], np.float32), np.array([3.0], np.float32))[0] == 0.125
