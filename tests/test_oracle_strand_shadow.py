"""Strand optical depth in shadow rays pinned by closed forms of the
reference's own formula (primitives/sweep.h:97-148 strand_segment_tau /
strand_tau, bvh.h:928-960 strand_bvh_tau, render_kernel.cu:275-280
transmittance exp(-tau) with sigma = 6), each pin checked to fail with the
planted mistakes it targets (pin_util.run_pin).

A sun along v = (1, 1, 0)/sqrt(2) lights a Lambert floor; a near-orthographic
camera looks straight down at the floor point p = 0, whose shadow ray starts
at O = p + 0.002 n.  With w = O - p0 and u = p1 - p0 the reference solves
s = (B E - C D) / (A C - B^2), t = (B s + E) / C (A = u.u, B = u.v,
C = v.v, D = u.w, E = v.w), falls back to t = 0, s = clamp(-D/A) when
t < 0 (or s = 0 when the denominator vanishes) and measures
|w + u s - v t|: the distance between the segment O + s u and the half-line
p0 + t v, which is the usual closest approach with the signs of s and t
reversed.  Two closed forms follow:
* a strand lying along the shadow ray, from distance c to c + a (c <= a),
  of constant radius r: distance 0, tau = 6 * 2 r;
* a strand from O + c v + g z to that point + a v + b z (z perpendicular to
  v, 0 <= g <= b, a g / b >= c): s = g / b, t = a g / b - c, distance 0, so
  tau = 6 * 2 r(g / b) with r(s) = r0 + (r1 - r0) s;
* a strand crossing the shadow ray at right angles farther than its radius
  from O: the distance is at least that far, tau = 0 (the reference casts
  no strand shadow there; reproduced, not corrected).
pixel = A a + A I cos45 exp(-tau); depth stops adding once it exceeds 6, so
three strands of depth 4 give exp(-8), not exp(-12).
"""
import math

import numpy as np
import pytest

from sheen_amd import contract as K
from pin_util import camera, f32, gpu_render, hquad, light, mat, run_pin, scene, unit
from util import f3

A = np.array([0.8, 0.6, 0.4])
AMB = 0.1
L = unit((1.0, 1.0, 0.0))
O_Y = 0.002  # shadow-ray origin offset along the floor normal


def _strand(p0, p1, r0, r1):
    return (tuple(p0), tuple(p1), r0, r1)


def _along(c, a, r0, r1):
    """A strand on the shadow ray from distance c to c + a."""
    o = np.array([0.0, O_Y, 0.0])
    v = f32(L)
    return _strand(o + c * v, o + (c + a) * v, r0, r1)


def _oblique(c, alpha, beta, gamma, r0, r1):
    """A strand from O + c v + gamma z to that point + alpha v + beta z: the
    reference's s = gamma / beta, t = alpha gamma / beta - c >= 0, distance 0,
    so tau = 6 * 2 r(gamma / beta) (well conditioned: A C - B^2 = beta^2)."""
    o = np.array([0.0, O_Y, 0.0])
    v, z = f32(L), np.array([0.0, 0.0, 1.0])
    p0 = o + c * v + gamma * z
    return _strand(p0, p0 + alpha * v + beta * z, r0, r1)


def _across(t0, half=1.0, r=0.05):
    """A strand along z crossing the shadow ray at distance t0."""
    q = np.array([0.0, O_Y, 0.0]) + t0 * f32(L)
    return _strand(q - (0, 0, half), q + (0, 0, half), r, r)


def _strand_scene(strands):
    pts, st = [], []
    for (p0, p1, r0, r1) in strands:
        s = K.SweepStrand()
        s.point_offset, s.point_count, s.material_id = len(pts), 2, 1
        for q, r in ((p0, r0), (p1, r1)):
            p = K.SweepPoint()
            p.position, p.radius = f3(*q), r
            pts.append(p)
        st.append(s)
    sun = light("sun", direction=tuple(-np.float32(L)), intensity=1.0, angle=0.0)
    ps = scene([mat(color=A, amb=AMB), mat(color=(0.3, 0.3, 0.3))], hquad(0.0, -50, 50, -50, 50, 0), [sun],
               strands=(pts, st))
    camera(ps, (0, 0, 0), (0, -1, 0), dist=0.5)
    return ps


def _value(tau):
    return A * AMB + A * float(f32(L)[1]) * math.exp(-tau)


CASES = [
    # (strands, tau)
    ([_along(0.3, 1.0, 0.05, 0.05)], 6 * 2 * 0.05),
    ([_oblique(0.3, 2.0, 0.08, 0.02, 0.05, 0.10)], 6 * 2 * (0.05 + 0.25 * 0.05)),
    ([_oblique(0.2, 1.5, 0.10, 0.06, 0.08, 0.02)], 6 * 2 * (0.08 - 0.6 * 0.06)),
    ([_along(0.3, 1.0, 0.05, 0.05), _along(0.5, 2.0, 0.1, 0.1)], 6 * 2 * 0.15),
    ([_along(1.0, 1.5, 1 / 3, 1 / 3), _along(2.0, 2.5, 1 / 3, 1 / 3), _along(3.0, 3.5, 1 / 3, 1 / 3)], 8.0),
    ([_across(1.0)], 0.0),
    ([_across(0.5, r=0.2)], 0.0),
]


def test_strand_transmittance():
    def check(mode):
        for strands, tau in CASES:
            img = mode.render(_strand_scene(strands))
            np.testing.assert_allclose(img, np.broadcast_to(_value(tau), img.shape), rtol=2e-5,
                                       err_msg=str(strands))

    run_pin(check, ["CHORD_HALF", "RADIUS_REV", "TAU_NO_STOP", "SIGMA"])


@pytest.mark.gpu
def test_strand_pins_gpu_matches_oracle():
    from pin_util import Mode
    for strands, _ in CASES:
        ps = _strand_scene(strands)
        got = gpu_render(ps, 8, 8, 2)
        ref = Mode().render(ps, 8, 8, 2)
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-7, err_msg=str(strands))
