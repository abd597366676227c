"""Crossing test of a segment with a cylinder (xy) or sphere surface.

Pins the oracle's hit_test (restating Cylinder_XY::intersect,
src/primitives/cylinderXY.cuh:82-122, and Sphere::intersect, Sphere.cuh:102-139)
against closed-form roots computed in float64, and against special cases:
pure-z steps never cross a cylinder, tangent and inward/outward starts, and
the side-tracking rules that replace the reference's epsilon test.
"""
import math

import numpy as np

from axonexa_amd import abi
from oracle import pyoracle as po


def _roots64(ptype, c, R, s, u):
    s, u, c = np.asarray(s, float), np.asarray(u, float), np.asarray(c, float)
    o = s - c
    if ptype == abi.PRIM_CYLINDER_XY:
        o, u = o[:2], u[:2]
    A, B, C = u @ u, o @ u, o @ o - R * R
    disc = B * B - A * C
    if A == 0 or disc < 0:
        return None
    sq = math.sqrt(disc)
    return (-B - sq) / A, (-B + sq) / A, C


def _rand_unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def test_hit_test_matches_float64_roots():
    rng = np.random.default_rng(1)
    checked = 0
    for ptype in (abi.PRIM_CYLINDER_XY, abi.PRIM_SPHERE):
        for _ in range(4000):
            c = rng.uniform(-5, 5, 3)
            R = rng.uniform(0.5, 4.0)
            u = _rand_unit(rng)
            ell = rng.uniform(0.01, 1.0)
            # start near the surface, inside or outside
            d = _rand_unit(rng)
            if ptype == abi.PRIM_CYLINDER_XY:
                d[2] = 0.0
                d /= np.linalg.norm(d)
            rad = R + rng.uniform(-ell, ell)
            s = c + d * rad + (np.array([0, 0, rng.uniform(-3, 3)]) if ptype == 0 else 0)
            inside = rad < R
            r = _roots64(ptype, c, R, s, u)
            hit, t = po.hit_test(ptype, c.astype(np.float32), np.float32(R), inside,
                                 s.astype(np.float32), u.astype(np.float32), np.float32(ell))
            if r is None:
                assert not hit or t == 0.0
                continue
            t1, t2, C = r
            expect = t2 if inside else t1
            exp_hit = (0.0 <= expect < ell) if inside else (t1 >= 0.0 and t1 < ell)
            margin = 1e-4 * (1 + abs(expect))
            if abs(expect - ell) < margin or abs(expect) < margin:
                continue  # float32 decides boundary cases
            assert hit == exp_hit, (ptype, s, u, c, R, ell, inside, expect)
            if hit:
                assert abs(t - expect) <= 2e-5 * (1 + R / max(np.linalg.norm(u[:2] if ptype == 0 else u), 1e-3))
                checked += 1
    assert checked > 1000


def test_pure_z_step_never_crosses_cylinder():
    hit, _ = po.hit_test(abi.PRIM_CYLINDER_XY, (0, 0, 0), 1.0, True, (0.999, 0, 0), (0, 0, 1), 0.5)
    assert not hit
    hit, _ = po.hit_test(abi.PRIM_CYLINDER_XY, (0, 0, 0), 1.0, False, (1.001, 0, 0), (0, 0, -1), 0.5)
    assert not hit


def test_side_clamp_on_surface():
    # On the surface (numerically either side), tracked inside, moving outward:
    # reflect immediately (t = 0); moving inward: exit at the far chord end.
    for x in (0.9999999, 1.0, 1.0000001):
        hit, t = po.hit_test(abi.PRIM_SPHERE, (0, 0, 0), 1.0, True, (x, 0, 0), (1, 0, 0), 0.1)
        assert hit and t < 1e-6
        hit, t = po.hit_test(abi.PRIM_SPHERE, (0, 0, 0), 1.0, True, (x, 0, 0), (-1, 0, 0), 3.0)
        assert hit and abs(t - 2.0) < 1e-5
        # tracked outside, moving away: no crossing; moving in: t = 0
        hit, t = po.hit_test(abi.PRIM_SPHERE, (0, 0, 0), 1.0, False, (x, 0, 0), (1, 0, 0), 0.1)
        assert not hit
        hit, t = po.hit_test(abi.PRIM_SPHERE, (0, 0, 0), 1.0, False, (x, 0, 0), (-1, 0, 0), 0.1)
        assert hit and t < 1e-6


def test_chord_through_both_endpoints_outside():
    # segment passes through the circle with both ends outside: entry root
    hit, t = po.hit_test(abi.PRIM_CYLINDER_XY, (0, 0, 0), 1.0, False, (-2, 0.5, 0), (1, 0, 0), 4.0)
    assert hit and abs(t - (2 - math.sqrt(0.75))) < 1e-6
    # misses: tangent line outside
    hit, _ = po.hit_test(abi.PRIM_CYLINDER_XY, (0, 0, 0), 1.0, False, (-2, 1.5, 0), (1, 0, 0), 4.0)
    assert not hit
