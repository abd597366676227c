"""Philox4x32-10, the uniform conversion and the step direction of the oracle.

Pins: the Random123 known-answer vectors (Salmon et al., SC'11), rocRAND's
header-only engine (fixture tests/golden/philox_rocrand.txt, generated by
tests/golden/gen_philox_rocrand.cpp), closed-form properties of the uniform
and of sin/cos of 2*pi*u.  Replaces curand XORWOW + acos/sin/cos of
src/kernels/kernelLattice.cuh:52-57.
"""
import math
import os

import numpy as np
import pytest
from scipy import stats

from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))

# Random123 kat_vectors, philox4x32 10 rounds: ctr[4], key[2] -> out[4]
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_random123_kat(ctr, key, out):
    assert list(po.philox(ctr, key)) == out


def test_philox_matches_rocrand_fixture():
    rows = [l.split() for l in open(os.path.join(HERE, "golden", "philox_rocrand.txt")) if l.strip()]
    assert len(rows) >= 6
    for r in rows:
        v = [int(x, 16) for x in r]
        assert list(po.philox(v[:4], v[4:6])) == v[6:], r


def test_uniform_exact_and_open_interval():
    # u = ((w >> 8) | 1) * 2^-24: odd multiples of 2^-24 in (0, 1)
    for w in (0, 1, 255, 256, 0x7FFFFFFF, 0xFFFFFFFF, 0x80000000):
        u = po.unif(w)
        k = ((w >> 8) | 1)
        assert u == k * 2.0 ** -24
        assert 0.0 < u < 1.0
    assert po.unif(0) == 2.0 ** -24
    assert po.unif(0xFFFFFFFF) == 1.0 - 2.0 ** -24


def test_sincos_turn_whole_grid_against_libm():
    """sin, cos of 2 pi u(w) from the azimuth table, at every one of the
    2^23 values u = k 2^-24 (k odd) against libm in fp64: the table's fp32
    rounding, the turn by d (|d| <= pi/256, truncation < 1e-9) and the
    angle-sum products stay within 2.5e-7; the unit norm within 5e-7."""
    k = np.arange(1, 1 << 24, 2, dtype=np.uint64)
    s, c = po.sincos_turn_n((k << 8).astype(np.uint32))
    phi = 2.0 * np.pi * k.astype(np.float64) * 2.0 ** -24
    es = np.abs(s.astype(np.float64) - np.sin(phi))
    ec = np.abs(c.astype(np.float64) - np.cos(phi))
    assert max(es.max(), ec.max()) < 2.5e-7, (es.max(), ec.max())
    n2 = s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2
    assert np.abs(n2 - 1.0).max() < 5e-7
    # the low 8 bits of the word are not used
    s2, c2 = po.sincos_turn_n((k << 8 | 0x7F).astype(np.uint32)[::4096])
    assert np.array_equal(s2, s[::4096]) and np.array_equal(c2, c[::4096])


def test_cylinder_walk_directions_keep_an_xy_component():
    """The cylinder-only crossing test takes its decision without the term
    A > 0 (A = ux^2 + uy^2, the quadratic's leading coefficient): every step
    generator's direction has A well above zero, over the whole grids of its
    words, and a reflection about a normal with nz = 0 keeps A to rounding."""
    # isotropic: cos(theta) on its whole grid of 2^23 values, sin theta the
    # correctly rounded square root of fmaf(-ct, ct, 1) (exact in fp64 here)
    m = np.arange(1 << 23, dtype=np.float64)
    ct = m * 2.0 ** -22 + (2.0 ** -23 - 1.0)
    sth = np.sqrt((1.0 - ct * ct).astype(np.float32)).astype(np.float64)
    k = np.arange(1, 1 << 24, 2, dtype=np.uint64)
    s, c = po.sincos_turn_n((k << 8).astype(np.uint32))
    n2 = s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2
    a_iso = sth.min() ** 2 * n2.min() * (1.0 - 2.0 ** -20)   # the products' roundings
    assert a_iso > 2.0 ** -23, a_iso
    # axis signs: two components of +-0.57735026
    assert 2 * float(np.float32(0.57735026)) ** 2 > 0.66
    # Gaussian: A = r1^2 (c1^2 + s1^2) / |n|^2 with |n|^2 <= r1^2 + r2^2
    y = po.neg2log_unif((k << 8).astype(np.uint32)).astype(np.float64)
    a_gauss = y.min() * n2.min() / (2.0 * y.max()) * (1.0 - 2.0 ** -18)
    assert a_gauss > 1e-9, a_gauss
    # reflection u' = u - 2 (u.n) n about n = (p - c) / R, nz = 0, in fp32,
    # from directions at the isotropic minimum of A up to A = 1
    rng = np.random.default_rng(11)
    n = 200000
    f32 = np.float32
    R = rng.uniform(0.2, 8.0, n).astype(f32)
    ang = rng.uniform(0.0, 2 * np.pi, n)
    cx, cy = rng.uniform(-50, 50, n).astype(f32), rng.uniform(-50, 50, n).astype(f32)
    px = (cx + R * np.cos(ang)).astype(f32)
    py = (cy + R * np.sin(ang)).astype(f32)
    inv = (f32(1.0) / R).astype(f32)
    nx, ny = ((px - cx) * inv).astype(f32), ((py - cy) * inv).astype(f32)
    st = np.where(rng.random(n) < 0.5, sth.min(), rng.uniform(sth.min(), 1.0, n)).astype(f32)
    phi = rng.uniform(0.0, 2 * np.pi, n)
    ux, uy = (st * np.cos(phi)).astype(f32), (st * np.sin(phi)).astype(f32)
    a0 = (uy * uy + ux * ux).astype(np.float64)
    kd = (ny * uy + nx * ux).astype(f32)
    k2 = (kd + kd).astype(f32)
    ux2, uy2 = (ux - k2 * nx).astype(f32), (uy - k2 * ny).astype(f32)
    a1 = (uy2 * uy2 + ux2 * ux2).astype(np.float64)
    assert np.abs(a1 / a0 - 1.0).max() < 1e-4 and a1.min() > 2.0 ** -23


def test_azimuth_table_nodes():
    """Entry j is the quarter-turn polynomial at 2 pi (j + 1/2)/256, within
    2e-7 of libm; a table centred on j/256 instead (a plausible off-by-half)
    would leave the turned values 0.012 off."""
    c, s = po.azimuth_table()
    j = np.arange(256)
    ang = 2.0 * np.pi * (j + 0.5) / 256.0
    assert np.abs(c - np.cos(ang)).max() < 2e-7 and np.abs(s - np.sin(ang)).max() < 2e-7
    for jj in (0, 1, 63, 64, 127, 128, 200, 255):
        ps, pc = po.sincos_u24((2 * jj + 1) << 15)
        assert (np.float32(pc), np.float32(ps)) == (c[jj], s[jj])
    # the polynomial itself over the quarter-turn boundaries
    for v in (0, 1, (1 << 22) - 1, 1 << 22, 1 << 23, 3 << 22, (1 << 24) - 1):
        ps, pc = po.sincos_u24(v)
        a = 2.0 * math.pi * v * 2.0 ** -24
        assert abs(ps - math.sin(a)) < 2e-7 and abs(pc - math.cos(a)) < 2e-7
    # off-by-half planted in a float64 restatement (entries at j/256, turn
    # still measured from the bin centre): the bound above rejects it
    kk = np.arange(1, 1 << 24, 2 * 997, dtype=np.int64)
    d = 2.0 * np.pi * ((kk & 0xFFFF) - 32768) * 2.0 ** -24
    a_wrong = 2.0 * np.pi * (kk >> 16) / 256.0 + d
    assert np.abs(np.cos(a_wrong) - np.cos(2.0 * np.pi * kk * 2.0 ** -24)).max() > 1e-2


def test_direction_distribution_is_isotropic():
    # cos(theta) = 2u - 1 must be uniform on (-1, 1) and phi uniform on (0, 2 pi):
    # draw the step words exactly as the walk does (pair counter, words x/y)
    n = 20000
    cts, phis = [], []
    for wid in range(n // 2):
        o = po.philox([wid, 0, 0, 0], [1234, 0])
        for wa, wb in ((o[0], o[1]), (o[2], o[3])):
            cts.append(2.0 * po.unif(int(wb)) - 1.0)
            s, c = po.sincos_turn(int(wa))
            phis.append(math.atan2(s, c) % (2 * math.pi))
    cts, phis = np.array(cts), np.array(phis)
    assert stats.kstest(cts, "uniform", args=(-1, 2)).pvalue > 1e-3
    assert stats.kstest(phis, "uniform", args=(0, 2 * math.pi)).pvalue > 1e-3
    # mean of the unit step is 0 and E[ux^2] = 1/3
    st = np.sqrt(np.maximum(1 - cts ** 2, 0))
    ux = st * np.cos(phis)
    assert abs(ux.mean()) < 4 / math.sqrt(n)
    assert abs((ux ** 2).mean() - 1 / 3) < 4 * math.sqrt(4 / 45 / n)


def test_neg2log_unif_whole_grid_within_3_ulp():
    """-2 ln u of the Gaussian steps against libm (fp64) at every one of the
    2^23 values u = k 2^-24, k odd: at most 3 ulp of the fp32 result (the
    atanh series to s^9 truncates below 4e-10, far under one ulp)."""
    k = np.arange(1, 1 << 24, 2, dtype=np.uint64)
    y = po.neg2log_unif((k << 8).astype(np.uint32)).astype(np.float64)
    ref = -2.0 * np.log(k.astype(np.float64) * 2.0 ** -24)
    ulp = np.abs(y - ref) / np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert ulp.max() < 3.0, ulp.max()
    assert y.min() > 0.0 and abs(y.max() - ref.max()) < 1e-5
    # the low byte of the word is ignored
    assert np.array_equal(po.neg2log_unif(np.array([0x12345600, 0x123456FF], np.uint32))[0:1],
                          po.neg2log_unif(np.array([0x123456FF], np.uint32)))
