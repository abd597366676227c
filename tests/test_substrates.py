"""Substrate builders (src/substrate/RPSinitializer.h, examples/hexPhase.cu)
and their validation: non-overlap under the minimum image, exact area
fractions, the permeability for a target transmission probability."""
import math

import numpy as np

from axonexa_amd import abi
from axonexa_amd import substrates as sb
from axonexa_amd import waveforms as wf
from oracle import pyoracle as po


def _min_image_min_gap(sub):
    c = sub.centers[:, :2]
    L = np.array([sub.a, sub.b])
    d = c[:, None, :] - c[None, :, :]
    d -= L * np.round(d / L)
    dist = np.hypot(d[..., 0], d[..., 1])
    np.fill_diagonal(dist, np.inf)
    return np.min(dist - (sub.radii[:, None] + sub.radii[None, :]))


def test_hex_lattice_geometry():
    s = sb.hex_lattice(2.5, 6.0, 20, 25, D_in=1.7, D_ex=2.0)
    assert s.n_prims == 1000
    assert abs(s.a - 120.0) < 1e-9 and abs(s.b - 25 * 6 * math.sqrt(3)) < 1e-9
    assert abs(sb.area_fraction(s) - sb.hex_intra_fraction(2.5, 6.0)) < 1e-12
    assert abs(_min_image_min_gap(s) - 1.0) < 1e-9


def test_gamma_packing_reaches_target_fraction_without_overlap():
    from scipy import stats
    s = sb.gamma_packing(400, 5.0, 0.4, 0.65, D_in=1.7, D_ex=2.0, seed=4, rmin=0.25)
    assert s.n_prims == 400
    assert abs(sb.area_fraction(s) - 0.65) < 1e-9
    assert _min_image_min_gap(s) >= -1e-9
    # the benchmark packing keeps its surfaces a step length apart
    from axonexa_amd import workloads as W
    c3 = W.c3_gamma(n_walkers=16, n_steps=2).sub
    assert abs(sb.area_fraction(c3) - 0.65) < 1e-9
    assert _min_image_min_gap(c3) >= W.C3_GAP - 1e-9
    # radii follow Gamma(k=5, theta=0.4) (truncation at 0.25 um is negligible)
    assert stats.kstest(s.radii, "gamma", args=(5.0, 0.0, 0.4)).pvalue > 1e-3


def test_validation_rejects_overlaps_and_large_steps():
    acq = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0)], 10, 1e-3, 2.675e5)
    bad = abi.SubstrateSpec(a=10, b=10, c=10, D_lat=2.0, types=[0, 0],
                            centers=[[2, 2, 0], [4, 2, 0]], radii=[1.5, 1.5], D=2.0, T2=np.inf)
    assert po.validate(bad, acq) == -5
    wrap = abi.SubstrateSpec(a=10, b=10, c=10, D_lat=2.0, types=[0, 0],
                             centers=[[0.5, 5, 0], [9.5, 5, 0]], radii=[0.6, 0.6], D=2.0, T2=np.inf)
    assert po.validate(wrap, acq) == -5        # overlap across the periodic edge (defect D8)
    ok = sb.single_cylinder(10.0, 3.0, D_in=2.0, D_ex=2.0)
    assert po.validate(ok, acq) == 0
    big_step = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0)], 10, 10.0, 2.675e5)
    assert po.validate(ok, big_step) == -5


def test_kappa_for_probability():
    dt, D = 2e-4, 1.7
    k = sb.kappa_for_probability(1e-3, D, dt)
    assert abs(4 * k / math.sqrt(6 * D / dt) - 1e-3) < 1e-15
