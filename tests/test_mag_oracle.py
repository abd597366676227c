"""Pins of the magnetization (Bloch) oracle, oracle/axx_oracle.c::orc_walk_mag
(the restatement of updateWalkersMag, src/kernels/kernelMag.cuh:1-141, and
of the integrators in src/bloch/blochdiff.cuh).

What fixes each part:
  * sine/cosine and square root: libm in double (<= 1 ulp; exact);
  * rotation: closed-form precession in a uniform field;
  * field evaluation, positions and sign conventions: with B0 = 0 and a
    gradient along x the magnetization of every walker equals the phase of
    the phase path (an independent code path pinned to closed forms) for the
    same walkers: Mx = cos(phi), My = -sin(phi);
  * RK4 and Dormand-Prince: closed-form precession with T2 / T1 relaxation,
    and their convergence orders (4 and 5) under a rotating RF field, which
    fail if a sample time or a stage coefficient is wrong;
  * substeps: a rotation split into s substeps equals one rotation in a
    static field; functors: their affine form equals the reference's
    piecewise definitions point by point.
"""
import math

import numpy as np
import pytest

from axonexa_amd import abi
from axonexa_amd import fields as fl
from axonexa_amd import substrates as sb
from oracle import pyoracle as po

G = fl.GAMMA


def test_sincos64_and_sqrt_against_libm():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-4, 4, 100000), rng.uniform(-1e5, 1e5, 100000),
                        np.array([0.0, math.pi / 4, math.pi / 2, -math.pi, 1e-300])])
    s, c = po.sincos64(x)
    ulp_s = np.abs(s - np.sin(x)) / np.spacing(np.abs(np.sin(x)) + 1e-300)
    ulp_c = np.abs(c - np.cos(x)) / np.spacing(np.abs(np.cos(x)) + 1e-300)
    assert ulp_s.max() <= 1.0 and ulp_c.max() <= 1.0
    assert np.abs(s - np.sin(x)).max() < 2.3e-16 and np.abs(c - np.cos(x)).max() < 2.3e-16
    y = np.concatenate([rng.uniform(0, 1e6, 50000) ** 2, rng.uniform(0, 1e-12, 50000)])
    assert np.array_equal(po.sqrt_cr(y), np.sqrt(y))


def _still(n=32):
    # walkers that do not move (D = 0)
    return sb.single_sphere(20.0, 5.0, D_in=0.0, D_ex=0.0), abi.RunSpec(n_walkers=n, seed=3)


@pytest.mark.parametrize("integ", [abi.MAG_ROTATE, abi.MAG_RK4, abi.MAG_DOPRI])
def test_uniform_field_precesses_at_the_larmor_rate(integ):
    sub, run = _still()
    B0, dt, n = 2e-3, 1e-4, 400
    f = fl.build_field([fl.RECT_GRE(B0, (0, 0, 0), 1.0, 0.1)], n, dt, integrator=integ,
                       substeps=np.full(n, 4, np.int32), m0=(0.6, 0.0, 0.8))
    M = abi.mag_from_state(po.walk_mag(sub, f, run), run.n_walkers, 1)[:, 0]
    ang = G * B0 * n * dt
    want = np.array([0.6 * math.cos(ang), -0.6 * math.sin(ang), 0.8])
    tol = {abi.MAG_ROTATE: 1e-13, abi.MAG_RK4: 1e-6, abi.MAG_DOPRI: 1e-9}[integ]
    assert np.abs(M - want).max() < tol, np.abs(M - want).max()


def test_substeps_of_a_static_field_compose():
    sub, run = _still(8)
    B = fl.RECT_GRE(1e-3, (0, 0, 0), 1.0, 0.1)
    one = fl.build_field([B], 50, 1e-3, m0=(1.0, 0.0, 0.0))
    many = fl.build_field([B], 50, 1e-3, substeps=np.full(50, 7, np.int32), m0=(1.0, 0.0, 0.0))
    a = abi.mag_from_state(po.walk_mag(sub, one, run), 8, 1)
    b = abi.mag_from_state(po.walk_mag(sub, many, run), 8, 1)
    assert np.abs(a - b).max() < 1e-13


@pytest.mark.parametrize("integ", [abi.MAG_RK4, abi.MAG_DOPRI])
def test_relaxation_closed_form(integ):
    sub = sb.single_sphere(20.0, 5.0, D_in=0.0, D_ex=0.0, T2_in=3.0, T2_ex=3.0)
    run = abi.RunSpec(n_walkers=16, seed=4, relaxation=1, init_mode=abi.INIT_INSIDE_PRIM)
    B0, dt, n, T1 = 1e-5, 1e-2, 300, 5.0
    f = fl.build_field([fl.RECT_GRE(B0, (0, 0, 0), 1.0, 0.1)], n, dt, integrator=integ,
                       m0=(1.0, 0.0, 0.0), T1=np.array([T1, T1]))
    M = abi.mag_from_state(po.walk_mag(sub, f, run), run.n_walkers, 1)[:, 0]
    t = n * dt
    ang = G * B0 * t
    want = np.array([math.exp(-t / 3.0) * math.cos(ang), -math.exp(-t / 3.0) * math.sin(ang),
                     1.0 - math.exp(-t / T1)])
    assert np.abs(M - want).max() < 1e-6, np.abs(M - want).max()


# RF on for the whole walk (39 steps of 1 us inside a 40 us pulse), so the
# field is smooth in time and the orders are not spoilt by the pulse edge
def _rf_error(integ, s, ref):
    sub, run = _still(4)
    uc = fl.UCN_lowfield(5e-4, 0.0, 0.04, 1e-3, 1e-3)
    f = fl.build_field([uc], 39, 1e-3, integrator=integ, substeps=np.full(39, s, np.int32), m0=(0, 0, 1.0))
    M = abi.mag_from_state(po.walk_mag(sub, f, run), 4, 1)[0, 0]
    return np.abs(M - ref).max()


def test_convergence_orders_under_a_rotating_rf_field():
    sub, run = _still(4)
    uc = fl.UCN_lowfield(5e-4, 0.0, 0.04, 1e-3, 1e-3)
    fine = fl.build_field([uc], 39, 1e-3, integrator=abi.MAG_DOPRI, substeps=np.full(39, 64, np.int32),
                          m0=(0, 0, 1.0))
    ref = abi.mag_from_state(po.walk_mag(sub, fine, run), 4, 1)[0, 0]
    # 39/40 of a pi/2 pulse tips M from z nearly into the transverse plane
    assert abs(ref[2] - math.cos(math.pi / 2 * 39 / 40)) < 1e-3 and np.hypot(ref[0], ref[1]) > 0.99
    e1, e2 = _rf_error(abi.MAG_RK4, 2, ref), _rf_error(abi.MAG_RK4, 4, ref)
    assert e1 / e2 > 12.0, (e1, e2)          # order 4: 16x
    d1, d2 = _rf_error(abi.MAG_DOPRI, 1, ref), _rf_error(abi.MAG_DOPRI, 2, ref)
    assert d1 / d2 > 24.0, (d1, d2)          # order 5: 32x
    # piecewise-constant rotation is first order in time for a time-varying field
    r1, r2 = _rf_error(abi.MAG_ROTATE, 8, ref), _rf_error(abi.MAG_ROTATE, 16, ref)
    assert 1.5 < r1 / r2 < 2.6, (r1, r2)


def test_zero_static_field_equals_the_phase_path():
    """B = (0, 0, G x g(t)): every walker's Mx + i My is exp(-i phi) with phi
    the phase path's phase of the same walker (same walk, same numbers)."""
    dt = 2.0 ** -10
    d, s = 16 * dt, 8 * dt          # lobes [0, 16 dt) and [24 dt, 40 dt): balanced samples
    n = 48
    sub = sb.hex_lattice(0.03, 0.07, 2, 1, D_in=2e-4, D_ex=4e-4, kappa=0.0)
    run = abi.RunSpec(n_walkers=3000, seed=9)
    Gx = 2e4
    f = fl.build_field([fl.RECT_GRE(0.0, (Gx, 0, 0), d, s)], n, dt, gamma=1.0, m0=(1.0, 0.0, 0.0))
    M = abi.mag_from_state(po.walk_mag(sub, f, run), run.n_walkers, 1)[:, 0]
    t = np.arange(n) * dt
    prof = np.where(t < d, 1.0, np.where((t >= s + d) & (t < 2 * d + s), -1.0, 0.0)).astype(np.float32)
    assert prof.sum() == 0.0
    acq = abi.AcqSpec(n_steps=n, dt=dt, gamma=1.0, profiles=prof, wf_profile=[0], wf_amplitude=[Gx],
                      wf_dir=[[1.0, 0.0, 0.0]])
    st = po.walk(sub, acq, run)
    v = abi.state_view(st, run.n_walkers, 1)
    phi = dt * Gx * v["mom"][:, 0, 0].astype(np.float64)
    assert np.abs(phi).max() > 0.5                     # the gradient does dephase
    assert np.abs(M[:, 0] - np.cos(phi)).max() < 2e-5
    assert np.abs(M[:, 1] + np.sin(phi)).max() < 2e-5
    assert np.abs(M[:, 2]).max() < 1e-12
    # same walkers: positions of both walks are identical
    sm = po.walk_mag(sub, f, run).reshape(abi.mag_state_fields(1), run.n_walkers)
    assert np.array_equal(sm[:abi.F_MOM], st.reshape(-1, run.n_walkers)[:abi.F_MOM])


def test_field_groups_and_chunks_are_consistent():
    sub = sb.single_sphere(0.01, 0.0025, D_in=2.5e-6, D_ex=2.5e-6)
    run = abi.RunSpec(n_walkers=500, seed=2, init_mode=abi.INIT_INSIDE_PRIM)
    fs = [fl.QSC_GRE(5e-4, g * 4e-4, 0.02, 0.01) for g in range(11)]
    n, dt = 500, 1e-4
    f = fl.build_field(fs, n, dt, origin=(0.005, 0.005, 0.005))
    full = po.run_mag(sub, f, run)
    # group 1 holds fields 8..10; the same fields alone give the same walkers
    f2 = fl.build_field(fs[8:], n, dt, origin=(0.005, 0.005, 0.005))
    alone = po.run_mag(sub, f2, run)
    assert np.array_equal(full.mx[8:], alone.mx) and np.array_equal(full.mz[8:], alone.mz)
    # a walk in two chunks equals one walk
    st = po.walk_mag(sub, f, run, step_end=123)
    st = po.walk_mag(sub, f, run, step_begin=123, state=st)
    assert np.array_equal(st, po.walk_mag(sub, f, run))
    # G = 0: no dephasing beyond B0 precession, |Mxy| = N
    assert abs(math.hypot(full.mx[0], full.my[0]) - run.n_walkers) < 1e-9 * run.n_walkers
    assert full.transverse[10] < full.transverse[0]


def test_functors_match_piecewise_definitions():
    rng = np.random.default_rng(5)
    r = rng.normal(size=3)
    cases = [
        (fl.QSC_GRE(0.1, 2.0, 3.0, 1.0),
         lambda t: (0, 2 * r[2], 0.1 + 2 * r[1]) if t < 3 else
         ((0, -2 * r[2], 0.1 - 2 * r[1]) if 4 <= t < 7 else (0, 0, 0.1))),
        (fl.UCN_lowfield(0.2, 3.0, 0.5, 0.1, 0.2, gamma=2.0),
         lambda t: ((math.pi / 2) * math.cos(0.4 * t) - 1.5 * r[0], -(math.pi / 2) * math.sin(0.4 * t) - 1.5 * r[1],
                    3 * r[2] + 0.2) if t < 0.5 else (-1.5 * r[0], -1.5 * r[1], 3 * r[2] + 0.2)),
        (fl.COS_GRE(0.3, (1.0, 2.0, 0.5), 2.0, 1.0, 2),
         lambda t: (0, 0, 0.3 + (r[0] + 2 * r[1] + 0.5 * r[2]) * math.cos(2 * math.pi * t / 1.0)) if t < 2 else
         ((0, 0, 0.3 - (r[0] + 2 * r[1] + 0.5 * r[2]) * math.cos(2 * math.pi * (t - 3) / 1.0)) if 3 <= t < 5
          else (0, 0, 0.3))),
        (fl.QSC_PGSE(0.1, 2.0, 1.0, 2.0, 0.25, 0.01, 0.1, 4.0, gamma=3.0),
         lambda t: (0, 2 * r[2], 0.1 + 2 * r[1]) if t < 1 else
         ((math.pi / 0.75 * math.cos(0.3 * t), -math.pi / 0.75 * math.sin(0.3 * t), 0.1) if 2 <= t < 2.25 else
          ((0, 2 * r[2], 0.1 + 2 * r[1]) if 3 <= t < 4 else (0, 0, 0.1)))),
    ]
    for fn, ref in cases:
        for t in np.linspace(0, 8, 161):
            assert np.allclose(fn(r, t), ref(t), rtol=1e-13, atol=1e-13), (type(fn).__name__, t)
    tr = fl.QSC_GRE_TRAP(0.0, 2.0, 1.0, 0.5, 0.25)
    # ramp up multiplies by (G / ramp) t: at t = ramp the gradient is G^2
    assert np.isclose(tr(r, 0.25 - 1e-12, )[2], 2.0 * 2.0 * r[1], rtol=1e-9)
    assert np.isclose(tr(r, 0.5)[2], 2.0 * r[1])
    pg = fl.QSC_PGSE(0.1, 2.0, 1.0, 2.0, 0.25, 0.01, 0.1, 4.0)
    # dtmax 0.1 outside the pulse, dtmin 0.01 in it (it starts at TE / 2 = 2.0)
    assert np.array_equal(fl.auto_substeps([pg], 5, 0.5), [5, 5, 5, 5, 50])
