"""GPU runs at the benchmark workloads' real sizes, checked through
size-independent properties (the oracle checks walkers one by one on samples):

  * c1_free (1e5 walkers x 20,000 steps): the exact sinc-product signal of
    fixed-length steps, every gradient within 3 sigma;
  * c1_cylinder (intra-axonal walkers): Gaussian-phase (Neuman / van
    Gelderen) restricted signal within 3 sigma + 2 % of (1 - E);
  * c2_hex without permeability: parallel (z) signal equals the T2-weighted
    mixture of the two compartments' exact free signals, within 3 sigma;
  * c2_hex as benchmarked, a 2048-walker sample over all 250,000 steps:
    walker-by-walker identical to the oracle;
  * c3_gamma as benchmarked (400 gamma-radius cylinders at 65 %, 64 profiles,
    5760 waveforms), a 2048-walker sample over all 80,000 steps: walker by
    walker identical to the oracle;
  * c3_gamma at its full 4e6 walkers x 80,000 steps with 16 extra waveforms
    along z on its PGSE profiles: no walker changes compartment (impermeable)
    and the z signals equal the T2-weighted mixture of the compartments'
    exact free signals within 3 sigma; no step reaches the bounce cap.
"""
import math

import numpy as np
import pytest

from axonexa_amd import abi
from axonexa_amd import substrates as sb
from axonexa_amd import waveforms as wf
from axonexa_amd import workloads as W
from oracle import pyoracle as po
from tests import analytic as an
from tests.helpers import compare_states

pytestmark = pytest.mark.gpu


def _engine():
    from axonexa_amd import engine
    return engine


def test_c1_free_full_size_matches_sinc_product():
    print("[gpu] c1_free 1e5 x 20000", flush=True)
    w = W.c1_free()
    sig = _engine().run_phase(w.sub, w.acq, w.run)
    exact = np.array([wf.free_signal_exact(w.acq, j, 2.0, cell=(20, 20, 20)).real
                      for j in range(w.acq.n_waveforms)])
    z = np.abs(sig.E_re - exact)[1:] / sig.sigma_re[1:]
    print("[gpu] c1_free max |dE|/sigma", z.max(), flush=True)
    assert np.all(z < 3.0)


def test_c1_cylinder_full_size_matches_gpd():
    print("[gpu] c1_cylinder 1e5 x 20000", flush=True)
    w = W.c1_cylinder()
    w.run.init_mode, w.run.init_arg = abi.INIT_INSIDE_PRIM, 0
    sig = _engine().run_phase(w.sub, w.acq, w.run)
    worst = 0.0
    for j, g in enumerate(np.linspace(0.0, 0.3, 16) * 1e-6):
        e = an.gpd_pgse(w.acq.gamma, g, 5.0, 15.0, 2.0, 4.0, "cylinder")
        d = abs(sig.E_re[j] - e)
        worst = max(worst, d)
        assert d <= 3 * sig.sigma_re[j] + 0.02 * (1 - e) + 1e-9, (j, sig.E_re[j], e)
    print("[gpu] c1_cylinder max |E - GPD|", worst, flush=True)
    # also uniform walkers over the whole cell: the intra part is the same signal
    w2 = W.c1_cylinder()
    sig2 = _engine().run_phase(w2.sub, w2.acq, w2.run)
    frac = sig2.n_intra / sig2.n_walkers
    assert abs(frac - math.pi * 16 / 400) < 4 * math.sqrt(0.13 / 1e5)


def test_c2_parallel_signal_is_compartment_mixture():
    print("[gpu] c2_hex impermeable, z gradients", flush=True)
    w = W.c2_hex(n_walkers=1000000, n_steps=50001)
    sub = w.sub
    sub.kappa[:] = 0.0
    run = abi.RunSpec(n_walkers=w.run.n_walkers, seed=1234, boundary_mode=abi.BOUNDARY_SPECULAR,
                      relaxation=1)
    fs = [wf.CosGFunc(g * 1e-6, 4.0, 1.0, 1.0, (0, 0, 1)) for g in (0.0, 0.1, 0.2, 0.4)]
    acq = wf.build_acq(fs, 50001, w.acq.dt, w.acq.gamma)
    sig = _engine().run_phase(sub, acq, run)
    n = acq.n_steps
    fin = sig.n_intra / sig.n_walkers
    win = math.exp(-(n - 1) * acq.dt / 70.0)
    wex = math.exp(-(n - 1) * acq.dt / 50.0)
    for j in range(len(fs)):
        e = (fin * win * wf.free_signal_exact(acq, j, 1.7, cell=(sub.a, sub.b, sub.c)).real
             + (1 - fin) * wex * wf.free_signal_exact(acq, j, 2.0, cell=(sub.a, sub.b, sub.c)).real)
        assert abs(sig.E_re[j] - e) < 3 * sig.sigma_re[j] + 1e-12, (j, sig.E_re[j], e)


def test_c2_sample_full_steps_matches_oracle():
    w = W.c2_hex(n_walkers=2048)
    print("[gpu] c2_hex 2048 walkers x 250000 steps vs oracle", flush=True)
    eng = _engine()
    with eng.Plan(w.sub, w.acq, w.run) as p:
        p.walk().reduce()
        gsig = p.results()
        gst = p.get_state()
    ost = po.walk(w.sub, w.acq, w.run, bin_width=1.5)
    osig = po.reduce(w.sub, w.acq, w.run, ost)
    info = compare_states(gst, ost, w.run.n_walkers, w.acq.n_profiles)
    assert info["bitwise_fraction"] == 1.0
    assert np.allclose(gsig.sum_cos, osig.sum_cos, rtol=1e-12, atol=1e-9)
    assert gsig.transmissions == osig.transmissions and gsig.transmissions > 0
    assert gsig.cap_hits == 0 and osig.cap_hits == 0
    # the same walk as tasks of 8192 steps (31 chunks handed between waves)
    with eng.Plan(w.sub, w.acq, w.run) as p:
        p.set_schedule(8192)
        p.walk()
        assert p.schedule()["last_task_steps"] == 8192
        gst2 = p.get_state()
    assert np.array_equal(gst2, gst)


def test_c2_full_size_tasks_equal_whole_walks():
    # 800,000 walkers fill the resident waves about twice: the automatic
    # schedule walks them in time chunks handed between waves; every word of
    # the state equals the run with one task per walker group
    w = W.c2_hex(n_walkers=800000, n_steps=4001)
    print("[gpu] c2_hex 800000 walkers x 4001 steps: automatic schedule vs whole walks", flush=True)
    eng = _engine()
    states = []
    for sched in (0, -1):
        with eng.Plan(w.sub, w.acq, w.run) as p:
            p.set_schedule(sched)
            p.walk().reduce()
            info = p.schedule()
            print(f"[gpu] schedule {sched}: {info}", flush=True)
            if sched == 0:
                assert info["last_task_steps"] < 4000
            else:
                assert info["last_task_steps"] == 4000
            states.append(p.get_state())
    assert np.array_equal(states[0], states[1])


def test_c3_sample_full_steps_matches_oracle():
    w = W.c3_gamma(n_walkers=2048)
    print(f"[gpu] c3_gamma 2048 walkers x {w.acq.n_steps} steps, P={w.acq.n_profiles} "
          f"M={w.acq.n_waveforms} vs oracle", flush=True)
    eng = _engine()
    with eng.Plan(w.sub, w.acq, w.run) as p:
        p.walk().reduce()
        gsig = p.results()
        gst = p.get_state()
    ost = po.walk(w.sub, w.acq, w.run, bin_width=1.5)
    osig = po.reduce(w.sub, w.acq, w.run, ost)
    info = compare_states(gst, ost, w.run.n_walkers, w.acq.n_profiles)
    print(f"[gpu] c3 sample: {info}, collisions {gsig.collisions}", flush=True)
    assert info["bitwise_fraction"] == 1.0
    for k in ("sum_cos", "sum_sin", "intra_cos", "intra_sin"):
        assert np.allclose(getattr(gsig, k), getattr(osig, k), rtol=1e-12, atol=1e-9), k
    assert gsig.collisions == osig.collisions > 0
    assert gsig.cap_hits == 0 and osig.cap_hits == 0


def test_c3_full_size_parallel_signal_is_compartment_mixture():
    w = W.c3_gamma()
    zs = [wf.RectGrad(0.1 * W.T_PER_M, 4.0 + 1.0 * k, 40.0 - 2.0 * (4.0 + 1.0 * k) - 1.0, (0, 0, 1))
          for k in range(16)]
    acq = wf.build_acq(W.c3_functors() + zs, w.acq.n_steps, w.acq.dt, w.acq.gamma)
    assert acq.n_profiles == 64, "the z waveforms reuse the PGSE profiles"
    print(f"[gpu] c3_gamma full size {w.run.n_walkers} x {acq.n_steps}, M={acq.n_waveforms}", flush=True)
    sig = _engine().run_phase(w.sub, acq, w.run)
    n, N = acq.n_steps, sig.n_walkers
    f = sb.area_fraction(w.sub)
    fin = sig.n_intra / N
    assert abs(fin - f) < 4 * math.sqrt(f * (1 - f) / N), (fin, f)
    # a step reaches the bounce cap only when a walker left within fp32
    # rounding of a surface slides along it at a grazing angle (none of this
    # run's 3.2e11 steps does; tests/test_oracle_physics.py builds such
    # walkers): allow one
    assert sig.transmissions == 0 and sig.cap_hits <= 1 and sig.collisions > 0
    win = math.exp(-(n - 1) * acq.dt / 70.0)
    wex = math.exp(-(n - 1) * acq.dt / 50.0)
    cell = (w.sub.a, w.sub.b, w.sub.c)
    worst = 0.0
    for j in range(acq.n_waveforms - len(zs), acq.n_waveforms):
        e = (fin * win * wf.free_signal_exact(acq, j, 1.7, cell=cell).real
             + (1 - fin) * wex * wf.free_signal_exact(acq, j, 2.0, cell=cell).real)
        z = abs(sig.E_re[j] - e) / sig.sigma_re[j]
        worst = max(worst, z)
        assert z < 3.0, (j, sig.E_re[j], e)
    print(f"[gpu] c3 full size: parallel mixture max |dE|/sigma {worst:.2f}, "
          f"collisions {sig.collisions}", flush=True)
