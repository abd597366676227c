"""Signal post-processing (axonexa_amd/analysis.py): the ADC fit recovers D
exactly from exact exponentials and, within Monte Carlo error, from the
oracle's free-diffusion signals; replica statistics; the driver table."""
import math

import numpy as np

from axonexa_amd import abi
from axonexa_amd import analysis as A
from axonexa_amd import substrates as sb
from axonexa_amd import waveforms as wf
from oracle import pyoracle as po

GAMMA = 2.675e5


def test_adc_of_exact_exponentials():
    b = np.linspace(0.0, 2.0, 9)
    assert abs(A.adc_fit(b, 0.7 * np.exp(-1.37 * b)) - 1.37) < 1e-12


def test_adc_of_free_diffusion_is_D():
    # PGSE in a free periodic box: ln E = -b D (continuum), the discrete table
    # gives it to O(dt / delta); fit against the sampled-table b-values
    gs = np.linspace(0.02, 0.12, 6) * 1e-6
    fs = [wf.RectGrad(g, 5.0, 10.0) for g in gs]
    acq = wf.build_acq(fs, 20000, 1e-3, GAMMA)
    sub = sb.empty_box(20.0, 2.0)
    run = abi.RunSpec(n_walkers=20000, seed=3)
    sig, _ = po.run_phase(sub, acq, run, bin_width=1.0)
    b = np.array([wf.b_discrete(acq.profiles[:, acq.wf_profile[j]], acq.wf_amplitude[j], acq.dt, GAMMA)
                  for j in range(acq.n_waveforms)])
    E = A.signal(sig.sum_cos, sig.sum_sin, run.n_walkers)
    assert np.allclose(E, sig.E, rtol=1e-12)
    D = A.adc_fit(b, E)
    assert abs(D - 2.0) < 0.03, D


def test_repeat_stats_and_table():
    r = A.RepeatStats()
    for v in (1.0, 2.0, 3.0, 4.0):
        r.add(v)
    assert r.points == 4 and r.mean == 2.5
    assert abs(r.error() - math.sqrt(1.25)) < 1e-12
    t = A.table([0.05], [1e-7], [10.0], [-2.0]).split()
    assert [float(x) for x in t] == [0.05, 1e-7, 10.0, -2.0]
