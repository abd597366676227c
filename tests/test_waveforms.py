"""Gradient waveforms: closed-form b-values (src/functors/gfunctors.cuh getB)
against quadrature of b = int (gamma int_0^t G)^2 dt (the definition used by
src/external/bvalue.h:3-30), the discrete b of the sampled tables, balance of
the sampled profiles, and profile sharing between functors."""
import math

import numpy as np
import pytest

from axonexa_amd import waveforms as wf

G = 0.3e-6          # T/um
GAMMA = 2.675e5


@pytest.mark.parametrize("f", [
    wf.RectGrad(G, 5.0, 10.0),
    wf.CosGFunc(G, 20.0, 2.0, 3),
    wf.SinGFunc(G, 20.0, 2.0, 3),
    wf.SWOGSEFunc(G, 20.0, 2.0, 3),
])
def test_closed_form_b_matches_quadrature(f):
    bq = wf.b_quadrature(f, GAMMA, n=400000)
    bc = f.getB(GAMMA)
    assert abs(bc - bq) <= 2e-3 * bq, (type(f).__name__, bc, bq)


def test_apodised_cosine_closed_form_does_not_describe_its_waveform():
    # gfunctors.cuh:220-228 apodises with sin(4 pi t / period) on the first and
    # last quarter periods, which leaves each lobe a net moment of -period/pi
    # (times G); its getB (:196-199) assumes a balanced lobe.  The engine
    # integrates the waveform as written, so its b is the quadrature value.
    f = wf.ApodCosGFunc(G, 20.0, 2.0, 4)
    t = np.linspace(0.0, 20.0, 400001)
    m0 = np.trapezoid(f.profile(t), t)
    assert abs(m0 - (-(20.0 / 4) / math.pi)) < 1e-3
    bq = wf.b_quadrature(f, GAMMA, n=400000)
    assert f.getB(GAMMA) < 0.5 * bq
    dt = 1e-3
    n = int(round(42.0 / dt)) + 1
    acq = wf.build_acq([f], n, dt, GAMMA)
    assert abs(wf.b_discrete(acq.profiles[:, 0], f.G, dt, GAMMA) - bq) <= 2e-3 * bq


def test_discrete_b_converges_to_continuum():
    f = wf.CosGFunc(G, 20.0, 5.0, 2.5)  # real-valued number of periods
    dt = 2e-3
    n = int(round((2 * 20.0 + 5.0) / dt)) + 1
    acq = wf.build_acq([f], n, dt, GAMMA)
    bd = wf.b_discrete(acq.profiles[:, 0], f.G, dt, GAMMA)
    bq = wf.b_quadrature(f, GAMMA, n=400000)
    assert abs(bd - bq) <= 2e-3 * bq


def test_pgse_table_is_balanced_and_exact():
    dt = 1e-3
    acq = wf.build_acq([wf.RectGrad(G, 5.0, 10.0)], 20000, dt, GAMMA)
    p = acq.profiles[:, 0].astype(np.float64)
    assert set(np.unique(p)) <= {-1.0, 0.0, 1.0}
    assert (p == 1).sum() == (p == -1).sum() == 5000
    assert p.sum() == 0.0


def test_square_wave_is_finite_and_signed():
    # the reference's pow(-1, 2tv) is NaN for non-integer 2tv (gfunctors.cuh:281-283)
    f = wf.SWOGSEFunc(G, 10.0, 1.0, 2.5)
    t = np.linspace(0, 21, 10001)
    p = f.profile(t)
    assert np.all(np.isfinite(p))
    assert set(np.unique(p)) <= {-1.0, 0.0, 1.0}


def test_profiles_are_shared_by_shape():
    fs = [wf.CosGFunc(g, 20.0, 5.0, 2.0, ax) for g in (1e-7, 2e-7) for ax in ((1, 0, 0), (0, 0, 1))]
    fs.append(wf.CosGFunc(1e-7, 20.0, 5.0, 3.0))
    acq = wf.build_acq(fs, 1000, 0.05, GAMMA)
    assert acq.n_profiles == 2
    assert list(acq.wf_profile) == [0, 0, 0, 0, 1]
    assert np.allclose(acq.wf_amplitude, [1e-7, 1e-7, 2e-7, 2e-7, 1e-7])
