"""Host-side pieces of the surrogate (no GPU): parameter box, waveform
features, instance construction (validated by the oracle's substrate and
acquisition checks), training-box test."""
import numpy as np

from axonexa_amd import surrogate as S
from axonexa_amd import waveforms as wf
from oracle import pyoracle as po


def test_sobol_points_fill_the_box_and_map_back():
    u = S.sobol_unit(256, seed=3)
    assert u.shape == (256, S.N_SCALARS) and u.min() >= 0 and u.max() < 1
    v = S.physical(u)
    for k, (_, lo, hi) in enumerate(S.SCALARS):
        assert lo <= v[:, k].min() and v[:, k].max() <= hi
    assert np.allclose(S.unit(v), u)


def test_waveform_samples_follow_the_functor_and_are_balanced():
    proto = S.Protocol()
    u = S.sobol_unit(64, seed=4)
    v = S.physical(u)
    fams = set()
    for q in range(64):
        w = S.waveform_samples(v[q], proto)
        assert w.shape == (S.N_WAVE,) and np.abs(w).max() <= 1.0 + 1e-6
        f = S.functor(v[q], proto)
        fams.add(f.kind)
        # the second lobe mirrors the first: the zeroth moment of the full
        # sampled profile vanishes to sampling accuracy
        t = np.arange(proto.n_steps) * proto.dt
        assert abs(f.profile(t).sum()) * proto.dt < 0.05 * f.T
    assert fams == {"rect", "cos", "sin", "sw"}


def test_instances_are_valid_engine_inputs():
    proto = S.Protocol()
    u = S.sobol_unit(16, seed=5)
    v = S.physical(u)
    geoms = S.geometries(u[:, :3], seed=5)
    for q in range(16):
        sub, acq, run = S.instance(v[q], geoms[q], proto, 100, 1, q)
        assert po.validate(sub, acq) == 0
        assert acq.n_profiles == 1 and acq.n_waveforms == S.N_DIRS
        box, _, radii = geoms[q]
        frac = np.pi * np.sum(radii ** 2) / box ** 2
        assert abs(frac - v[q, 2]) < 1e-9


def test_training_box():
    u = S.sobol_unit(8, seed=6)
    w = np.zeros((8, S.N_WAVE), np.float32)
    inside = S.Surrogate.in_hull(u, w)
    assert inside.all()
    u[3, 5] = 1.01
    w[6, 10] = 1.5
    inside = S.Surrogate.in_hull(u, w)
    assert inside.tolist() == [True, True, True, False, True, True, False, True]


def test_features_layout():
    u = S.sobol_unit(4, seed=7)
    w = np.ones((4, S.N_WAVE), np.float32)
    X = S.features(u, w)
    assert X.shape == (4, S.N_SCALARS + S.N_WAVE)
    assert np.allclose(X[:, :S.N_SCALARS], 2 * u - 1, atol=1e-7)
    assert np.all(X[:, S.N_SCALARS:] == 1)
    assert wf.RectGrad(1.0, 1.0, 1.0).kind == "rect"
