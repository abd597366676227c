"""The learned surrogate end to end on the GPU: data from the batched
engine, training with the project's MFMA kernels, held-out error against
Monte Carlo, the Monte Carlo fallback outside the training box, and what the
network gains over the engine (queries/s at matched noise).

Sizes are reduced (16,384 training queries of a 10 ms protocol, 2,000
walkers each) so the test runs in well under a minute; the held-out truth
uses 40,000 walkers (noise sigma = 1/sqrt(2N) = 0.0035, mean absolute noise
0.0028), so the 1 % gate on the mean absolute error is dominated by the
network's own error.  The full-size figures come from
``bench.py --workload c5_surrogate``.
"""
import math

import numpy as np
import pytest

from axonexa_amd import surrogate as S

pytestmark = pytest.mark.gpu


def _data(proto, G, variants, seed, walkers, first_instance=0):
    u_geo = S.sobol_unit(G, seed=seed, dims=3)
    u = np.concatenate([np.repeat(u_geo, variants, axis=0),
                        S.sobol_unit(G * variants, seed=seed + 1, dims=S.N_SCALARS - 3)], axis=1)
    v = S.physical(u)
    geoms = S.geometries(u_geo, seed=seed, processes=8)
    Y = S.simulate(v, geoms, np.repeat(np.arange(G), variants), proto, walkers, seed=1234,
                   first_instance=first_instance)
    X = S.features(u, np.stack([S.waveform_samples(v[q], proto) for q in range(len(u))]))
    return u, X, Y


def test_surrogate_learns_the_engine_and_falls_back_outside_its_box():
    proto = S.Protocol(TE=10.0, dt=5e-3)
    u, X, Y = _data(proto, 4096, 4, seed=31, walkers=2000)
    print(f"[surrogate] {len(X)} training queries simulated", flush=True)
    uh, Xh, Yh = _data(proto, 128, 1, seed=41, walkers=40000, first_instance=10 ** 6)
    sur = S.Surrogate(proto, max_batch=4096)
    try:
        sur.init(0)
        trace = sur.fit(X, Y.astype(np.float32), steps=3000, batch=4096, lr=1e-3)
        P = sur.predict_features(Xh)
        mae = float(np.abs(P - Yh).mean())
        base = float(np.abs(Y.mean(axis=0) - Yh).mean())      # predict the mean signal
        print(f"[surrogate] loss {trace[:10].mean():.4f} -> {trace[-10:].mean():.5f}; held-out MAE "
              f"{mae:.4f} vs mean-predictor {base:.4f}", flush=True)
        assert mae < 0.01 and mae < 0.25 * base
        # a query outside the training box is answered by Monte Carlo, exactly
        uo = uh[:3].copy()
        uo[1, 2] = 1.1                                          # packing fraction beyond the box
        E, used = sur.predict(uo, mc_walkers=2000, seed=5)
        assert used.tolist() == [False, True, False]
        vo = S.physical(uo[1:2])
        geo = S.geometries(uo[1:2, :3], seed=5)
        ref = S.simulate(vo, geo, np.arange(1), proto, 2000, seed=5)
        assert np.array_equal(E[1], ref[0])
        assert np.allclose(E[[0, 2]], sur.predict_features(
            S.features(uo[[0, 2]], np.stack([S.waveform_samples(v, proto) for v in S.physical(uo[[0, 2]])]))))
    finally:
        sur.close()


def test_surrogate_is_over_1000x_faster_than_monte_carlo_at_matched_noise():
    from axonexa_amd import engine

    proto = S.Protocol()                       # the production protocol: 30 ms, 12,001 steps
    sur = S.Surrogate(proto, max_batch=65536)
    try:
        sur.init(1)
        X = np.random.default_rng(2).uniform(-1, 1, size=(65536, S.N_SCALARS + S.N_WAVE)).astype(np.float32)
        sur.net.forward(X)
        ms = float(np.median([sur.net.time_forward(65536) for _ in range(3)]))
        sur_qps = 65536 / (ms * 1e-3)
    finally:
        sur.close()
    # Monte Carlo with the walker count whose noise is 1 % (sigma = 1/sqrt(2N))
    walkers = int(math.ceil(1.0 / (2 * 0.01 ** 2)))
    u = S.sobol_unit(16, seed=51)
    v = S.physical(u)
    geoms = S.geometries(u[:, :3], seed=51)
    insts = [S.instance(v[q], geoms[q], proto, walkers, 3, q) for q in range(16)]
    with engine.Batch(insts) as b:
        b.run().sync()
        mc_ms = float(np.median([b.time() for _ in range(2)]))
    mc_qps = 16 / (mc_ms * 1e-3)
    print(f"[surrogate] network {sur_qps:.3e} queries/s, Monte Carlo {mc_qps:.3e} queries/s, "
          f"ratio {sur_qps / mc_qps:.0f}", flush=True)
    assert sur_qps > 1000 * mc_qps
