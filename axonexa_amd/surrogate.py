"""Learned surrogate of the Monte Carlo engine, with the engine as its
fallback and judge.

Query: 12 scalars (gamma radius mean and shape, packing fraction, membrane
transmission probability, D_in, D_ex, T2_in, T2_ex, gradient amplitude,
lobe fraction, OGSE periods, waveform family) plus the 256-sample unit
gradient waveform.  Answer: the normalised signal (T2 weighting included) of
30 gradient directions (a Fibonacci set) for cylinders packed with
Gamma(k, r_mean / k) radii at that fraction.

Training data come only from this simulator: Sobol points of the 12
scalars, a gamma packing per geometry, one batched engine run per chunk of
instances (``engine.Batch``).  The network is ``engine.MLP``
(268 -> 1024 x 6 GELU -> 30, BF16 on MFMA, trained on the GPU by the
project's own kernels).  Queries outside the training box of the scalars (or
with waveform samples outside [-1, 1]) are answered by the Monte Carlo
engine instead.

No counterpart in the reference; its nearest relatives are the gamma-packing
drivers (examples/gammadist.cu) whose signals the surrogate reproduces.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import abi
from . import substrates as sb
from . import waveforms as wf

GAMMA = 2.675e5
T_PER_M = 1e-6

# (name, low, high): the training box of the 12 scalars
SCALARS: List[Tuple[str, float, float]] = [
    ("r_mean", 0.8, 3.0),          # um, gamma mean
    ("k_shape", 2.0, 10.0),        # gamma shape
    ("fraction", 0.35, 0.62),      # intra-cylinder area fraction
    ("p_transmit", 0.0, 0.02),     # transmission probability per collision (from inside)
    ("D_in", 0.5, 2.5),            # um^2 / ms
    ("D_ex", 1.0, 3.0),
    ("T2_in", 40.0, 120.0),        # ms
    ("T2_ex", 30.0, 100.0),
    ("G_max", 0.02, 0.3),          # T/m
    ("lobe_frac", 0.3, 0.95),      # lobe length / half echo time
    ("periods", 1.0, 4.0),         # OGSE periods per lobe
    ("family", 0.0, 1.0),          # [0, .25) PGSE, cos, sin, [.75, 1) square wave
]
N_SCALARS = len(SCALARS)
N_WAVE = 256
N_DIRS = 30
N_CYL = 24
FAMILIES = ("pgse", "cos", "sin", "square")


@dataclass
class Protocol:
    """Echo time and time step of every query (fixed for one surrogate)."""
    TE: float = 30.0          # ms
    dt: float = 2.5e-3        # ms

    @property
    def n_steps(self) -> int:
        return int(round(self.TE / self.dt)) + 1


def directions(n: int = N_DIRS) -> np.ndarray:
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(1.0 - z * z)
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def sobol_unit(n: int, seed: int, dims: int = N_SCALARS) -> np.ndarray:
    """n scrambled Sobol points in [0, 1)^dims."""
    from scipy.stats import qmc

    return qmc.Sobol(d=dims, scramble=True, seed=seed).random(n)


def physical(u: np.ndarray) -> np.ndarray:
    """Unit coordinates -> physical values (columns as SCALARS)."""
    lo = np.array([s[1] for s in SCALARS])
    hi = np.array([s[2] for s in SCALARS])
    return lo + np.asarray(u, np.float64) * (hi - lo)


def unit(values: np.ndarray) -> np.ndarray:
    lo = np.array([s[1] for s in SCALARS])
    hi = np.array([s[2] for s in SCALARS])
    return (np.asarray(values, np.float64) - lo) / (hi - lo)


def functor(v: np.ndarray, proto: Protocol, ghat=(1.0, 0.0, 0.0)):
    """Gradient functor of one query (physical scalars v)."""
    half = 0.5 * (proto.TE - 2.0 * proto.dt)
    lobe = v[9] * half
    gap = proto.TE - 2.0 * proto.dt - 2.0 * lobe
    G = v[8] * T_PER_M
    fam = FAMILIES[min(3, int(v[11] * 4.0))]
    if fam == "pgse":
        return wf.RectGrad(G, lobe, gap, ghat)
    if fam == "cos":
        return wf.CosGFunc(G, lobe, gap, v[10], ghat)
    if fam == "sin":
        return wf.SinGFunc(G, lobe, gap, v[10], ghat)
    return wf.SWOGSEFunc(G, lobe, gap, v[10], ghat)


def waveform_samples(v: np.ndarray, proto: Protocol) -> np.ndarray:
    """The 256-sample unit waveform (midpoints of 256 equal bins of the echo
    time) that the network sees."""
    f = functor(v, proto)
    t = (np.arange(N_WAVE) + 0.5) * proto.TE / N_WAVE
    return f.profile(t).astype(np.float32)


def features(u: np.ndarray, waves: np.ndarray) -> np.ndarray:
    """Network input: unit scalars mapped to [-1, 1], then the waveform."""
    return np.concatenate([2.0 * np.asarray(u, np.float32) - 1.0, np.asarray(waves, np.float32)], axis=1)


def _pack_one(args):
    """One packing; a draw of radii that does not relax to a non-overlapping
    state (about 2 % at the top fraction) is replaced by the next seed."""
    r_mean, k, frac, seed = args
    for attempt in range(32):
        try:
            sub = sb.gamma_packing(N_CYL, k, r_mean / k, frac, D_in=1.0, D_ex=1.0,
                                   seed=seed + 7919 * attempt, rmin=0.3 * r_mean)
            return sub.a, sub.centers[:, :2].copy(), sub.radii.copy()
        except RuntimeError:
            continue
    raise RuntimeError(f"no packing for r_mean={r_mean}, k={k}, fraction={frac}")


def geometries(u_geom: np.ndarray, seed: int, processes: int = 0):
    """Gamma packings for rows of unit (r_mean, k_shape, fraction)."""
    v = physical(np.concatenate([u_geom, np.zeros((len(u_geom), N_SCALARS - 3))], axis=1))
    args = [(v[i, 0], v[i, 1], v[i, 2], seed * 1000003 + i) for i in range(len(v))]
    if processes and len(args) > 8:
        from multiprocessing import get_context

        with get_context("fork").Pool(processes) as pool:
            return pool.map(_pack_one, args, chunksize=8)
    return [_pack_one(a) for a in args]


def instance(v: np.ndarray, geom, proto: Protocol, n_walkers: int, seed: int, inst_id: int):
    """(substrate, acquisition, run) of one query for the engine."""
    box, centres, radii = geom
    n = len(radii)
    cs = np.zeros((n, 3))
    cs[:, :2] = centres
    dt = proto.dt
    sub = abi.SubstrateSpec(a=box, b=box, c=box, D_lat=v[5], T2_lat=v[7],
                            types=np.full(n, abi.PRIM_CYLINDER_XY), centers=cs, radii=radii,
                            D=v[4], T2=v[6], kappa=sb.kappa_for_probability(v[3], v[4], dt))
    fs = [functor(v, proto, d) for d in directions()]
    acq = wf.build_acq(fs, proto.n_steps, dt, GAMMA)
    run = abi.RunSpec(n_walkers=int(n_walkers), seed=int(seed), instance=int(inst_id),
                      boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
    return sub, acq, run


def simulate(values: np.ndarray, geoms: Sequence, geom_index: np.ndarray, proto: Protocol,
             n_walkers: int, seed: int = 1234, chunk: int = 512, ctx=None,
             first_instance: int = 0) -> np.ndarray:
    """Monte Carlo signals [Q][30] of queries (physical values, geometry
    index per query) on the batched engine."""
    from . import engine

    Q = len(values)
    out = np.zeros((Q, N_DIRS), np.float64)
    for c0 in range(0, Q, chunk):
        idx = range(c0, min(Q, c0 + chunk))
        insts = [instance(values[q], geoms[geom_index[q]], proto, n_walkers, seed, first_instance + q)
                 for q in idx]
        with engine.Batch(insts, ctx=ctx) as b:
            b.run()
            for k, q in enumerate(idx):
                out[q] = b.results(k).E
    return out


class Surrogate:
    """Network + training box + Monte Carlo fallback."""

    def __init__(self, proto: Protocol = Protocol(), hidden: int = 1024, depth: int = 6,
                 max_batch: int = 65536, ctx=None):
        from . import engine

        self.proto = proto
        self.dims = [N_SCALARS + N_WAVE] + [hidden] * depth + [N_DIRS]
        self.net = engine.MLP(self.dims, max_batch=max_batch, ctx=ctx)
        self.ctx = ctx

    def init(self, seed: int = 0):
        rng = np.random.default_rng(seed)
        parts = []
        for a, b in zip(self.dims[:-1], self.dims[1:]):
            lim = math.sqrt(6.0 / a) / 2.0     # He-uniform, halved for the GELU stack
            parts.append(rng.uniform(-lim, lim, size=a * b))
            parts.append(np.zeros(b))
        self.net.set_params(np.concatenate(parts).astype(np.float32))

    def fit(self, X: np.ndarray, Y: np.ndarray, steps: int, batch: int = 4096, lr: float = 1e-3,
            seed: int = 1, lr_final: float = 1e-5) -> np.ndarray:
        """Adam with a step-wise exponential learning-rate decay from lr to
        lr_final over `steps` (in 10 stages)."""
        self.net.set_data(X, Y)
        traces = []
        stages = 10
        per = max(1, steps // stages)
        for s in range(stages):
            lr_s = lr * (lr_final / lr) ** (s / max(1, stages - 1))
            traces.append(self.net.train(per, batch=batch, lr=lr_s, seed=seed + s))
        return np.concatenate(traces)

    def predict_features(self, X: np.ndarray) -> np.ndarray:
        out = np.empty((len(X), N_DIRS), np.float32)
        mb = 65536
        for i in range(0, len(X), mb):
            out[i:i + mb] = self.net.forward(X[i:i + mb])
        return out

    @staticmethod
    def in_hull(u: np.ndarray, waves: np.ndarray) -> np.ndarray:
        u = np.asarray(u)
        return (np.all((u >= 0.0) & (u <= 1.0), axis=1)
                & np.all(np.abs(np.asarray(waves)) <= 1.0 + 1e-6, axis=1))

    def predict(self, u: np.ndarray, waves: Optional[np.ndarray] = None, mc_walkers: int = 5000,
                seed: int = 99) -> Tuple[np.ndarray, np.ndarray]:
        """Signals of queries given as unit scalars u [Q][12] (waveforms
        derived from them when not given).  Returns (E [Q][30], used_mc [Q])."""
        u = np.asarray(u, np.float64)
        v = physical(u)
        if waves is None:
            waves = np.stack([waveform_samples(v[q], self.proto) for q in range(len(u))])
        inside = self.in_hull(u, waves)
        E = np.zeros((len(u), N_DIRS), np.float64)
        if inside.any():
            E[inside] = self.predict_features(features(u[inside], waves[inside]))
        out = np.nonzero(~inside)[0]
        if out.size:
            geoms = geometries(u[out][:, :3].clip(0.0, None), seed=seed)
            E[out] = simulate(v[out], geoms, np.arange(out.size), self.proto, mc_walkers, seed=seed,
                              ctx=self.ctx)
        return E, ~inside

    def close(self):
        self.net.close()
