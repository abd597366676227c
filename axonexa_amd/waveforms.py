"""Gradient waveforms: the reference's functors, sampled into unit profiles.

The reference evaluates every functor ``G_j(t)`` once per walker per step on
the device (src/kernels/kernelLattice.cuh:82-85 calling
src/functors/gfunctors.cuh).  Here each functor is split into an amplitude, a
direction and a unit-amplitude time profile; the profile is sampled once on
the host at ``t_i = i*dt`` (in float64, exactly like ``G[j](par->timestep*i)``)
and cast to float32.  Functors that share a time shape share one profile, so
the kernel accumulates one moment per (profile, axis) and the phase of every
waveform follows by linearity.

Shapes (gfunctors.cuh):
  RectGrad     :3-66    PGSE, lobes [0, d) and [gap+d, 2d+gap)
  CosGFunc     :68-120  cos(2 pi t / period), period = T / N
  SinGFunc     :122-174 sin(2 pi t / period)
  ApodCosGFunc :176-234 apodised cosine
  SWOGSEFunc   :236-290 square wave; the reference's pow(-1, 2 t v) is NaN for
                        non-integer 2tv (defect D5); here (-1)^floor(2 v t).
``N`` (periods per lobe) may be real-valued, so OGSE can be specified by
frequency with a fixed lobe duration.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Iterable, List, Sequence, Tuple

import numpy as np

from .abi import AcqSpec

# Default gyromagnetic ratio in the reference's units, rad / (ms * T)
# (examples/Cos_OneCyl.cu:5).
GAMMA_MS_T = 267500.0


def _unit(v) -> np.ndarray:
    v = np.asarray(v, dtype=np.float64).reshape(3)
    n = np.linalg.norm(v)
    if n == 0.0:
        raise ValueError("gradient direction must be non-zero")
    return v / n


class _TwoLobe:
    """G(t) = G*m(t) on [0,T), -G*m(t-gap-T) on [gap+T, 2T+gap), else 0."""

    kind = "base"

    def __init__(self, G: float, T: float, gap: float, N: float = 1.0, ghat=(1.0, 0.0, 0.0)):
        self.G = float(G)
        self.T = float(T)
        self.gap = float(gap)
        self.N = float(N)
        self.ghat = _unit(ghat)

    # amplitude modulation on [0, T)
    def amp_mod(self, t: np.ndarray) -> np.ndarray:  # pragma: no cover - abstract
        raise NotImplementedError

    def profile(self, t: np.ndarray) -> np.ndarray:
        """Unit-amplitude G(t)/G as float64 (gfunctors.cuh operator())."""
        t = np.asarray(t, dtype=np.float64)
        out = np.zeros_like(t)
        l1 = t < self.T
        l2 = (t >= self.gap + self.T) & (t < 2.0 * self.T + self.gap)
        out[l1] = self.amp_mod(t[l1])
        out[l2] = -self.amp_mod(t[l2] - self.gap - self.T)
        return out

    def __call__(self, t: float) -> np.ndarray:
        return self.G * self.profile(np.array([t]))[0] * self.ghat

    def key(self) -> Tuple:
        return (self.kind, self.T, self.gap, self.N)

    def getG(self) -> float:
        return self.G

    def getFreq(self) -> float:
        return self.N / self.T


class RectGrad(_TwoLobe):
    """PGSE (gfunctors.cuh:3-66): duration delta, spacing gap."""

    kind = "rect"

    def __init__(self, G: float, delta: float, gap: float, ghat=(1.0, 0.0, 0.0)):
        super().__init__(G, delta, gap, 1.0, ghat)

    def amp_mod(self, t):
        return np.ones_like(t)

    def getB(self, gamma: float = GAMMA_MS_T) -> float:
        # gfunctors.cuh:31-33
        d = self.T
        return self.G ** 2 * d ** 2 * gamma ** 2 * ((2.0 / 3.0) * d + self.gap)

    def getFreq(self) -> float:
        return 0.0

    def getDiffusionTime(self) -> float:
        return self.gap + self.T


class CosGFunc(_TwoLobe):
    """Cosine OGSE (gfunctors.cuh:68-120)."""

    kind = "cos"

    def amp_mod(self, t):
        return np.cos(2.0 * math.pi * t * self.N / self.T)

    def getB(self, gamma: float = GAMMA_MS_T) -> float:
        # gfunctors.cuh:89-91 (valid for integer N)
        return gamma ** 2 * 0.25 * self.G ** 2 * self.T ** 3 / (math.pi ** 2 * self.N ** 2)


class SinGFunc(_TwoLobe):
    """Sine OGSE (gfunctors.cuh:122-174); b prefactor 3/4 (defect D6)."""

    kind = "sin"

    def amp_mod(self, t):
        return np.sin(2.0 * math.pi * t * self.N / self.T)

    def getB(self, gamma: float = GAMMA_MS_T) -> float:
        return gamma ** 2 * 0.75 * self.G ** 2 * self.T ** 3 / (math.pi ** 2 * self.N ** 2)


class ApodCosGFunc(_TwoLobe):
    """Apodised cosine OGSE (gfunctors.cuh:176-234)."""

    kind = "apodcos"

    def amp_mod(self, t):
        period = self.T / self.N
        mid = (t >= period / 4.0) & (t < self.T - period / 4.0)
        return np.where(mid, np.cos(2.0 * math.pi * t / period), np.sin(4.0 * math.pi * t / period))

    def getB(self, gamma: float = GAMMA_MS_T) -> float:
        return (gamma ** 2 * 0.25 * self.G ** 2 * self.T ** 3 / (math.pi ** 2 * self.N ** 2)
                * (1.0 - 1.0 / (8.0 * self.N)))


class SWOGSEFunc(_TwoLobe):
    """Square-wave OGSE (gfunctors.cuh:236-290) with (-1)^floor(2 v t)."""

    kind = "sw"

    def amp_mod(self, t):
        v = self.N / self.T
        return np.where(np.floor(2.0 * v * t).astype(np.int64) % 2 == 0, 1.0, -1.0)

    def getB(self, gamma: float = GAMMA_MS_T) -> float:
        # Integer N: in every half period h = 1/(2v) the dephasing q ramps
        # linearly 0 -> gamma*G*h -> 0, so each lobe gives
        # 2N * gamma^2 G^2 h^3 / 3 = gamma^2 G^2 T / (12 v^2); two lobes give
        # gamma^2 G^2 T / (6 v^2) — gfunctors.cuh:257-263 with a1 = T.
        v = self.N / self.T
        return gamma ** 2 * self.G ** 2 * self.T / (6.0 * v ** 2)


def sample_times(n_steps: int, dt: float) -> np.ndarray:
    """t_i = dt * i in float64, as G[j](par->timestep * i)."""
    return dt * np.arange(n_steps, dtype=np.float64)


def build_acq(functors: Sequence[_TwoLobe], n_steps: int, dt: float,
              gamma: float = GAMMA_MS_T) -> AcqSpec:
    """Factor functors into shared unit profiles and sample them."""
    t = sample_times(n_steps, dt)
    keys: List[Tuple] = []
    cols: List[np.ndarray] = []
    wf_p, wf_a, wf_d = [], [], []
    for f in functors:
        k = f.key()
        if k not in keys:
            keys.append(k)
            cols.append(f.profile(t).astype(np.float32))
        wf_p.append(keys.index(k))
        wf_a.append(f.G)
        wf_d.append(f.ghat)
    prof = np.stack(cols, axis=1) if cols else np.zeros((n_steps, 1), np.float32)
    return AcqSpec(n_steps=n_steps, dt=dt, gamma=gamma, profiles=prof,
                   wf_profile=np.array(wf_p, np.int32), wf_amplitude=np.array(wf_a),
                   wf_dir=np.array(wf_d).reshape(-1, 3))


def q_tail(profile: np.ndarray) -> np.ndarray:
    """Q_k = sum_{i>=k} g_i for k = 0..n-1 (float64 sums of the float32 table)."""
    g = np.asarray(profile, dtype=np.float64)
    return np.cumsum(g[::-1])[::-1]


def b_discrete(profile: np.ndarray, amplitude: float, dt: float, gamma: float) -> float:
    """b of the sampled waveform as the engine integrates it.

    phi = gamma*dt*sum_{k>=1} Q_k . Delta_k for unit-variance steps; for
    Gaussian steps of variance 2*D*dt per axis, Var(phi) = 2*b*D with
    b = gamma^2 amp^2 dt^3 sum_{k=1}^{n-1} Q_k^2 (the discrete analogue of
    b = int (gamma int_0^t G)^2 dt, src/external/bvalue.h:3-30)."""
    Q = q_tail(profile)[1:]
    return float(gamma ** 2 * amplitude ** 2 * dt ** 3 * np.sum(Q * Q))


def b_quadrature(functor: _TwoLobe, gamma: float = GAMMA_MS_T, n: int = 200000) -> float:
    """b = int_0^TE (gamma int_0^t G(s) ds)^2 dt by the trapezoid rule
    (numeric definition of src/external/bvalue.h:3-30)."""
    TE = 2.0 * functor.T + functor.gap
    t = np.linspace(0.0, TE, n + 1)
    g = functor.G * functor.profile(t)
    g[-1] = 0.0
    q = gamma * np.concatenate([[0.0], np.cumsum(0.5 * (g[1:] + g[:-1]) * np.diff(t))])
    return float(np.trapezoid(q * q, t))


def free_signal_exact(acq: AcqSpec, j: int, D: float, cell=None) -> complex:
    """Exact mean of e^{i phi_j} for free diffusion with fixed-length steps.

    E = E_r0[exp(i gamma dt Qtilde_0 . r0)] * prod_{k>=1} sinc(gamma dt |Q_k| L),
    L = sqrt(6 D dt), sinc(x) = sin(x)/x; r0 uniform in the cell (if given)."""
    p = int(acq.wf_profile[j])
    amp = float(acq.wf_amplitude[j])
    d = acq.wf_dir[j]
    Q = q_tail(acq.profiles[:, p]) * amp
    L = math.sqrt(6.0 * D * acq.dt)
    x = acq.gamma * acq.dt * np.abs(Q[1:]) * L
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(x > 0, np.sin(x) / x, 1.0)
    val = complex(np.prod(s))
    if cell is not None and Q[0] != 0.0:
        for ax in range(3):
            kappa = acq.gamma * acq.dt * Q[0] * d[ax]
            if kappa == 0.0:
                continue
            Lc = cell[ax]
            val *= (np.exp(1j * kappa * Lc) - 1.0) / (1j * kappa * Lc)
    return val
