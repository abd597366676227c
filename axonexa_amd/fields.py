"""Magnetic-field functors of the Bloch path and the field tables they make.

Each class restates one functor of src/functors/bfunctors.cuh: the field
B(r, t) a walker at position r sees at time t (static field B0 along z plus
gradient, concomitant-gradient and RF terms).  Every one of them is affine in
r, so each is given here as ``affine(t) -> (b0 [n, 3], J [n, 3, 3])`` with
B(r, t) = b0(t) + J(t) r, evaluated for an array of times; ``__call__``
evaluates B at one point for checks.  ``build_field`` samples a list of
functors at the times an integrator needs and packs the ``FieldSpec`` the
C ABI takes (``axx_field`` in include/axonexa.h).

Names, constructor arguments and piecewise definitions follow the reference,
including its unusual ones (noted per class).
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np

from . import abi

GAMMA = 267500.0          # ms^-1 T^-1 (the reference drivers' GAMMA)


def _zeros(t):
    t = np.asarray(t, dtype=np.float64).reshape(-1)
    return t, np.zeros((t.size, 3)), np.zeros((t.size, 3, 3))


class FieldFunctor:
    def affine(self, t):
        raise NotImplementedError

    def __call__(self, r, t) -> np.ndarray:
        b0, J = self.affine(np.array([float(t)]))
        return b0[0] + J[0] @ np.asarray(r, dtype=np.float64)

    def step_time_min(self, t) -> Optional[np.ndarray]:
        """Substep length wanted at t (getStepTimeMin), or None."""
        return None


class QSC_GRE(FieldFunctor):
    """Quadrupolar saddle coil, gradient echo (bfunctors.cuh:5-53):
    B = (0, G z, B0 + G y) in the first lobe [0, d), its negative gradient in
    the second [s + d, 2 d + s), (0, 0, B0) otherwise."""

    def __init__(self, B0, G, gradientDuration, gradientSpacing):
        self.B0, self.G = float(B0), float(G)
        self.d, self.s = float(gradientDuration), float(gradientSpacing)

    def affine(self, t):
        t, b0, J = _zeros(t)
        b0[:, 2] = self.B0
        l1 = t < self.d
        l2 = ~l1 & (t >= self.s + self.d) & (t < 2.0 * self.d + self.s)
        g = np.where(l1, self.G, np.where(l2, -self.G, 0.0))
        J[:, 1, 2] = g
        J[:, 2, 1] = g
        return b0, J

    def getB(self):
        return self.G ** 2 * self.d ** 2 * GAMMA ** 2 * ((2.0 / 3.0) * self.d + self.s)

    def getDiffusionTime(self):
        return self.s + self.d


class QSC_PGSE(FieldFunctor):
    """Saddle-coil PGSE with a 180-degree RF pulse at TE/2
    (bfunctors.cuh:55-124): first lobe (0, G z, B0 + G y); during
    [TE/2, TE/2 + rf) the rotating RF field (B1 cos wt, -B1 sin wt, B0) with
    B1 = pi / |gamma rf|, w = gamma B0; second lobe [s + d, 2 d + s) with the
    same sign as the first; (0, 0, B0) otherwise.  Substeps: dtmin during the
    pulse, dtmax elsewhere."""

    def __init__(self, B0, G, gradientDuration, gradientSpacing, rfDuration, dtmin, dtmax, TE, gamma=GAMMA):
        self.B0, self.G = float(B0), float(G)
        self.d, self.s = float(gradientDuration), float(gradientSpacing)
        self.rf, self.dtmin, self.dtmax, self.TE = float(rfDuration), float(dtmin), float(dtmax), float(TE)
        self.gamma = float(gamma)
        self.B1 = math.pi / abs(self.gamma * self.rf)

    def affine(self, t):
        t, b0, J = _zeros(t)
        b0[:, 2] = self.B0
        l1 = t < self.d
        rf = ~l1 & (t >= self.TE / 2.0) & (t < self.TE / 2.0 + self.rf)
        l2 = ~l1 & ~rf & (t >= self.s + self.d) & (t < 2.0 * self.d + self.s)
        w = self.gamma * self.B0
        b0[rf, 0] = self.B1 * np.cos(w * t[rf])
        b0[rf, 1] = -self.B1 * np.sin(w * t[rf])
        g = np.where(l1 | l2, self.G, 0.0)
        J[:, 1, 2] = g
        J[:, 2, 1] = g
        return b0, J

    def step_time_min(self, t):
        t = np.asarray(t, dtype=np.float64)
        on = (t >= self.TE / 2.0) & (t < self.TE / 2.0 + self.rf)
        return np.where(on, self.dtmin, self.dtmax)

    def getB(self):
        return self.G ** 2 * self.d ** 2 * GAMMA ** 2 * ((2.0 / 3.0) * self.d + self.s)

    def getDiffusionTime(self):
        return self.s + self.d


class QSC_COS_GRE(FieldFunctor):
    """Saddle coil, cosine gradient echo (bfunctors.cuh:126-196): the
    QSC_GRE field with the gradient scaled by cos(2 pi t / T), T = d / N, in
    each lobe (time measured from the lobe start), negative in the second."""

    def __init__(self, B0, G, gradientDuration, gradientSpacing, numOfPeriods):
        self.B0, self.G = float(B0), float(G)
        self.d, self.s = float(gradientDuration), float(gradientSpacing)
        self.n = float(numOfPeriods)
        self.b = GAMMA * GAMMA * 0.25 * self.G ** 2 * self.d ** 3 / (math.pi ** 2 * self.n ** 2)
        self.period = self.d / self.n

    def affine(self, t):
        t, b0, J = _zeros(t)
        b0[:, 2] = self.B0
        l1 = t < self.d
        l2 = ~l1 & (t >= self.s + self.d) & (t < 2.0 * self.d + self.s)
        m1 = np.cos(2.0 * math.pi * t / self.period)
        m2 = np.cos(2.0 * math.pi * (t - self.s - self.d) / self.period)
        g = np.where(l1, self.G * m1, np.where(l2, -self.G * m2, 0.0))
        J[:, 1, 2] = g
        J[:, 2, 1] = g
        return b0, J

    def getB(self):
        return self.b

    def getFreq(self):
        return 1.0 / self.period


class QSC_GRE_TRAP(FieldFunctor):
    """Saddle coil with trapezoidal lobes (bfunctors.cuh:198-284), as the
    reference defines it: ramps multiply the gradient by ampMod1(t) = (G/ramp) t
    and ampMod2(t) = G - (G/ramp) t (so G enters twice), the ramp-down
    interval is [ramp + d, 3 ramp + d), and the second lobe (from
    t5 = s + 3 ramp + d) repeats the first with the same sign."""

    def __init__(self, B0, G, gradientDuration, gradientSpacing, gradientRampTime):
        self.B0, self.G = float(B0), float(G)
        self.d, self.s, self.ramp = float(gradientDuration), float(gradientSpacing), float(gradientRampTime)

    def affine(self, t):
        t, b0, J = _zeros(t)
        b0[:, 2] = self.B0
        G, r = self.G, self.ramp
        t1, t2 = r, r + self.d
        t3 = 2.0 * t1 + t2
        t5 = self.s + t3
        up = lambda u: (G / r) * u
        down = lambda u: G - (G / r) * u
        conds = [t < t1, (t >= t1) & (t < t2), (t >= t2) & (t < t3),
                 (t >= t5) & (t < t5 + t1), (t >= t5 + t1) & (t < t5 + t2), (t >= t5 + t2) & (t < t5 + t3)]
        vals = [G * up(t), np.full_like(t, G), G * down(t - t2),
                G * up(t - t5), np.full_like(t, G), G * down(t - t5 - t2)]
        g = np.select(conds, vals, 0.0)
        J[:, 1, 2] = g
        J[:, 2, 1] = g
        return b0, J


class _LongitudinalLobes(FieldFunctor):
    """B = (0, 0, B0 + m(t) G.r) in the first lobe, (0, 0, B0 - m2(t) G.r) in
    the second, (0, 0, B0) otherwise."""

    def __init__(self, B0, G, gradientDuration, gradientSpacing):
        self.B0 = float(B0)
        self.Gv = np.asarray(G, dtype=np.float64).reshape(3)
        self.d, self.s = float(gradientDuration), float(gradientSpacing)

    def _mods(self, t):
        return np.ones_like(t), np.ones_like(t)

    def affine(self, t):
        t, b0, J = _zeros(t)
        b0[:, 2] = self.B0
        l1 = t < self.d
        l2 = ~l1 & (t >= self.s + self.d) & (t < 2.0 * self.d + self.s)
        m1, m2 = self._mods(t)
        g = np.where(l1, m1, np.where(l2, -m2, 0.0))
        J[:, 2, :] = g[:, None] * self.Gv[None, :]
        return b0, J


class SIN_GRE(_LongitudinalLobes):
    """Sine gradient echo along G (bfunctors.cuh:286-330): lobe factors
    sin(2 pi t / period), sin(2 pi (t - s - d) / period)."""

    def __init__(self, B0, G, gradientDuration, gradientSpacing, period):
        super().__init__(B0, G, gradientDuration, gradientSpacing)
        self.period = float(period)

    def _mods(self, t):
        return (np.sin(2.0 * math.pi * t / self.period),
                np.sin(2.0 * math.pi * (t - self.s - self.d) / self.period))


class COS_GRE(_LongitudinalLobes):
    """Cosine gradient echo along G (bfunctors.cuh:332-391): period d / N,
    b = gamma^2 |G|^2 d^3 / (4 pi^2 N^2).  COS_GRE(B0, Gvec, d, s, N) or
    COS_GRE(B0, G, ghat, d, s, N)."""

    def __init__(self, B0, *args):
        if len(args) == 5:
            Gm, ghat, d, s, n = args
            Gv = np.asarray(ghat, dtype=np.float64) * float(Gm)
            gmag = float(Gm)
        else:
            Gv, d, s, n = args
            Gv = np.asarray(Gv, dtype=np.float64)
            gmag = float(np.linalg.norm(Gv))
        super().__init__(B0, Gv, d, s)
        self.period = self.d / float(n)
        self.b = GAMMA * GAMMA * 0.25 * gmag * gmag * self.d ** 3 / (math.pi ** 2 * float(n) ** 2)

    def _mods(self, t):
        return (np.cos(2.0 * math.pi * t / self.period),
                np.cos(2.0 * math.pi * (t - self.s - self.d) / self.period))

    def getB(self):
        return self.b

    def getFreq(self):
        return 1.0 / self.period


class RECT_GRE(_LongitudinalLobes):
    """Rectangular gradient echo along G (bfunctors.cuh:393-425)."""


class UCN_lowfield(FieldFunctor):
    """Low-field cell with a pi/2 RF pulse (bfunctors.cuh:427-476): during
    [0, rf) B = (B1 cos wt - G x / 2, -B1 sin wt - G y / 2, G z + B0) with
    B1 = pi / (2 |gamma rf|), w = gamma B0; afterwards the same without the
    RF terms (a divergence-free linear gradient).  Substeps: dtmin for
    t <= rf, dtmax after."""

    def __init__(self, B0, G, rfDuration, dtmin, dtmax, gamma=GAMMA):
        self.B0, self.G, self.rf = float(B0), float(G), float(rfDuration)
        self.dtmin, self.dtmax = float(dtmin), float(dtmax)
        self.gamma = float(gamma)
        self.B1 = math.pi / (2.0 * abs(self.gamma * self.rf))
        self.w = self.gamma * self.B0

    def affine(self, t):
        t, b0, J = _zeros(t)
        on = t < self.rf
        b0[on, 0] = self.B1 * np.cos(self.w * t[on])
        b0[on, 1] = -self.B1 * np.sin(self.w * t[on])
        b0[:, 2] = self.B0
        J[:, 0, 0] = -0.5 * self.G
        J[:, 1, 1] = -0.5 * self.G
        J[:, 2, 2] = self.G
        return b0, J

    def step_time_min(self, t):
        t = np.asarray(t, dtype=np.float64)
        return np.where(t <= self.rf, self.dtmin, self.dtmax)


DOPRI_NODES = (0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0)


def sample_times(n_steps: int, dt: float, integrator: int, substeps: np.ndarray) -> np.ndarray:
    """Every sample time, step after step (the order of the field table)."""
    out = []
    for i in range(n_steps):
        s = int(substeps[i])
        h = dt / s
        ti = i * dt
        if integrator == abi.MAG_ROTATE:
            out.append(ti + np.arange(s) * h)
        elif integrator == abi.MAG_RK4:
            out.append(ti + np.arange(2 * s + 1) * (h / 2.0))
        else:
            k = np.repeat(np.arange(s, dtype=np.float64), 6)
            c = np.tile(np.array(DOPRI_NODES), s)
            out.append(ti + (k + c) * h)
    return np.concatenate(out)


def auto_substeps(functors: Sequence[FieldFunctor], n_steps: int, dt: float) -> np.ndarray:
    """Substeps per step from the functors' getStepTimeMin at the step start:
    ceil(dt / hmin) (the reference's multistep loop takes steps of hmin until
    dt is covered)."""
    t = np.arange(n_steps) * dt
    s = np.ones(n_steps, dtype=np.int64)
    for f in functors:
        h = f.step_time_min(t)
        if h is None:
            continue
        s = np.maximum(s, np.ceil(dt / np.asarray(h) - 1e-9).astype(np.int64))
    return np.clip(s, 1, 4096).astype(np.int32)


def build_field(functors: Sequence[FieldFunctor], n_steps: int, dt: float, gamma: float = GAMMA,
                integrator: int = abi.MAG_ROTATE, substeps=None, m0=(1.0, 0.0, 0.0),
                origin=(0.0, 0.0, 0.0), T1=None) -> abi.FieldSpec:
    """Sample every functor for a walk of n_steps steps of dt.  substeps:
    None (one per step), "auto" (from getStepTimeMin) or an array."""
    if substeps is None:
        sub = np.ones(n_steps, dtype=np.int32)
        sub_arg = None
    else:
        sub = auto_substeps(functors, n_steps, dt) if isinstance(substeps, str) else \
            np.asarray(substeps, dtype=np.int32).reshape(n_steps)
        sub_arg = sub
    t = sample_times(n_steps, dt, integrator, sub)
    samples = np.empty((t.size, len(functors), 12))
    for j, f in enumerate(functors):
        b0, J = f.affine(t)
        samples[:, j, :3] = b0
        samples[:, j, 3:] = J.reshape(-1, 9)
    return abi.FieldSpec(n_steps=n_steps, dt=dt, gamma=gamma, samples=samples, integrator=integrator,
                         substeps=sub_arg, origin=tuple(origin), m0=tuple(m0), T1=T1)
