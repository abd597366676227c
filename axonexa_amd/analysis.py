"""Post-processing of simulated signals (the acquisition side of the
reference: src/acquisition/phaseAcquisition.cuh:46-85 and
src/util/misc.cuh:41-59).

  * ``signal``       |S_j| = sqrt(Mx_j^2 + My_j^2) (phaseAcquisition.cuh:46),
                     normalised by the walker count when asked;
  * ``adc_fit``      ADC = -slope of the least-squares line of ln|S| on b
                     (calcADC, phaseAcquisition.cuh:54-77, with the closed
                     normal equations of linear_regression, misc.cuh:41-59);
  * ``RepeatStats``  running mean and mean square of repeated fits, the
                     reference's points / ADC / ADCmeansquare / getADCError;
  * ``table``        the drivers' stdout table (Cos_OneCyl.cu:153-164):
                     one line "freq G Mx My" per waveform.
"""
from __future__ import annotations

import math
from typing import Iterable, Sequence

import numpy as np


def signal(mx: Sequence[float], my: Sequence[float], n_walkers: int = 1) -> np.ndarray:
    mx = np.asarray(mx, np.float64)
    my = np.asarray(my, np.float64)
    return np.sqrt(mx * mx + my * my) / float(n_walkers)


def adc_fit(b: Sequence[float], s: Sequence[float]) -> float:
    """-slope of ln s against b by least squares (s > 0)."""
    x = np.asarray(b, np.float64)
    y = np.log(np.asarray(s, np.float64))
    if x.size != y.size or x.size < 2:
        raise ValueError("adc_fit needs matching b and signal arrays of length >= 2")
    n = float(x.size)
    sxy, sx, sy, sxx = float(np.sum(x * y)), float(np.sum(x)), float(np.sum(y)), float(np.sum(x * x))
    delta = n * sxx - sx * sx
    if delta == 0.0:
        raise ValueError("adc_fit: all b-values are equal")
    return -(n * sxy - sx * sy) / delta


class RepeatStats:
    """Running mean / mean square over repeated estimates (replicas)."""

    def __init__(self):
        self.points = 0
        self.mean = 0.0
        self.mean_square = 0.0

    def add(self, v: float) -> None:
        self.points += 1
        k = self.points
        self.mean = (self.mean * (k - 1) + v) / k
        self.mean_square = (self.mean_square * (k - 1) + v * v) / k

    def error(self) -> float:
        """Standard deviation of the estimates (getADCError)."""
        return math.sqrt(max(self.mean_square - self.mean * self.mean, 0.0))


def table(freqs: Iterable[float], G: Iterable[float], mx: Iterable[float], my: Iterable[float]) -> str:
    """'freq G Mx My' lines with 20 significant digits."""
    out = []
    for f, g, a, b in zip(freqs, G, mx, my):
        out.append(f"{f:.20g} {g:.20g} {a:.20g} {b:.20g} ")
    return "\n".join(out)
