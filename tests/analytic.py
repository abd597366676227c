"""Analytic references for the restricted-diffusion checks.

Gaussian-phase approximation (GPD) for PGSE perpendicular to an impermeable
cylinder or inside a sphere (Murday & Cotts 1968; Neuman 1974; written as in
van Gelderen et al., J. Magn. Reson. B 103:255, 1994):

  ln E = -2 gamma^2 g^2 sum_m [2 D a_m^2 d - 2 + 2 e^{-D a_m^2 d} + 2 e^{-D a_m^2 Delta}
                               - e^{-D a_m^2 (Delta - d)} - e^{-D a_m^2 (Delta + d)}]
                              / [D^2 a_m^6 (R^2 a_m^2 - nu)]

with a_m R the roots of J1'(x) = 0 and nu = 1 (cylinder), or of
x j1'(x)... i.e. x J'_{3/2}(x) - J_{3/2}(x)/2 = 0 and nu = 2 (sphere).
Long-pulse limits d >> R^2/D (Neuman 1974):
  cylinder ln E = -(7/96) gamma^2 g^2 R^4 / D * (2 d - 99 R^2 / (112 D))
  sphere   ln E = -(8/175) gamma^2 g^2 R^4 / D * (2 d - 581 R^2 / (840 D)).
"""
import math

import numpy as np
from scipy import optimize, special


def cylinder_roots(n=60):
    return special.jnp_zeros(1, n)


def sphere_roots(n=60):
    f = lambda x: x * special.jvp(1.5, x) - 0.5 * special.jv(1.5, x)
    roots, x = [], 0.5
    grid = np.linspace(0.5, 400, 40000)
    v = f(grid)
    for i in range(len(grid) - 1):
        if v[i] == 0 or v[i] * v[i + 1] < 0:
            roots.append(optimize.brentq(f, grid[i], grid[i + 1]))
            if len(roots) == n:
                break
    return np.array(roots)


def gpd_pgse(gamma, g, delta, Delta, D, R, shape="cylinder", n=60):
    x = cylinder_roots(n) if shape == "cylinder" else sphere_roots(n)
    nu = 1.0 if shape == "cylinder" else 2.0
    a = x / R
    Da2 = D * a * a
    num = (2 * Da2 * delta - 2 + 2 * np.exp(-Da2 * delta) + 2 * np.exp(-Da2 * Delta)
           - np.exp(-Da2 * (Delta - delta)) - np.exp(-Da2 * (Delta + delta)))
    den = D * D * a ** 6 * (R * R * a * a - nu)
    return math.exp(-2 * gamma * gamma * g * g * float(np.sum(num / den)))


def neuman_limit(gamma, g, delta, D, R, shape="cylinder"):
    if shape == "cylinder":
        k, c = 7.0 / 96.0, 99.0 / 112.0
    else:
        k, c = 8.0 / 175.0, 581.0 / 840.0
    return math.exp(-k * gamma * gamma * g * g * R ** 4 / D * (2.0 * delta - c * R * R / D))


def gpd_pgse_planes(gamma, g, delta, Delta, D, l, n_terms=400):
    """GPD for PGSE normal to two parallel reflecting planes a distance l
    apart: the position autocorrelation is sum_{n odd} 8 l^2/(n^4 pi^4)
    exp(-D (n pi / l)^2 t), so
      ln E = -gamma^2 g^2 sum_{n odd} [8 l^2/(n^4 pi^4)] F(a_n) / (D^2 a_n^4),
    F(a) = 2 D a^2 d - 2 + 2e^{-D a^2 d} + 2e^{-D a^2 Delta}
           - e^{-D a^2 (Delta - d)} - e^{-D a^2 (Delta + d)}, a_n = n pi / l
    (the same form as the cylinder formula above with its coefficients)."""
    n = np.arange(1, 2 * n_terms, 2, dtype=float)
    a = n * math.pi / l
    Da2 = D * a * a
    F = (2 * Da2 * delta - 2 + 2 * np.exp(-Da2 * delta) + 2 * np.exp(-Da2 * Delta)
         - np.exp(-Da2 * (Delta - delta)) - np.exp(-Da2 * (Delta + delta)))
    B = 8 * l * l / (n ** 4 * math.pi ** 4)
    return math.exp(-gamma * gamma * g * g * float(np.sum(B * F / (D * D * a ** 4))))
