"""The benchmark and validation workloads, as concrete synthetic inputs.

Engine units: um, ms, T; gamma = 2.675e5 rad/(ms T); gradients in T/um
(1 T/m = 1e-6 T/um).

  c1_free      periodic 20 um box, D = 2.0, PGSE d = 5 ms, Delta = 15 ms,
               16 gradients 0..0.3 T/m along x, 1e5 walkers, dt = 1e-3 ms,
               20,000 steps, seed 1234
  c1_cylinder  the same box with one impermeable cylinder R = 4 um
  c2_hex       1000 cylinders R = 2.5 um, spacing 6 um, hexagonal, periodic
               120 x 259.8 um cell; transmission probability 1e-3 per collision
               (from inside); T2 70/50 ms; D 1.7/2.0; 1e6 walkers;
               dt = 2e-4 ms, 250,000 steps; cosine OGSE, lobes 20 ms, gap 5 ms,
               8 frequencies 25..200 Hz x 8 amplitudes to 0.4 T/m, along x
               (perpendicular) and z (parallel)
  c3_gamma     400 cylinders, radii Gamma(k=5, theta=0.4 um), packed to 65 %
               area fraction; 4e6 walkers, dt = 5e-4 ms, 80,000 steps;
               64 profiles (PGSE, cos, sin, square wave x 16 timings) x
               30 directions x 3 amplitudes to 0.3 T/m

  c4_batch     256 instances from a seeded Philox sampler: 128 single spheres
               R ~ U(1, 8) um (walkers inside), 64 hexagonal cylinder
               lattices R ~ U(1, 5) um, 64 gamma packings (40 cylinders);
               walkers log-uniform in [1e4, 1e6]; each instance its own PGSE
               or cosine-OGSE set of 4-32 waveforms; dt = 2e-3 ms, 10,000
               steps

The reference drivers these restate: examples/Cos_OneCyl.cu (one cylinder,
cosine OGSE), examples/hexPhase.cu (hexagonal lattice), examples/gammadist.cu
(gamma packing), examples/planePhase.cu / spherePhase.cu (validation).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

import numpy as np

from . import abi
from . import substrates as sb
from . import waveforms as wf

GAMMA = 2.675e5          # rad / (ms T)
T_PER_M = 1e-6           # T/um per T/m


@dataclass
class Workload:
    name: str
    sub: abi.SubstrateSpec
    acq: abi.AcqSpec
    run: abi.RunSpec
    description: str

    @property
    def walker_steps(self) -> int:
        return int(self.run.n_walkers) * (self.acq.n_steps - 1)

    def flops_per_walker_step(self, rho: float) -> float:
        """Algorithmic FP32 work per walker-step:
        F = 22 + 16*beta + 35*rho + tau + 2*P*A (FMA = 2, sqrt = 1)."""
        beta = 1 if self.sub.n_prims > 0 else 0
        tau = 1 if self.run.relaxation and (np.isfinite(self.sub.T2_lat) or
                                            np.any(np.isfinite(self.sub.T2))) else 0
        A = 0
        for a in range(3):
            if np.any((self.acq.wf_dir[:, a] != 0) & (self.acq.wf_amplitude != 0)):
                A += 1
        return 22 + 16 * beta + 35 * rho + tau + 2 * self.acq.n_profiles * A


def c1_acq(n_steps: int = 20000, dt: float = 1e-3) -> abi.AcqSpec:
    fs = [wf.RectGrad(g * T_PER_M, 5.0, 10.0, (1, 0, 0)) for g in np.linspace(0.0, 0.3, 16)]
    return wf.build_acq(fs, n_steps, dt, GAMMA)


def c1_free(n_walkers: int = 100000, n_steps: int = 20000) -> Workload:
    return Workload("c1_free", sb.empty_box(20.0, 2.0), c1_acq(n_steps),
                    abi.RunSpec(n_walkers=n_walkers, seed=1234),
                    "free diffusion, periodic 20 um box, PGSE 16 gradients")


def c1_cylinder(n_walkers: int = 100000, n_steps: int = 20000) -> Workload:
    return Workload("c1_cylinder", sb.single_cylinder(20.0, 4.0, D_in=2.0, D_ex=2.0), c1_acq(n_steps),
                    abi.RunSpec(n_walkers=n_walkers, seed=1234),
                    "one impermeable cylinder R=4 um in a periodic 20 um box, PGSE 16 gradients")


def c2_functors(lobe: float = 20.0, gap: float = 5.0) -> List[wf.CosGFunc]:
    fs = []
    for f_hz in np.arange(1, 9) * 25.0:
        periods = f_hz * 1e-3 * lobe          # cycles per ms x ms
        for g in np.arange(1, 9) * 0.05:
            for ax in ((1, 0, 0), (0, 0, 1)):
                fs.append(wf.CosGFunc(g * T_PER_M, lobe, gap, periods, ax))
    return fs


def c2_hex(n_walkers: int = 1000000, n_steps: int = 250000, dt: float = 2e-4) -> Workload:
    sub = sb.hex_lattice(2.5, 6.0, 20, 25, D_in=1.7, D_ex=2.0, T2_in=70.0, T2_ex=50.0,
                         kappa=sb.kappa_for_probability(1e-3, 1.7, dt), c=120.0)
    acq = wf.build_acq(c2_functors(), n_steps, dt, GAMMA)
    run = abi.RunSpec(n_walkers=n_walkers, seed=1234, boundary_mode=abi.BOUNDARY_PERMEABLE,
                      relaxation=1)
    return Workload("c2_hex", sub, acq, run,
                    "hex lattice 1000 cylinders R=2.5um s=6um, p=1e-3, T2 70/50 ms, D 1.7/2.0, "
                    "cosine OGSE 8 freq x 8 amp x {x,z}")


def fibonacci_directions(n: int) -> np.ndarray:
    """n nearly uniform unit vectors (Fibonacci sphere)."""
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(1.0 - z * z)
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def c3_functors(n_dirs: int = 30, amps=(0.1, 0.2, 0.3), total: float = 40.0):
    fs = []
    dirs = fibonacci_directions(n_dirs)
    for k in range(16):
        lobe = 4.0 + 1.0 * k                  # 4..19 ms lobes
        gap = total - 2.0 * lobe - 1.0
        for kind in range(4):
            for d in dirs:
                for g in amps:
                    G = g * T_PER_M
                    if kind == 0:
                        fs.append(wf.RectGrad(G, lobe, gap, d))
                    elif kind == 1:
                        fs.append(wf.CosGFunc(G, lobe, gap, 2.0, d))
                    elif kind == 2:
                        fs.append(wf.SinGFunc(G, lobe, gap, 2.0, d))
                    else:
                        fs.append(wf.SWOGSEFunc(G, lobe, gap, 2.0, d))
    return fs


C3_GAP = 0.08   # um, minimum surface-to-surface distance of the gamma packing


def c3_gamma(n_walkers: int = 4000000, n_steps: int = 80000, dt: float = 5e-4,
             seed: int = 4) -> Workload:
    # surfaces at least one extra-axonal step length (0.0775 um) apart, so no
    # walker is caught in the cusp between two nearly touching cylinders
    sub = sb.gamma_packing(400, 5.0, 0.4, 0.65, D_in=1.7, D_ex=2.0, seed=seed, T2_in=70.0,
                           T2_ex=50.0, rmin=0.25, gap=C3_GAP)
    acq = wf.build_acq(c3_functors(), n_steps, dt, GAMMA)
    run = abi.RunSpec(n_walkers=n_walkers, seed=1234, relaxation=1)
    return Workload("c3_gamma", sub, acq, run,
                    "400 gamma-radius cylinders at 65% packing, 64 profiles x 30 dirs x 3 amps")


def _c4_waveforms(rng: np.random.Generator, n_steps: int, dt: float):
    """4..32 waveforms: PGSE (one timing, random directions, linear
    amplitudes) or cosine OGSE (2, 4, 8, 16 or 32 frequencies, at most one
    per waveform, x amplitudes along a random axis)."""
    T = (n_steps - 1) * dt
    M = int(rng.integers(4, 33))
    gmax = rng.uniform(0.05, 0.3) * T_PER_M
    fs = []
    if rng.random() < 0.5:
        delta = rng.uniform(0.1, 0.4) * T
        gap = rng.uniform(0.02, 1.0) * (T - 2 * delta - 2 * dt)
        dirs = rng.normal(size=(M, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        for j in range(M):
            fs.append(wf.RectGrad(gmax * (j + 1) / M, delta, gap, dirs[j]))
    else:
        nf = min(int(rng.choice([2, 4, 8, 16, 32])), M)
        lobe = rng.uniform(0.3, 0.48) * T
        gap = rng.uniform(0.0, 1.0) * (T - 2 * lobe - 2 * dt)
        axis = np.zeros(3)
        axis[int(rng.integers(0, 3))] = 1.0
        for j in range(M):
            periods = 1.0 + (j % nf)
            fs.append(wf.CosGFunc(gmax * (1 + j // nf) / math.ceil(M / nf), lobe, gap, periods, axis))
    return wf.build_acq(fs, n_steps, dt, GAMMA)


def c4_batch(seed: int = 4, n_inst: int = 256, scale: float = 1.0, n_steps: int = 10000,
             dt: float = 2e-3) -> List[Workload]:
    """The batched uneven workload: instance parameters from a Philox
    generator (numpy's Philox bit generator, key = seed); walker counts
    log-uniform in [1e4, 1e6] times `scale`."""
    rng = np.random.Generator(np.random.Philox(seed))
    out = []
    n_sph, n_hex = n_inst // 2, n_inst // 4
    for k in range(n_inst):
        N = max(64, int(round(10.0 ** rng.uniform(4.0, 6.0) * scale)))
        acq = _c4_waveforms(rng, n_steps, dt)
        D_in, D_ex = rng.uniform(1.0, 2.5), rng.uniform(1.5, 3.0)
        if k < n_sph:
            R = rng.uniform(1.0, 8.0)
            sub = sb.single_sphere(2.5 * R, R, D_in=D_in, D_ex=D_ex)
            run = abi.RunSpec(n_walkers=N, seed=1234, instance=k, init_mode=abi.INIT_INSIDE_PRIM,
                              init_arg=0)
            name = f"sphere R={R:.2f}"
        elif k < n_sph + n_hex:
            R = rng.uniform(1.0, 5.0)
            s = 2.0 * R * rng.uniform(1.05, 1.3)
            p = rng.uniform(0.0, 1e-2)
            sub = sb.hex_lattice(R, s, 2, 2, D_in=D_in, D_ex=D_ex, T2_in=rng.uniform(50, 100),
                                 T2_ex=rng.uniform(40, 80), kappa=sb.kappa_for_probability(p, D_in, dt))
            run = abi.RunSpec(n_walkers=N, seed=1234, instance=k,
                              boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
            name = f"hex R={R:.2f} s={s:.2f} p={p:.1e}"
        else:
            theta = rng.uniform(0.25, 0.5)
            frac = rng.uniform(0.4, 0.6)
            p = rng.uniform(0.0, 1e-2)
            sub = sb.gamma_packing(40, 5.0, theta, frac, D_in=D_in, D_ex=D_ex,
                                   seed=int(rng.integers(1, 2 ** 31)), T2_in=rng.uniform(50, 100),
                                   T2_ex=rng.uniform(40, 80),
                                   kappa=sb.kappa_for_probability(p, D_in, dt), rmin=0.5)
            run = abi.RunSpec(n_walkers=N, seed=1234, instance=k,
                              boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
            name = f"gamma theta={theta:.2f} f={frac:.2f} p={p:.1e}"
        out.append(Workload(f"c4_{k}", sub, acq, run, name))
    return out


# ------------------------------------------------------------ Bloch path
# fp64 flops of one rotation of one field (field value 18, |B|^2 5, angle 1,
# sine/cosine 45, projection 7, sine / |B| 1, C 3, B x M 9, update 15;
# FMA = 2) plus one square root and one reciprocal.
MAG_ROTATE_FLOPS = 104.0


@dataclass
class MagWorkload:
    name: str
    sub: abi.SubstrateSpec
    field: abi.FieldSpec
    run: abi.RunSpec
    description: str

    @property
    def walker_steps(self) -> int:
        return int(self.run.n_walkers) * (self.field.n_steps - 1)

    @property
    def field_updates(self) -> int:
        """Field updates of one walk: every walker, every field, t_0 .. t_{n-1}."""
        return int(self.run.n_walkers) * self.field.n_steps * self.field.n_fields


def mag_lf(n_walkers: int = 917504, n_fields: int = 2, integrator: int = abi.MAG_ROTATE) -> MagWorkload:
    """The longest acquisition of examples/magLF_spherNeuman.cu (mm, ms, T):
    sphere R = 2.5 um, D = 2.5e-6 mm^2/ms, dt = 1e-4 ms, walkers start
    inside with M = x; QSC_GRE fields with B0 = 0.5 mT, lobes 0.05 ms,
    spacing 0.01 ms (1100 steps), gradients 0 .. 6e-6 T/mm over n_fields
    fields (the reference's two: 0 and 6e-6).  Walkers: 8 x the reference's
    114,688 per acquisition."""
    from . import fields as fl
    R, D, dt = 0.0025, 2.5e-6, 1e-4
    L = 6.0 * R
    sub = sb.single_sphere(L, R, D_in=D, D_ex=D)
    d, s = 0.05, 0.01
    n = int((2.0 * d + s) / dt)
    fs = [fl.QSC_GRE(5e-4, 6e-6 * g / max(n_fields - 1, 1), d, s) for g in range(n_fields)]
    f = fl.build_field(fs, n, dt, integrator=integrator, origin=(L / 2,) * 3, m0=(1.0, 0.0, 0.0))
    run = abi.RunSpec(n_walkers=n_walkers, seed=11, init_mode=abi.INIT_INSIDE_PRIM)
    return MagWorkload("mag_lf", sub, f, run,
                       f"low-field QSC_GRE echo in a 2.5 um sphere, {n_fields} fields, {n} steps of 0.1 us, "
                       f"{n_walkers} walkers, integrator {integrator}")


def mag_lf_stream(n_acq: int = 1000, n_walkers: int = 114688):
    """Every acquisition of examples/magLF_spherNeuman.cu: acquisition i has
    lobes of (i + 1) 0.05 / n_acq ms, spacing 0.01 ms, two QSC_GRE fields
    (B0 = 0.5 mT, G = 0 and 6e-6 T/mm), 114,688 walkers in a 2.5 um sphere,
    dt = 1e-4 ms; seeds i * 11.  Returns [(substrate, field, run)]."""
    from . import fields as fl
    R, D, dt = 0.0025, 2.5e-6, 1e-4
    L = 6.0 * R
    sub = sb.single_sphere(L, R, D_in=D, D_ex=D)
    items = []
    for i in range(n_acq):
        d = (i + 1) * 0.05 / n_acq
        n = int((2.0 * d + 0.01) / dt)
        fs = [fl.QSC_GRE(5e-4, 0.0, d, 0.01), fl.QSC_GRE(5e-4, 6e-6, d, 0.01)]
        f = fl.build_field(fs, n, dt, origin=(L / 2,) * 3, m0=(1.0, 0.0, 0.0))
        items.append((sub, f, abi.RunSpec(n_walkers=n_walkers, seed=i * 11, instance=i,
                                          init_mode=abi.INIT_INSIDE_PRIM)))
    return items
