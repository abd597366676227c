"""The ported reference drivers (axonexa_amd/drivers, over
include/axonexa_compat.hpp) run on the GPU and give exactly what a direct
C-ABI run of the same experiment gives."""
import math
import os
import subprocess

import numpy as np
import pytest

from axonexa_amd import abi
from axonexa_amd import waveforms as wf

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "axonexa_amd", "bin")


def _run(name, *args):
    print(f"[gpu] driver {name} {args}", flush=True)
    out = subprocess.run([os.path.join(BIN, name), *map(str, args)], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    i = [k for k, l in enumerate(lines) if "Signals" in l][0]
    rows = np.array([[float(x) for x in l.split()] for l in lines[i + 1:] if l.strip()])
    return lines, rows


def test_cos_onecyl_equals_direct_run():
    N, frac = 4096, 0.05
    _, rows = _run("cos_onecyl", N, frac, 1234)
    assert rows.shape == (400, 4)
    # G = 0 waveforms: every phase is 0, Mx = N exactly
    assert np.all(rows[rows[:, 1] == 0.0][:, 2] == N)
    # the same experiment through the C ABI
    from axonexa_amd import engine
    dt, gamma = 0.001, 267500.0
    steps = max(2, int((2 * 20.0 + 2.0) / dt * frac))
    fs = [wf.CosGFunc(8.0 * i * 0.0000025 * (1 + 10 * j), 20.0, 2.0, 1 + 10 * j, (1, 0, 0))
          for j in range(20) for i in range(20)]
    acq = wf.build_acq(fs, steps, dt, gamma)
    R, f = .006, .75
    a = math.sqrt(math.pi * R * R / f)
    sub = abi.SubstrateSpec(a=a, b=a, c=a, D_lat=2.5e-6, T2_lat=200.0, types=[0],
                            centers=[[a / 2, a / 2, 0.0]], radii=[R], D=1e-6, T2=200.0, kappa=0.0)
    sig = engine.run_phase(sub, acq, abi.RunSpec(n_walkers=N, seed=1234))
    assert np.allclose(rows[:, 2], sig.sum_cos, rtol=1e-9, atol=1e-6)
    assert np.allclose(rows[:, 3], sig.sum_sin, rtol=1e-9, atol=1e-6)


def test_gammadist_runs_replicas():
    lines, rows = _run("gammadist", 2048, 0.02, 2)
    pf = float([l for l in lines if "Packing Fraction" in l][0].split("=")[1])
    assert 0.05 < pf < 0.9
    assert rows.shape == (400, 4)
    assert np.all(np.isfinite(rows))
    assert np.all(np.abs(rows[:, 2]) <= 2048 + 1e-6)


def test_wccompare_stream_equals_single_runs():
    print("[gpu] driver wccompare 2048 0.05 4", flush=True)
    out = subprocess.run([os.path.join(BIN, "wccompare"), "2048", "0.05", "4", "7"], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    rows = [l.split() for l in lines if l and l[0].isdigit()]
    assert [int(r[0]) for r in rows] == [2, 4]
    assert all(float(x) > 0 for r in rows for x in r[1:])
    assert lines[-1] == "max |stream - single| = 0", lines[-1]


def test_maglf_equals_direct_magnetization_runs():
    """The low-field driver (magAcquisitionStream over axx_run_mag) gives the
    ADC that direct runs of the same fields, seeds and pore give."""
    from axonexa_amd import engine
    from axonexa_amd import fields as fl
    from axonexa_amd import substrates as sb
    from axonexa_amd import analysis
    N, NOA, REP, SEED = 2048, 3, 2, 11
    print(f"[gpu] driver maglf {N} {NOA} {REP} {SEED}", flush=True)
    out = subprocess.run([os.path.join(BIN, "maglf"), str(N), str(NOA), str(REP), str(SEED)],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    rows = [[float(x) for x in l.split()] for l in out.stdout.splitlines() if l.strip() and "B0" not in l]
    assert len(rows) == NOA
    R, D, dt, B0 = 0.0025, 2.5e-6, 1e-4, 5e-4
    L = 6.0 * R
    sub = sb.single_sphere(L, R, D_in=D, D_ex=D)
    for i in range(NOA):
        d = (i + 1) * 0.05 / NOA
        n = int((2.0 * d + 0.01) / dt)
        fs = [fl.QSC_GRE(B0, 0.0, d, 0.01), fl.QSC_GRE(B0, 6e-6, d, 0.01)]
        f = fl.build_field(fs, n, dt, origin=(L / 2,) * 3, m0=(1.0, 0.0, 0.0))
        adcs = []
        for r in range(REP):
            seed = i * SEED if r == 0 else SEED * (r + 1) * (i + 1)
            m = engine.run_mag(sub, f, abi.RunSpec(n_walkers=N, seed=seed, instance=i,
                                                   init_mode=abi.INIT_INSIDE_PRIM))
            adcs.append(analysis.adc_fit([x.getB() for x in fs], np.hypot(m.mx, m.my)))
        assert abs(rows[i][0] - (0.01 + d)) < 1e-12
        # ln|S| moves by ~1e-9 between the two b-values, so the normal-equation
        # fit (n sum xy - sum x sum y) cancels ~12 digits: the two fits agree
        # to ~1e-5 relative although the walkers are the same
        assert np.isclose(rows[i][1], np.mean(adcs), rtol=1e-4, atol=0.0), (i, rows[i], adcs)
        assert np.isclose(rows[i][2], np.std(adcs), rtol=1e-2, atol=0.0), (i, rows[i], adcs)


def test_cos_onesphere_merges_periodic_copies():
    """The FCC driver lists 14 spheres (corners and face centres); the
    compat layer passes the four distinct ones, and the signals equal direct
    runs of those four."""
    from axonexa_amd import engine
    N, frac, seed = 4096, 0.02, 1234
    lines, rows = _run("cos_onesphere", N, frac, seed)
    assert rows.shape == (300, 3)
    R, f = .0015, .4
    a = ((16. / 3.) * math.pi * R ** 3 / f) ** (1. / 3.)
    centers = [[0, 0, 0], [0.5 * a, 0, 0.5 * a], [a, 0.5 * a, 0.5 * a], [0.5 * a, 0.5 * a, a]]
    sub = abi.SubstrateSpec(a=a, b=a, c=a, D_lat=2.5e-6, T2_lat=0.0, types=[abi.PRIM_SPHERE] * 4,
                            centers=centers, radii=[R] * 4, D=1e-6, T2=0.0, kappa=0.0)
    steps = max(2, int(42.0 / 0.001 * frac))
    for j in (0, 7):
        n_per = 1 + 10 * j
        fs = [wf.CosGFunc(i * 0.0000025 * n_per, 20.0, 2.0, n_per, (1, 0, 0)) for i in range(30)]
        acq = wf.build_acq(fs, steps, 0.001, 267500.0)
        sig = engine.run_phase(sub, acq, abi.RunSpec(n_walkers=N, seed=j * seed, instance=j,
                                                     init_mode=abi.INIT_COMPARTMENT, init_arg=0))
        want = np.hypot(sig.sum_cos, sig.sum_sin)
        got = rows[30 * j:30 * (j + 1), 2]
        assert np.allclose(got, want, rtol=1e-12, atol=1e-9), (j, np.max(np.abs(got - want)))
        assert sig.n_intra == 0 and sig.collisions > 0


def _blocks(lines, tag):
    """Numeric rows after each line containing tag."""
    out, cur = [], None
    for l in lines:
        if tag in l:
            cur = []
            out.append(cur)
        elif cur is not None and l.strip():
            cur.append([float(x) for x in l.split()])
    return [np.array(b) for b in out]


def test_hex_phase_equals_direct_runs():
    """Four listed cylinders, two distinct: the driver's signals equal direct
    runs of the two."""
    from axonexa_amd import engine
    N, frac, seed = 2048, 0.05, 1234
    print(f"[gpu] driver hex_phase {N} {frac} {seed}", flush=True)
    out = subprocess.run([os.path.join(BIN, "hex_phase"), str(N), str(frac), str(seed)], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    blocks = _blocks(out.stdout.splitlines(), "f = ")
    assert len(blocks) == 4 and all(b.shape == (6, 3) for b in blocks)
    s3 = 0.86602540378443864676372317075294
    R, f = .0025, .6
    d = math.sqrt(2.0 * math.pi * R * R / (math.sqrt(3.0) * f))
    sub = abi.SubstrateSpec(a=d, b=2.0 * d * s3, c=d, D_lat=2.5e-6, T2_lat=200.0, types=[0, 0],
                            centers=[[d / 2.0, 0.0, 0.0], [0.0, d * s3, 0.0]], radii=[R, R], D=1e-6, T2=200.0,
                            kappa=0.0)
    steps = max(2, int(22.0 / 0.001 * frac))
    for j in range(3):
        fs = [wf.CosGFunc(i * 0.0000025 * (1 + j), 10.0, 2.0, 1 + j, (1, 0, 0)) for i in range(2)]
        acq = wf.build_acq(fs, steps, 0.001, 267500.0)
        sig = engine.run_phase(sub, acq, abi.RunSpec(n_walkers=N, seed=j * seed, instance=j))
        want = np.hypot(sig.sum_cos, sig.sum_sin)
        assert np.allclose(blocks[0][2 * j:2 * j + 2, 2], want, rtol=1e-12, atol=1e-9), j


def test_plane_phase_equals_direct_runs():
    """A slab bounded by x = 0 and x = l (one plane under periodicity): the
    driver's replica-averaged ADC equals direct runs with the same seeds."""
    from axonexa_amd import engine
    from axonexa_amd import analysis
    N, frac, seed = 2048, 0.05, 1234
    print(f"[gpu] driver plane_phase {N} {frac} {seed}", flush=True)
    out = subprocess.run([os.path.join(BIN, "plane_phase"), str(N), str(frac), str(seed)], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    blocks = _blocks(out.stdout.splitlines(), "PLANE LENGTH")
    assert len(blocks) == 4 and all(b.shape == (10, 2) for b in blocks)
    # the first length (seeds carry over from one length to the next)
    l = .001
    sub = abi.SubstrateSpec(a=l, b=l, c=l, D_lat=2.5e-6, T2_lat=200.0, types=[abi.PRIM_PLANE],
                            centers=[[0.0, 0.0, 0.0]], radii=[0.0], D=0.0, T2=200.0, kappa=0.0,
                            normals=[[1.0, 0.0, 0.0]])
    steps = max(2, int(22.0 / 0.001 * frac))
    for j in (0, 6):
        fs = [wf.CosGFunc(i * 0.0000025 * (1 + j), 10.0, 2.0, 1 + j, (1, 0, 0)) for i in range(2)]
        acq = wf.build_acq(fs, steps, 0.001, 267500.0)
        adcs = []
        for r in range(3):
            sd = j * seed if r == 0 else seed * r * (j + 1)
            sig = engine.run_phase(sub, acq, abi.RunSpec(n_walkers=N, seed=sd, instance=j))
            adcs.append(analysis.adc_fit([x.getB(267500.0) for x in fs], np.hypot(sig.sum_cos, sig.sum_sin)))
        assert np.isclose(blocks[0][j, 1], np.mean(adcs), rtol=1e-6, atol=0.0), (j, blocks[0][j], adcs)
    # the slab restricts diffusion along x: the ADC of the narrowest slab at
    # the lowest frequency is well below D
    assert blocks[0][0, 1] < 0.5 * 2.5e-6
