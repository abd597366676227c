"""Batched persistent launch (axx_batch_*) vs the CPU oracle and vs single
runs.

Every instance of a batch keeps its own seed, instance id and walker ids, so
its walkers must be bitwise those of the oracle (and of its own single run):
the batch differs only in scheduling (tiles taken from a work counter by a
persistent grid) and in keeping moments for every axis (the oracle keeps the
axes in use; the others are masked before comparing).
"""
import numpy as np
import pytest

from axonexa_amd import abi
from axonexa_amd import waveforms as wf
from axonexa_amd import workloads as W
from oracle import pyoracle as po
from tests.helpers import compare_states, mask_moments

pytestmark = pytest.mark.gpu


def _engine():
    from axonexa_amd import engine
    return engine


def _check_instance(b, i, w, label):
    gsig = b.results(i)
    gst = b.get_state(i)
    ost = po.walk(w.sub, w.acq, w.run)
    osig = po.reduce(w.sub, w.acq, w.run, ost)
    n, P = w.run.n_walkers, w.acq.n_profiles
    info = compare_states(mask_moments(gst, n, P, po.axes_mask(w.acq)), ost, n, P)
    assert info["bitwise_fraction"] == 1.0, (label, info)
    for k in ("sum_cos", "sum_sin", "sum_cos2", "sum_sin2", "intra_cos", "intra_sin"):
        assert np.allclose(getattr(gsig, k), getattr(osig, k), rtol=1e-12, atol=1e-9 * n), (label, k)
    assert (gsig.collisions, gsig.transmissions, gsig.cap_hits, gsig.n_intra) == \
        (osig.collisions, osig.transmissions, osig.cap_hits, osig.n_intra), label
    return gsig


def test_batch_of_256_instances_matches_oracle():
    ws = W.c4_batch(scale=0.01, n_steps=300)
    print(f"[batch] 256 instances, {sum(w.run.n_walkers for w in ws)} walkers x 300 steps", flush=True)
    eng = _engine()
    with eng.Batch([(w.sub, w.acq, w.run) for w in ws]) as b:
        print(f"[batch] config {b.config()}", flush=True)
        b.run()
        coll = 0
        for i, w in enumerate(ws):
            coll += _check_instance(b, i, w, f"{w.name} {w.description}").collisions
    assert coll > 10000


def test_batch_equals_single_runs_and_graph_replays():
    ws = W.c4_batch(n_inst=12, scale=0.02, n_steps=500, seed=11)
    eng = _engine()
    single = [eng.run_phase(w.sub, w.acq, w.run) for w in ws]
    print("[batch] 12 instances: single runs done", flush=True)
    with eng.Batch([(w.sub, w.acq, w.run) for w in ws]) as b:
        for rep in range(3):
            b.run()
            for i, w in enumerate(ws):
                g = b.results(i)
                assert np.array_equal(g.sum_cos, single[i].sum_cos), (rep, i)
                assert np.array_equal(g.sum_sin, single[i].sum_sin), (rep, i)
                assert g.collisions == single[i].collisions
            # eager work between graph replays
            eng.run_phase(ws[0].sub, ws[0].acq, ws[0].run)
            print(f"[batch] replay {rep} identical", flush=True)


def test_batch_autotune_keeps_results():
    ws = W.c4_batch(n_inst=16, scale=0.02, n_steps=400, seed=5)
    eng = _engine()
    with eng.Batch([(w.sub, w.acq, w.run) for w in ws]) as b:
        b.run()
        before = [b.results(i).sum_cos for i in range(len(ws))]
        cfg = b.autotune(probe_steps=50)
        print(f"[batch] autotune {cfg} -> {b.config()}", flush=True)
        assert cfg["block"] in (256, 512) and cfg["wpt"] in (1, 2, 4)
        b.run()
        for i in range(len(ws)):
            assert np.array_equal(b.results(i).sum_cos, before[i]), i
        for blk, wpt in ((64, 1), (512, 3)):
            b.set_config(blk, wpt)
            b.run()
            for i in range(len(ws)):
                assert np.array_equal(b.results(i).sum_cos, before[i]), (blk, wpt, i)


def test_batch_mixes_policies_and_step_generators():
    """Rejection, reflecting and permeable instances with isotropic, axis-sign
    and Gaussian steps in one launch (the streaming acquisitions of a WC
    comparison), each bitwise its oracle walk."""
    from axonexa_amd import substrates as sb
    dt = 2e-3
    hexl = sb.hex_lattice(2.5, 6.0, 3, 2, D_in=1.7, D_ex=2.0, T2_in=70.0, T2_ex=50.0,
                          kappa=sb.kappa_for_probability(0.05, 1.7, dt))
    sph = sb.single_sphere(16.0, 5.0, D_in=2.0, D_ex=2.0)
    acq = wf.build_acq([wf.CosGFunc(g * 1e-6, 0.4, 0.1, 2.0, ax)
                        for g in (0.1, 0.4) for ax in ((1, 0, 0), (0, 0, 1))], 300, dt, W.GAMMA)

    class Case:
        pass

    ws = []
    for k, (sub, bmode, smode, imode) in enumerate([
            (hexl, abi.BOUNDARY_PERMEABLE, abi.STEP_ISOTROPIC, abi.INIT_UNIFORM),
            (sph, abi.BOUNDARY_REJECTION, abi.STEP_GAUSSIAN, abi.INIT_INSIDE_PRIM),
            (hexl, abi.BOUNDARY_SPECULAR, abi.STEP_AXES, abi.INIT_UNIFORM),
            (sph, abi.BOUNDARY_REJECTION, abi.STEP_ISOTROPIC, abi.INIT_INSIDE_PRIM),
            (hexl, abi.BOUNDARY_REJECTION, abi.STEP_AXES, abi.INIT_UNIFORM),
            (sph, abi.BOUNDARY_SPECULAR, abi.STEP_GAUSSIAN, abi.INIT_INSIDE_PRIM)]):
        c = Case()
        c.sub, c.acq = sub, acq
        c.run = abi.RunSpec(n_walkers=700 + 211 * k, seed=40 + k, instance=k, boundary_mode=bmode,
                            relaxation=1, step_mode=smode, init_mode=imode)
        ws.append(c)
    eng = _engine()
    with eng.Batch([(w.sub, w.acq, w.run) for w in ws]) as b:
        print(f"[batch] mixed policies, config {b.config()}", flush=True)
        b.run()
        for i, w in enumerate(ws):
            g = _check_instance(b, i, w, f"mixed {i}")
            assert g.collisions > 0, i


def test_batch_rejects_what_it_cannot_run():
    ws = W.c4_batch(n_inst=2, scale=0.01, n_steps=100, seed=3)
    eng = _engine()
    w = ws[1]
    # 33 distinct profiles: above the batch limit of 32
    fs = [wf.CosGFunc(1e-7, 0.08, 0.02, 1.0 + k, (1, 0, 0)) for k in range(33)]
    wide = wf.build_acq(fs, 100, 2e-3, W.GAMMA)
    assert wide.n_profiles == 33 > abi.BATCH_MAX_PROFILES
    with pytest.raises(eng.AxxError) as e:
        eng.Batch([(w.sub, wide, w.run)])
    assert e.value.code == -4


def _ogse_acq(n_freq, n_steps, dt, amps=(0.1, 0.4), axis=(1, 0, 0)):
    """Cosine OGSE at n_freq frequencies (1..n_freq periods per lobe) x amps:
    n_freq distinct profiles."""
    return wf.build_acq([wf.CosGFunc(g * 1e-6, 0.4, 0.1, float(f), axis) for f in range(1, n_freq + 1)
                         for g in amps], n_steps, dt, W.GAMMA)


@pytest.mark.parametrize("width", [1, 3, 8, 16, 32])
@pytest.mark.parametrize("kind", ["cylinders", "spheres_rejection", "spheres", "cylinders_rejection"])
def test_uniform_batch_runs_its_class_kernel(kind, width):
    """Instances that all share one class (profile width, geometry, boundary
    policy) with isotropic steps run the kernel fixed to that class (widths
    1, 4, 8) or the wide kernel (16, 32); each is still bitwise its oracle
    walk."""
    from axonexa_amd import substrates as sb
    dt = 2e-3
    acq = _ogse_acq(width, 300, dt)

    class Case:
        pass

    ws = []
    for k in range(5):
        c = Case()
        if kind.startswith("cylinders"):
            c.sub = sb.hex_lattice(1.5 + 0.5 * k, 4.0 + 1.2 * k, 3, 2, D_in=1.7, D_ex=2.0, T2_in=70.0,
                                   T2_ex=50.0, kappa=sb.kappa_for_probability(0.1, 1.7, dt))
            mode = abi.BOUNDARY_REJECTION if kind.endswith("rejection") else abi.BOUNDARY_PERMEABLE
            init = abi.INIT_UNIFORM
        else:
            c.sub = sb.single_sphere(12.0 + k, 3.0 + 0.5 * k, D_in=2.0, D_ex=2.0)
            mode = abi.BOUNDARY_REJECTION if kind.endswith("rejection") else abi.BOUNDARY_SPECULAR
            init = abi.INIT_INSIDE_PRIM
        c.acq = acq
        c.run = abi.RunSpec(n_walkers=600 + 173 * k, seed=60 + k, instance=k, boundary_mode=mode,
                            relaxation=1, init_mode=init)
        ws.append(c)
    eng = _engine()
    with eng.Batch([(w.sub, w.acq, w.run) for w in ws]) as b:
        b.run()
        for i, w in enumerate(ws):
            g = _check_instance(b, i, w, f"{kind} width {width} {i}")
            assert g.collisions > 0, i
    print(f"[batch] uniform {kind} width {width}: bitwise", flush=True)


def test_batch_mixes_narrow_and_wide_profile_sets():
    """32-frequency and 16-frequency cosine-OGSE instances (the wide kernel)
    with PGSE and 8-frequency ones (the narrow kernel) in one graph, each
    bitwise its oracle walk; axis-sign steps on one wide instance put both
    kernels on their step-generator instantiation."""
    from axonexa_amd import substrates as sb
    dt = 2e-3

    class Case:
        pass

    specs = [(32, abi.STEP_ISOTROPIC), (1, abi.STEP_ISOTROPIC), (16, abi.STEP_AXES), (8, abi.STEP_ISOTROPIC),
             (32, abi.STEP_ISOTROPIC), (20, abi.STEP_ISOTROPIC)]
    ws = []
    for k, (nf, smode) in enumerate(specs):
        c = Case()
        if k % 2 == 0:
            c.sub = sb.hex_lattice(2.0 + 0.3 * k, 5.0 + 0.5 * k, 3, 2, D_in=1.7, D_ex=2.0, T2_in=70.0,
                                   T2_ex=50.0, kappa=sb.kappa_for_probability(0.05, 1.7, dt))
            mode, init = abi.BOUNDARY_PERMEABLE, abi.INIT_UNIFORM
        else:
            c.sub = sb.single_sphere(14.0, 4.0 + 0.3 * k, D_in=2.0, D_ex=2.0)
            mode, init = abi.BOUNDARY_SPECULAR, abi.INIT_INSIDE_PRIM
        c.acq = (_ogse_acq(nf, 400, dt, amps=(0.2,), axis=(1, 1, 0)) if nf > 1 else
                 wf.build_acq([wf.RectGrad(g * 1e-6, 0.3, 0.1, (0, 0, 1)) for g in (0.1, 0.2, 0.3, 0.4)],
                              400, dt, W.GAMMA))
        assert c.acq.n_profiles == nf
        c.run = abi.RunSpec(n_walkers=900 + 157 * k, seed=70 + k, instance=k, boundary_mode=mode,
                            relaxation=1, init_mode=init, step_mode=smode)
        ws.append(c)
    eng = _engine()
    with eng.Batch([(w.sub, w.acq, w.run) for w in ws]) as b:
        print(f"[batch] narrow + wide profile sets, config {b.config()}", flush=True)
        for rep in range(2):
            b.run()
            for i, w in enumerate(ws):
                g = _check_instance(b, i, w, f"wide mix {i} (rep {rep})")
                assert g.collisions > 0, i


def test_tuning_cache_file_reloads_a_long_signature(tmp_path):
    """A 300-instance batch (signature of ~6 KB) tuned with AXX_TUNE_CACHE set
    is found again from the file by a fresh process."""
    import json
    import os
    import subprocess
    import sys
    path = tmp_path / "tune.txt"
    code = r"""
import json, os, sys
sys.path.insert(0, os.environ["AXX_ROOT"])
from axonexa_amd import abi, engine
from axonexa_amd import substrates as sb
from axonexa_amd import waveforms as wf
acq = wf.build_acq([wf.RectGrad(1e-7, 0.02, 0.01)], 30, 1e-3, 2.675e5)
items = [(sb.single_cylinder(10.0, 1.0 + 0.01 * k, D_in=2.0, D_ex=2.0), acq,
          abi.RunSpec(n_walkers=64 + k, seed=k, instance=k)) for k in range(300)]
with engine.Batch(items) as b:
    if sys.argv[1] == "tune":
        b.autotune(probe_steps=5)
    print(json.dumps(b.config()))
"""
    env = dict(os.environ, AXX_TUNE_CACHE=str(path), AXX_ROOT=os.path.dirname(os.path.dirname(__file__)))
    run = lambda mode: subprocess.run([sys.executable, "-c", code, mode], env=env, capture_output=True, text=True,
                                      timeout=300)
    r = run("tune")
    assert r.returncode == 0, r.stderr[-2000:]
    lines = path.read_text().split("\n")
    key, blk, wpt = lines[0].split()
    assert len(key) == 16 and len(lines[0]) < 64
    # a configuration the default never takes, so the reload is visible
    path.write_text(f"{key} 512 3\n")
    r = run("show")
    assert r.returncode == 0, r.stderr[-2000:]
    cfg = json.loads(r.stdout.strip().splitlines()[-1])
    assert (cfg["block"], cfg["wpt"]) == (512, 3), cfg
