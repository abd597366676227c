"""Magnetization (Bloch) path on the GPU (axx_mag_*, axx_run_mag) against the
CPU oracle (orc_walk_mag): every walker's position, counters and fp64
magnetization bitwise, sums to rounding, for each integrator, both kernel
geometries (cylinders only / spheres), all boundary policies, step
generators, several field groups, variable substeps and chunked walks; and
the B0 = 0 run against the phase path of the same walkers."""
import numpy as np
import pytest

from axonexa_amd import abi
from axonexa_amd import fields as fl
from axonexa_amd import substrates as sb
from axonexa_amd import waveforms as wf
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _engine():
    from axonexa_amd import engine
    return engine


def _check(label, sub, field, run, chunks=None, schedule=None):
    eng = _engine()
    print(f"[mag] {label}: N={run.n_walkers} steps={field.n_steps} fields={field.n_fields} "
          f"integ={field.integrator} schedule={schedule}", flush=True)
    with eng.Mag(sub, field, run) as m:
        if schedule is not None:
            m.set_schedule(schedule)
        if chunks:
            b = 0
            for e in chunks + [field.n_steps]:
                m.walk(b, e)
                b = e
        else:
            m.walk()
        got = m.results()
        for g in range(m.n_groups):
            gs = m.get_state(g)
            os_ = po.walk_mag(sub, field, run, group=g)
            same = np.array_equal(gs, os_)
            if not same:
                F = m.group_fields(g)
                a = gs.reshape(abi.mag_state_fields(F), run.n_walkers)
                b_ = os_.reshape(abi.mag_state_fields(F), run.n_walkers)
                bad = np.nonzero((a != b_).any(axis=0))[0]
                rows = np.nonzero((a != b_).any(axis=1))[0]
                raise AssertionError(f"{label} group {g}: {bad.size} walkers differ, fields {rows[:12]}")
    want = po.run_mag(sub, field, run)
    n = run.n_walkers
    for k in ("mx", "my", "mz"):
        assert np.allclose(getattr(got, k), getattr(want, k), rtol=1e-12, atol=1e-12 * n), (label, k)
    assert (got.collisions, got.transmissions, got.cap_hits) == (want.collisions, want.transmissions,
                                                                 want.cap_hits), label
    print(f"[mag] {label}: bitwise, collisions {got.collisions} transmissions {got.transmissions}", flush=True)
    return got


def test_rotation_hex_lattice_saddle_coil_fields():
    sub = sb.hex_lattice(0.0025, 0.006, 2, 1, D_in=1e-6, D_ex=2.5e-6, kappa=0.0)
    fs = [fl.QSC_GRE(5e-4, g * 2e-4, 0.01, 0.004) for g in range(3)]
    f = fl.build_field(fs, 240, 1e-4, origin=(0.003, 0.003, 0.0))
    run = abi.RunSpec(n_walkers=3000, seed=5)
    got = _check("rotate_hex", sub, f, run, chunks=[1, 77])
    assert got.collisions > 100


def test_rk4_relaxation_permeable_spheres():
    dt = 1e-4
    sub = sb.single_sphere(0.007, 0.0025, D_in=1e-6, D_ex=2e-6, T2_in=60.0, T2_ex=80.0,
                           kappa=sb.kappa_for_probability(0.05, 1e-6, dt))
    fs = [fl.COS_GRE(2e-4, (g * 3e-4, 0.0, g * 1e-4), 0.01, 0.003, 2) for g in range(2)]
    f = fl.build_field(fs, 230, dt, integrator=abi.MAG_RK4, substeps=np.full(230, 2, np.int32),
                       origin=(0.0035, 0.0035, 0.0035), T1=np.array([100.0, 70.0]))
    run = abi.RunSpec(n_walkers=2500, seed=6, boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
    got = _check("rk4_permeable_spheres", sub, f, run, chunks=[100])
    assert got.transmissions > 0


@pytest.mark.parametrize("smode", [abi.STEP_ISOTROPIC, abi.STEP_GAUSSIAN])
def test_dopri_single_pore_rejection_rf(smode):
    sub = sb.single_sphere(0.016, 0.0025, D_in=2.5e-6, D_ex=2.5e-6)
    fs = [fl.UCN_lowfield(5e-4, 3e-5, 0.006, 2e-5, 1e-4), fl.QSC_GRE(5e-4, 6e-6, 0.004, 0.002)]
    f = fl.build_field(fs, 120, 1e-4, integrator=abi.MAG_DOPRI, substeps="auto",
                       origin=(0.008, 0.008, 0.008), m0=(0.0, 0.0, 1.0))
    assert f.substeps.max() == 5 and f.substeps.min() == 1
    run = abi.RunSpec(n_walkers=2048, seed=7, boundary_mode=abi.BOUNDARY_REJECTION,
                      init_mode=abi.INIT_INSIDE_PRIM, step_mode=smode)
    _check(f"dopri_pore_{smode}", sub, f, run)


@pytest.mark.parametrize("integ,task_steps", [(abi.MAG_ROTATE, 3), (abi.MAG_RK4, 16), (abi.MAG_DOPRI, 7)])
def test_walk_tasks_of_a_few_steps(integ, task_steps):
    # every group's walk handed between waves every few steps (step 0 in
    # the first task), with chunked launches, permeable spheres and
    # relaxation: the fp64 magnetization of every walker bitwise
    dt = 1e-4
    sub = sb.single_sphere(0.007, 0.0025, D_in=1e-6, D_ex=2e-6, T2_in=60.0, T2_ex=80.0,
                           kappa=sb.kappa_for_probability(0.05, 1e-6, dt))
    fs = [fl.COS_GRE(2e-4, (g * 3e-4, 0.0, g * 1e-4), 0.01, 0.003, 2) for g in range(3)]
    f = fl.build_field(fs, 150, dt, integrator=integ, origin=(0.0035, 0.0035, 0.0035),
                       T1=np.array([100.0, 70.0]))
    run = abi.RunSpec(n_walkers=2000, seed=16, boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
    got = _check(f"tasks_{integ}_{task_steps}", sub, f, run, chunks=[1, 60], schedule=task_steps)
    assert got.transmissions > 0


def test_eleven_fields_two_groups_variable_substeps():
    sub = sb.single_sphere(0.01, 0.0025, D_in=2.5e-6, D_ex=2.5e-6)
    fs = [fl.QSC_GRE(5e-4, g * 4e-4, 0.02, 0.01) for g in range(11)]
    n = 500
    sub_n = np.where(np.arange(n) % 7 == 0, 3, 1).astype(np.int32)
    f = fl.build_field(fs, n, 1e-4, substeps=sub_n, origin=(0.005, 0.005, 0.005))
    run = abi.RunSpec(n_walkers=1500, seed=8, init_mode=abi.INIT_INSIDE_PRIM)
    got = _check("eleven_fields", sub, f, run)
    assert got.transverse[10] < got.transverse[0]


def test_zero_static_field_matches_phase_path():
    """B = (0, 0, G x g(t)): the magnetization sums are the phase path's
    (Mx = sum cos, My = -sum sin) for the same walkers."""
    eng = _engine()
    dt = 2.0 ** -10
    d, s = 16 * dt, 8 * dt
    n = 48
    sub = sb.hex_lattice(0.03, 0.07, 2, 1, D_in=2e-4, D_ex=4e-4)
    run = abi.RunSpec(n_walkers=20000, seed=9)
    Gx = 2e4
    f = fl.build_field([fl.RECT_GRE(0.0, (Gx, 0, 0), d, s)], n, dt, gamma=1.0)
    t = np.arange(n) * dt
    prof = np.where(t < d, 1.0, np.where((t >= s + d) & (t < 2 * d + s), -1.0, 0.0)).astype(np.float32)
    acq = abi.AcqSpec(n_steps=n, dt=dt, gamma=1.0, profiles=prof, wf_profile=[0], wf_amplitude=[Gx],
                      wf_dir=[[1.0, 0.0, 0.0]])
    mag = eng.run_mag(sub, f, run)
    ph = eng.run_phase(sub, acq, run)
    print(f"[mag] phase path {ph.sum_cos[0]:.4f} {ph.sum_sin[0]:.4f} mag {mag.mx[0]:.4f} {mag.my[0]:.4f}",
          flush=True)
    assert abs(mag.mx[0] - ph.sum_cos[0]) < 1e-4 * run.n_walkers
    assert abs(mag.my[0] + ph.sum_sin[0]) < 1e-4 * run.n_walkers
    assert ph.E[0] < 0.9


def test_invalid_field_descriptors_are_errors():
    eng = _engine()
    sub = sb.single_sphere(0.01, 0.0025, D_in=2.5e-6, D_ex=2.5e-6)
    run = abi.RunSpec(n_walkers=64, seed=1, init_mode=abi.INIT_INSIDE_PRIM)
    good = fl.build_field([fl.QSC_GRE(5e-4, 1e-4, 0.002, 0.001)], 60, 1e-4)
    bad = []
    f = fl.build_field([fl.QSC_GRE(5e-4, 1e-4, 0.002, 0.001)], 60, 1e-4)
    f.samples[3, 0, 5] = np.nan
    bad.append(f)
    f = fl.build_field([fl.QSC_GRE(5e-4, 1e-4, 0.002, 0.001)], 60, 1e-4, substeps=np.ones(60, np.int32))
    f.substeps[7] = 0
    bad.append(f)
    f = fl.build_field([fl.QSC_GRE(5e-4, 1e-4, 0.002, 0.001)], 60, 1e-4, integrator=abi.MAG_RK4,
                       T1=np.array([-1.0, 5.0]))
    bad.append(f)
    for k, f in enumerate(bad):
        with pytest.raises(eng.AxxError) as e:
            eng.run_mag(sub, f, run)
        assert e.value.code == -1, k
    m = eng.run_mag(sub, good, run)
    assert abs(np.hypot(m.mx[0], m.my[0]) / run.n_walkers - 1.0) < 0.05
    # a corrupt substrate descriptor is an error code, not a crash
    import ctypes as C
    ss, ff, rr = sub.struct(), good.struct(), run.struct()
    ss.n_prims = -3
    res = abi.MagResult()
    ctx = eng.default_context()
    assert eng.lib().axx_run_mag(ctx.handle, C.byref(ss), C.byref(ff), C.byref(rr), C.byref(res)) == -1


def test_batch_of_runs_equals_single_runs():
    """Acquisitions of a magAcquisitionStream in one launch: every run's
    walkers and sums are those of its own run_mag (and of the oracle)."""
    eng = _engine()
    items = []
    R = 0.0025
    for k in range(7):
        d = 0.004 * (k + 1)
        n = int((2 * d + 0.01) / 1e-4)
        L = 6 * R * (1.0 + 0.1 * k)
        sub = sb.single_sphere(L, R, D_in=2.5e-6, D_ex=2.5e-6) if k % 3 else \
            sb.hex_lattice(R, 0.006, 2, 1, D_in=1e-6, D_ex=2.5e-6)
        f = fl.build_field([fl.QSC_GRE(5e-4, 0.0, d, 0.01), fl.QSC_GRE(5e-4, 6e-5 * (k + 1), d, 0.01)], n, 1e-4,
                           origin=(L / 2,) * 3)
        run = abi.RunSpec(n_walkers=700 + 900 * k, seed=100 + k, instance=k,
                          init_mode=abi.INIT_INSIDE_PRIM if k % 3 else abi.INIT_UNIFORM,
                          boundary_mode=abi.BOUNDARY_REJECTION if k == 4 else abi.BOUNDARY_SPECULAR)
        items.append((sub, f, run))
    with eng.MagBatch(items) as b:
        b.run()
        for i, (sub, f, run) in enumerate(items):
            single = eng.run_mag(sub, f, run)
            got = b.results(i)
            assert np.array_equal(got.mx, single.mx) and np.array_equal(got.my, single.my), i
            assert np.array_equal(got.mz, single.mz) and got.collisions == single.collisions, i
            assert np.array_equal(b.get_state(i), po.walk_mag(sub, f, run)), i
        # a second run of the same batch gives the same bits
        b.run()
        assert np.array_equal(b.results(3).mx, eng.run_mag(*items[3]).mx)
    print(f"[mag] batch of {len(items)} runs bitwise equal to single runs and the oracle", flush=True)
    # mixed field counts are refused
    f3 = fl.build_field([fl.QSC_GRE(5e-4, 0.0, 0.004, 0.01)] * 3, 180, 1e-4)
    with pytest.raises(eng.AxxError):
        eng.MagBatch([items[0], (items[0][0], f3, items[0][2])])


@pytest.mark.parametrize("integ", [abi.MAG_ROTATE, abi.MAG_RK4, abi.MAG_DOPRI])
def test_eight_fields_single_and_batched(integ):
    """Eight fields in one group: the widest kernels, which run under a
    register budget with spills (single: four waves per SIMD, batched: three),
    keep every walker bitwise equal to the oracle."""
    eng = _engine()
    R = 0.0025
    L = 6 * R
    sub = sb.single_sphere(L, R, D_in=2.5e-6, D_ex=2.5e-6, kappa=1e-2)
    items = []
    for k in range(2):
        d = 0.003 * (k + 1)
        n = int((2 * d + 0.002) / 1e-4)
        fs = [fl.QSC_GRE(5e-4, 2e-5 * g, d, 0.002) for g in range(8)]
        f = fl.build_field(fs, n, 1e-4, integrator=integ, origin=(L / 2,) * 3, m0=(1.0, 0.0, 0.0),
                           substeps=None if integ == abi.MAG_ROTATE else np.full(n, 2, np.int32))
        run = abi.RunSpec(n_walkers=1500 + 700 * k, seed=40 + k, instance=k,
                          init_mode=abi.INIT_INSIDE_PRIM, boundary_mode=abi.BOUNDARY_PERMEABLE)
        items.append((sub, f, run))
    _check(f"eight_fields_integ{integ}", *items[0], chunks=[37])
    with eng.MagBatch(items) as b:
        b.run()
        for i, (s_, f, run) in enumerate(items):
            assert np.array_equal(b.get_state(i), po.walk_mag(s_, f, run)), (integ, i)
    print(f"[mag] eight fields, integrator {integ}: single and batched bitwise", flush=True)


def test_batch_with_mixed_step_generators():
    """A magnetization batch whose runs differ in step generator (isotropic,
    axis-sign, Gaussian) runs the per-run switching kernel; every run is still
    bitwise its oracle walk."""
    eng = _engine()
    R = 0.0025
    L = 6 * R
    sub = sb.single_sphere(L, R, D_in=2.5e-6, D_ex=2.5e-6)
    items = []
    for k, smode in enumerate((abi.STEP_ISOTROPIC, abi.STEP_AXES, abi.STEP_GAUSSIAN)):
        d = 0.003 * (k + 1)
        n = int((2 * d + 0.002) / 1e-4)
        f = fl.build_field([fl.QSC_GRE(5e-4, 0.0, d, 0.002), fl.QSC_GRE(5e-4, 3e-5, d, 0.002)], n, 1e-4,
                           origin=(L / 2,) * 3)
        run = abi.RunSpec(n_walkers=900 + 300 * k, seed=70 + k, instance=k, init_mode=abi.INIT_INSIDE_PRIM,
                          step_mode=smode)
        items.append((sub, f, run))
    with eng.MagBatch(items) as b:
        b.run()
        for i, (s_, f, run) in enumerate(items):
            assert np.array_equal(b.get_state(i), po.walk_mag(s_, f, run)), i
            assert np.array_equal(b.results(i).mx, eng.run_mag(s_, f, run).mx), i
    print("[mag] batch with mixed step generators: bitwise", flush=True)


@pytest.mark.parametrize("integ", [abi.MAG_ROTATE, abi.MAG_RK4, abi.MAG_DOPRI])
@pytest.mark.parametrize("cls", [0, 2, 3])
def test_uniform_batch_runs_its_class_kernel(integ, cls):
    """A magnetization batch whose runs share one class (bit 0: spheres or
    planes, bit 1: rejection) and take isotropic steps runs the kernel
    compiled for that class alone; classes 0 (cylinders), 2 (cylinders,
    rejection) and 3 (spheres, rejection) with two fields per run, every
    walker bitwise its oracle walk."""
    eng = _engine()
    R = 0.0025
    if cls & 1:
        L = 6 * R
        sub = sb.single_sphere(L, R, D_in=2.5e-6, D_ex=2.5e-6)
        origin, init = (L / 2,) * 3, abi.INIT_INSIDE_PRIM
    else:
        sub = sb.hex_lattice(R, 0.006, 2, 1, D_in=1e-6, D_ex=2.5e-6, kappa=0.0 if cls & 2 else 1e-2)
        origin, init = (sub.a / 2, sub.b / 2, 0.0), abi.INIT_UNIFORM
    boundary = abi.BOUNDARY_REJECTION if cls & 2 else (abi.BOUNDARY_SPECULAR if cls & 1 else abi.BOUNDARY_PERMEABLE)
    items = []
    for k in range(2):
        d = 0.003 * (k + 1)
        n = int((2 * d + 0.002) / 1e-4)
        f = fl.build_field([fl.QSC_GRE(5e-4, 0.0, d, 0.002), fl.QSC_GRE(5e-4, 4e-5 * (k + 1), d, 0.002)], n, 1e-4,
                           integrator=integ, origin=origin,
                           substeps=None if integ == abi.MAG_ROTATE else np.full(n, 2, np.int32))
        run = abi.RunSpec(n_walkers=800 + 500 * k, seed=90 + k, instance=k, init_mode=init, boundary_mode=boundary)
        items.append((sub, f, run))
    print(f"[mag] uniform batch class {cls}, integrator {integ}", flush=True)
    with eng.MagBatch(items) as b:
        b.run()
        for i, (s_, f, run) in enumerate(items):
            assert np.array_equal(b.get_state(i), po.walk_mag(s_, f, run)), (cls, integ, i)
    print(f"[mag] uniform batch class {cls}, integrator {integ}: bitwise", flush=True)
