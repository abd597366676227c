"""HIP walk vs CPU oracle, walker by walker (the parity gate).

Same seeded inputs on both sides; integer fields (compartment, collision,
transmission and bounce-cap counts) must be identical, float fields within
1e-5 relative; signal sums within 1e-12 relative.  Cases cover the
reference's boundary policies (kernelLattice.cuh:64-79,
boundaryCheck.cuh:34-64, :66-113), initial placements (kernelLattice.cuh:29-37),
cylinders and spheres, relaxation, chunked (checkpointed) walks.
"""
import math
import sys

import numpy as np
import pytest

from axonexa_amd import abi
from axonexa_amd import substrates as sb
from axonexa_amd import waveforms as wf
from oracle import pyoracle as po
from tests.helpers import compare_states

pytestmark = pytest.mark.gpu

GAMMA = 2.675e5


def _engine():
    from axonexa_amd import engine
    return engine


def _pgse(n_steps, dt, Gs=(0.0, 0.1, 0.3), d=None, gap=None, axis=(1, 0, 0)):
    T = n_steps * dt
    d = d if d is not None else 0.3 * T
    gap = gap if gap is not None else 0.3 * T
    return wf.build_acq([wf.RectGrad(g * 1e-6, d, gap, axis) for g in Gs], n_steps, dt, GAMMA)


def _run_both(name, sub, acq, run, chunks=None, schedule=None):
    print(f"[parity] {name}: N={run.n_walkers} steps={acq.n_steps} schedule={schedule}", flush=True)
    eng = _engine()
    with eng.Plan(sub, acq, run) as plan:
        if schedule is not None:
            plan.set_schedule(schedule)
        if chunks:
            prev = 0
            for c in list(chunks) + [acq.n_steps]:
                plan.walk(prev, c)
                prev = c
        else:
            plan.walk()
        plan.reduce()
        gsig = plan.results()
        gst = plan.get_state()
        if schedule is not None and schedule > 0:
            assert plan.schedule()["last_task_steps"] == min(schedule + (schedule & 1), acq.n_steps - 1)
    ost = po.walk(sub, acq, run)
    osig = po.reduce(sub, acq, run, ost)
    info = compare_states(gst, ost, run.n_walkers, acq.n_profiles)
    for k in ("sum_cos", "sum_sin", "sum_cos2", "sum_sin2", "intra_cos", "intra_sin"):
        a, b = getattr(gsig, k), getattr(osig, k)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-9 * run.n_walkers), k
    assert gsig.collisions == osig.collisions
    assert gsig.transmissions == osig.transmissions
    assert gsig.cap_hits == osig.cap_hits
    assert gsig.n_intra == osig.n_intra
    print(f"[parity] {name}: bitwise {info['bitwise_fraction']:.4f} max_rel {info['max_rel']:.2e} "
          f"collisions {gsig.collisions} transmissions {gsig.transmissions}", flush=True)
    return gsig, info


def test_free_box():
    sub = sb.empty_box(20.0, 2.0)
    acq = _pgse(2000, 1e-3)
    _run_both("free_box", sub, acq, abi.RunSpec(n_walkers=4096, seed=1234))


def test_cylinder_specular():
    sub = sb.single_cylinder(20.0, 4.0, D_in=2.0, D_ex=2.0)
    acq = _pgse(2000, 1e-2)
    sig, info = _run_both("cylinder_specular", sub, acq, abi.RunSpec(n_walkers=4096, seed=1234))
    assert sig.collisions > 1000


def test_cylinder_ragged_n_and_offset():
    sub = sb.single_cylinder(20.0, 4.0, D_in=1.7, D_ex=2.0)
    acq = _pgse(777, 1e-2, axis=(0.6, 0.8, 0.0))
    _run_both("cylinder_ragged", sub, acq,
              abi.RunSpec(n_walkers=1001, seed=99, walker_offset=123456, instance=7))


def test_hex_permeable_relaxation():
    dt = 2e-3
    sub = sb.hex_lattice(2.5, 6.0, 3, 2, D_in=1.7, D_ex=2.0, T2_in=70.0, T2_ex=50.0,
                         kappa=sb.kappa_for_probability(0.05, 1.7, dt))
    acq = wf.build_acq([wf.CosGFunc(g * 1e-6, 3.0, 0.5, 2.5, ax)
                        for g in (0.1, 0.4) for ax in ((1, 0, 0), (0, 0, 1))], 3500, dt, GAMMA)
    run = abi.RunSpec(n_walkers=4096, seed=1234, boundary_mode=abi.BOUNDARY_PERMEABLE,
                      relaxation=1)
    sig, _ = _run_both("hex_permeable", sub, acq, run)
    assert sig.transmissions > 100


def test_spheres_specular_3d():
    # 2 x 2 x 2 simple-cubic sphere lattice
    cs = [[2.5 + 5 * i, 2.5 + 5 * j, 2.5 + 5 * k] for i in range(2) for j in range(2) for k in range(2)]
    sub = abi.SubstrateSpec(a=10.0, b=10.0, c=10.0, D_lat=2.0, types=[abi.PRIM_SPHERE] * 8,
                            centers=cs, radii=[2.0] * 8, D=2.0, T2=np.inf, kappa=0.0)
    acq = _pgse(1500, 1e-2, axis=(0, 0, 1))
    sig, _ = _run_both("spheres", sub, acq, abi.RunSpec(n_walkers=2048, seed=5))
    assert sig.collisions > 1000


def test_sphere_permeable_inside_init():
    sub = sb.single_sphere(16.0, 5.0, D_in=1.0, D_ex=2.5, T2_in=40.0, T2_ex=80.0,
                           kappa=sb.kappa_for_probability(0.1, 1.0, 1e-2))
    acq = _pgse(1200, 1e-2, axis=(1, 1, 0))
    run = abi.RunSpec(n_walkers=2048, seed=11, init_mode=abi.INIT_INSIDE_PRIM, init_arg=0,
                      boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
    sig, _ = _run_both("sphere_perm_inside", sub, acq, run)
    assert sig.n_intra == run.n_walkers


def test_rejection_mode_compartment_init():
    sub = sb.single_cylinder(12.0, 3.0, D_in=2.0, D_ex=2.0)
    acq = _pgse(1000, 1e-2)
    run = abi.RunSpec(n_walkers=2048, seed=3, init_mode=abi.INIT_COMPARTMENT, init_arg=0,
                      boundary_mode=abi.BOUNDARY_REJECTION)
    sig, _ = _run_both("rejection", sub, acq, run)
    assert sig.n_intra == 0 and sig.collisions > 0


def test_chunked_walk_equals_oracle():
    sub = sb.single_cylinder(20.0, 4.0, D_in=2.0, D_ex=2.0)
    acq = _pgse(1000, 1e-2)
    _run_both("chunked", sub, acq, abi.RunSpec(n_walkers=2048, seed=8), chunks=[1, 2, 333, 600])


def test_chunked_permeable_walk_equals_oracle():
    # one Philox call serves four boundary events of a walker, numbered over
    # the whole walk: launches that start inside a block of four must resume
    # on the same words
    sub = sb.single_cylinder(20.0, 4.0, D_in=2.0, D_ex=2.0,
                             kappa=sb.kappa_for_probability(0.2, 2.0, 1e-2))
    acq = _pgse(1000, 1e-2)
    run = abi.RunSpec(n_walkers=2048, seed=9, boundary_mode=abi.BOUNDARY_PERMEABLE)
    sig, _ = _run_both("chunked_permeable", sub, acq, run, chunks=[1, 2, 3, 251, 252, 600])
    assert sig.transmissions > 1000


def test_bounce_cap_counted():
    sub = sb.single_cylinder(20.0, 4.0, D_in=2.0, D_ex=2.0)
    acq = _pgse(800, 1e-2)
    sig, _ = _run_both("bounce_cap", sub, acq, abi.RunSpec(n_walkers=2048, seed=2, bounce_cap=1))
    assert sig.cap_hits > 0


def test_planes_slab_with_permeable_sphere():
    # planes x = 4 and x = 12 (normals into the slab) around a permeable
    # sphere; planes reflect, never transmit and draw no collision numbers
    a = 16.0
    sub = abi.SubstrateSpec(a=a, b=a, c=a, D_lat=2.0, T2_lat=60.0,
                            types=[abi.PRIM_PLANE, abi.PRIM_PLANE, abi.PRIM_SPHERE],
                            centers=[[4.0, 0, 0], [12.0, 0, 0], [8.0, 8.0, 8.0]],
                            radii=[0.0, 0.0, 2.5], D=[0.0, 0.0, 1.0], T2=[np.inf, np.inf, 30.0],
                            kappa=[0.0, 0.0, sb.kappa_for_probability(0.05, 1.0, 1e-2)],
                            normals=[[1, 0, 0], [-1, 0, 0], [0, 0, 0]])
    acq = _pgse(1500, 1e-2, axis=(1, 0, 0))
    run = abi.RunSpec(n_walkers=2048, seed=17, boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
    sig, _ = _run_both("planes_slab_sphere", sub, acq, run)
    assert sig.collisions > 1000 and sig.transmissions > 10


def test_planes_at_cell_faces():
    # the reference's plane validation geometry (planePhase.cu:114-117):
    # planes at x = 0 and x = l, a cell of side l -- one plane twice under
    # periodicity; after a reflection the walker moves away from both copies,
    # so they act as one plane
    l = 10.0
    sub = abi.SubstrateSpec(a=l, b=l, c=l, D_lat=2.5, types=[abi.PRIM_PLANE] * 2,
                            centers=[[0.0, 0, 0], [l, 0, 0]], radii=[0.0, 0.0],
                            normals=[[1, 0, 0], [-1, 0, 0]])
    acq = wf.build_acq([wf.CosGFunc(g * 1e-6, 10.0, 2.0, 3, (1, 0, 0)) for g in (0.0, 0.05, 0.1)],
                       2200, 1e-2, GAMMA)
    sig, _ = _run_both("planes_cell_faces", sub, acq, abi.RunSpec(n_walkers=2048, seed=23))
    assert sig.collisions > 1000


def test_oblique_plane_with_cylinders():
    sub = sb.hex_lattice(2.0, 6.0, 2, 1, D_in=1.5, D_ex=2.0)
    sub = abi.SubstrateSpec(a=sub.a, b=sub.b, c=sub.c, D_lat=2.0,
                            types=list(sub.types) + [abi.PRIM_PLANE],
                            centers=np.vstack([sub.centers, [[6.0, 5.0, 1.0]]]),
                            radii=list(sub.radii) + [0.0], D=list(sub.D) + [0.0],
                            normals=np.vstack([np.zeros((sub.n_prims, 3)), [[1.0, 2.0, 0.5]]]))
    acq = _pgse(1500, 1e-2, axis=(0.6, 0.0, 0.8))
    sig, _ = _run_both("oblique_plane", sub, acq, abi.RunSpec(n_walkers=2048, seed=31))
    assert sig.collisions > 1000


def test_planes_rejection_mode():
    sub = sb.parallel_planes(10.0, 2.5, 7.5, 2.0)
    acq = _pgse(1200, 1e-2)
    sig, _ = _run_both("planes_rejection", sub, acq,
                       abi.RunSpec(n_walkers=2048, seed=41, boundary_mode=abi.BOUNDARY_REJECTION))
    assert sig.collisions > 100


@pytest.mark.parametrize("mode", [abi.BOUNDARY_SPECULAR, abi.BOUNDARY_PERMEABLE])
def test_wide_profile_set_gamma_packing(mode):
    # 64 profiles x 3 axes: the 512-thread, 256-VGPR instantiation (which
    # recomputes the transmission draws per event instead of keeping them)
    from axonexa_amd import workloads as W
    sub = sb.gamma_packing(30, 5.0, 0.4, 0.6, D_in=1.7, D_ex=2.0, seed=4, T2_in=70.0, T2_ex=50.0,
                           rmin=0.25, kappa=sb.kappa_for_probability(0.3, 1.7, 5e-3))
    fs = W.c3_functors(n_dirs=2, amps=(0.3,))
    acq = wf.build_acq(fs, 600, 5e-3, GAMMA)
    assert acq.n_profiles == 64
    run = abi.RunSpec(n_walkers=1500, seed=77, relaxation=1, boundary_mode=mode)
    sig, _ = _run_both(f"wide_profiles_{mode}", sub, acq, run)
    assert sig.collisions > 1000
    assert (sig.transmissions > 1000) == (mode == abi.BOUNDARY_PERMEABLE)


def test_dense_permeable_packing_many_events_per_step():
    # steps of 0.49 um against radii from 0.25 um with p = 0.6: several
    # boundary events per step, multi-candidate sub-bins, and transmission
    # draws that run through many blocks of four per walker
    dt = 2e-2
    sub = sb.gamma_packing(30, 5.0, 0.4, 0.6, D_in=1.7, D_ex=2.0, seed=7, T2_in=70.0, T2_ex=50.0,
                           rmin=0.25, kappa=sb.kappa_for_probability(0.6, 1.7, dt))
    acq = wf.build_acq([wf.RectGrad(g * 1e-6, 3.0, 2.0, (1, 0, 0)) for g in (0.1, 0.3)], 500, dt, GAMMA)
    run = abi.RunSpec(n_walkers=3000, seed=21, boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
    sig, _ = _run_both("dense_permeable", sub, acq, run)
    assert sig.collisions > 100000 and sig.transmissions > 50000


@pytest.mark.parametrize("task_steps", [2, 16, 64])
def test_walk_tasks_of_a_few_steps_equal_oracle(task_steps):
    # a walk launch split into tasks (64 walkers x task_steps steps): with
    # 47 groups every wave takes each group's next chunk from another wave,
    # so the publish / wait hand-over runs on nearly every task; many
    # boundary events per step and transmission draws in blocks of four
    dt = 2e-2
    sub = sb.gamma_packing(30, 5.0, 0.4, 0.6, D_in=1.7, D_ex=2.0, seed=7, T2_in=70.0, T2_ex=50.0,
                           rmin=0.25, kappa=sb.kappa_for_probability(0.6, 1.7, dt))
    acq = wf.build_acq([wf.RectGrad(g * 1e-6, 3.0, 2.0, (1, 0, 0)) for g in (0.1, 0.3)], 500, dt, GAMMA)
    run = abi.RunSpec(n_walkers=3000, seed=21, boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
    sig, _ = _run_both(f"tasks_{task_steps}", sub, acq, run, schedule=task_steps)
    assert sig.collisions > 100000 and sig.transmissions > 50000


@pytest.mark.parametrize("prob", [1.0, 3.0])
def test_certain_transmission(prob):
    # 4 kappa / v >= 1: infinite hazard, every boundary event transmits
    dt = 1e-2
    sub = sb.hex_lattice(2.5, 6.0, 2, 1, D_in=1.0, D_ex=1.0,
                         kappa=sb.kappa_for_probability(prob, 1.0, dt))
    acq = _pgse(400, dt)
    run = abi.RunSpec(n_walkers=1500, seed=31, boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
    sig, _ = _run_both(f"certain_transmission_{prob}", sub, acq, run)
    assert sig.transmissions == sig.collisions > 1000


@pytest.mark.parametrize("dt,group", [(1e-2, 2), (1.6667e-3, 4), (4.1667e-4, 8)])
def test_lookup_groups_equal_oracle(dt, group):
    # a hexagonal lattice with 1 um gaps walked with steps of 0.245, 0.1 and
    # 0.05 um: lookup groups of two, four and eight steps (every cylinder at
    # least 8 or 16 step lengths from every other); tasks of 10 steps and
    # launches that start inside a group
    sub = sb.hex_lattice(2.5, 6.0, 2, 1, D_in=1.0, D_ex=1.0,
                         kappa=sb.kappa_for_probability(0.3, 1.0, dt))
    L = math.sqrt(6.0 * 1.0 * dt)
    gap = 6.0 - 2 * 2.5
    assert {gap < 8 * L: 2, 8 * L <= gap < 16 * L: 4, 16 * L <= gap: 8}[True] == group
    acq = _pgse(600, dt)
    run = abi.RunSpec(n_walkers=2048, seed=17, boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
    sig, _ = _run_both(f"lookup_group_{group}", sub, acq, run, chunks=[1, 3, 6, 307], schedule=9)
    assert sig.collisions > 1000 and sig.transmissions > 100


def test_walk_tasks_inside_chunked_launches():
    # launches that begin on an even step: every task then starts inside a
    # step pair and takes the single-step path first
    sub = sb.single_cylinder(20.0, 4.0, D_in=2.0, D_ex=2.0,
                             kappa=sb.kappa_for_probability(0.2, 2.0, 1e-2))
    acq = _pgse(1000, 1e-2)
    run = abi.RunSpec(n_walkers=2048, seed=9, boundary_mode=abi.BOUNDARY_PERMEABLE)
    sig, _ = _run_both("tasks_in_chunks", sub, acq, run, chunks=[1, 2, 3, 251, 252, 600], schedule=9)
    assert sig.transmissions > 1000


@pytest.mark.parametrize("wid,cap", [(6, 0), (3, 0), (1, 0), (5, 0), (6, 50)])
def test_grazing_slides_equal_oracle(wid, cap):
    # a walker set within fp32 rounding of the inside of a cylinder with a
    # tangent step direction: 37-69 grazing reflections in one step, or the
    # bounce cap (tests/test_oracle_physics.py builds the state)
    from tests.test_oracle_physics import grazing_case, grazing_state
    sub, acq = grazing_case()
    run, st = grazing_state(sub, acq, wid, cap)
    print(f"[parity] grazing walker {wid} cap {cap}", flush=True)
    ost = po.walk(sub, acq, run, 1, 2, state=st.copy(), threads=1)
    with _engine().Plan(sub, acq, run) as plan:
        plan.walk(0, 1)
        assert np.array_equal(plan.get_state(), po.walk(sub, acq, run, 0, 1, threads=1))
        plan.set_state(st)
        plan.walk(1, 2)
        gst = plan.get_state()
    assert np.array_equal(gst, ost)
    v = abi.state_view(gst, 1, acq.n_profiles)
    assert int(v["ncoll"][0]) >= 37 and int(v["ncap"][0]) == (1 if cap or wid in (1, 5) else 0)


def test_permeable_sphere_lattice_many_events():
    dt = 2e-2
    cs = [[2.5 + 5 * i, 2.5 + 5 * j, 2.5 + 5 * k] for i in range(2) for j in range(2) for k in range(2)]
    sub = abi.SubstrateSpec(a=10.0, b=10.0, c=10.0, D_lat=2.0, types=[abi.PRIM_SPHERE] * 8, centers=cs,
                            radii=[2.2] * 8, D=1.5, T2=60.0, T2_lat=40.0,
                            kappa=sb.kappa_for_probability(0.5, 1.5, dt))
    acq = wf.build_acq([wf.RectGrad(g * 1e-6, 3.0, 2.0, (0, 0, 1)) for g in (0.1, 0.3)], 500, dt, GAMMA)
    run = abi.RunSpec(n_walkers=3000, seed=22, boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
    sig, _ = _run_both("sphere_lattice_permeable", sub, acq, run, chunks=[1, 250])
    assert sig.collisions > 100000 and sig.transmissions > 50000


def test_single_walker_large_ids():
    sub = sb.single_cylinder(12.0, 3.0, D_in=2.0, D_ex=2.0)
    acq = _pgse(900, 1e-2)
    _run_both("one_walker", sub, acq,
              abi.RunSpec(n_walkers=1, seed=2 ** 63 + 12345, walker_offset=2 ** 32 - 1, instance=2 ** 32 - 2))


def test_spheres_rejection_and_planes_cap():
    cs = [[2.5, 2.5, 2.5], [7.5, 7.5, 7.5]]
    sub = abi.SubstrateSpec(a=10.0, b=10.0, c=10.0, D_lat=2.0,
                            types=[abi.PRIM_SPHERE, abi.PRIM_SPHERE, abi.PRIM_PLANE],
                            centers=cs + [[0.0, 0.0, 5.0]], radii=[2.0, 2.0, 0.0], D=2.0,
                            normals=[[0, 0, 0], [0, 0, 0], [0.0, 0.3, 1.0]])
    acq = _pgse(1000, 1e-2, axis=(0, 1, 1))
    sig, _ = _run_both("spheres_rejection", sub, acq,
                       abi.RunSpec(n_walkers=2048, seed=13, boundary_mode=abi.BOUNDARY_REJECTION))
    assert sig.collisions > 100
    sig, _ = _run_both("spheres_planes_cap", sub, acq, abi.RunSpec(n_walkers=2048, seed=14, bounce_cap=2))
    assert sig.collisions > 100


def test_invalid_descriptors_are_errors():
    eng = _engine()
    sub = sb.single_cylinder(12.0, 3.0, D_in=2.0, D_ex=2.0)
    acq = _pgse(100, 1e-2)
    for run in (abi.RunSpec(n_walkers=0), abi.RunSpec(n_walkers=10, boundary_mode=7),
                abi.RunSpec(n_walkers=10, init_mode=abi.INIT_INSIDE_PRIM, init_arg=3)):
        with pytest.raises(eng.AxxError) as e:
            eng.Plan(sub, acq, run)
        assert e.value.code == -1
    big = sb.single_cylinder(12.0, 6.5, D_in=2.0, D_ex=2.0)           # wider than half the cell
    with pytest.raises(eng.AxxError) as e:
        eng.Plan(big, acq, abi.RunSpec(n_walkers=10))
    assert e.value.code == -5
    overlap = abi.SubstrateSpec(a=12.0, b=12.0, c=12.0, D_lat=2.0, types=[0, 0],
                                centers=[[3, 3, 0], [5, 3, 0]], radii=[1.5, 1.5], D=2.0)
    with pytest.raises(eng.AxxError) as e:
        eng.Plan(overlap, acq, abi.RunSpec(n_walkers=10))
    assert e.value.code == -5
    with eng.Plan(sub, acq, abi.RunSpec(n_walkers=10)) as p:
        with pytest.raises(eng.AxxError) as e:
            p.reduce()                       # before the walk finished
        assert e.value.code == -8
        with pytest.raises(eng.AxxError):
            p.walk(5, 10)                    # must restart at 0 or continue


def test_acquisition_limits_are_errors():
    # 64 profile columns is the widest kernel; 65 is refused as a limit, a
    # waveform naming a missing profile and non-finite samples as invalid
    eng = _engine()
    sub = sb.single_cylinder(12.0, 3.0, D_in=2.0, D_ex=2.0)
    n = 50
    run = abi.RunSpec(n_walkers=64)

    def acq(P, wf_p=None, bad=False):
        prof = np.zeros((n, P), np.float32)
        prof[: n // 2] = 1.0
        prof[n // 2:] = -1.0
        if bad:
            prof[3, 0] = np.nan
        wf_p = np.arange(P, dtype=np.int32) if wf_p is None else np.asarray(wf_p, np.int32)
        return abi.AcqSpec(n_steps=n, dt=1e-2, gamma=GAMMA, profiles=prof, wf_profile=wf_p,
                           wf_amplitude=np.full(wf_p.size, 1e-7), wf_dir=np.tile([1.0, 0, 0], (wf_p.size, 1)))

    with eng.Plan(sub, acq(64), run) as p:
        p.walk()
    for q, code in ((acq(65), -4), (acq(4, wf_p=[0, 1, 4]), -1), (acq(4, bad=True), -1)):
        with pytest.raises(eng.AxxError) as e:
            eng.Plan(sub, q, run)
        assert e.value.code == code


def test_restored_state_is_checked():
    eng = _engine()
    sub = sb.single_cylinder(12.0, 3.0, D_in=2.0, D_ex=2.0)
    acq = _pgse(300, 1e-2)
    run = abi.RunSpec(n_walkers=256, seed=3)
    with eng.Plan(sub, acq, run) as p:
        p.walk(0, 100)
        st = p.get_state()
        v = abi.state_view(st, 256, acq.n_profiles)
        for field, value in (("comp", 7), ("step", 0)):
            bad = st.copy()
            abi.state_view(bad, 256, acq.n_profiles)[field][5] = value
            with pytest.raises(eng.AxxError):
                p.set_state(bad)
        bad = st.copy()
        abi.state_view(bad, 256, acq.n_profiles)["pos"][3, 0] = np.nan
        with pytest.raises(eng.AxxError):
            p.set_state(bad)
        p.set_state(st)                     # a good state resumes exactly
        p.walk(100, 300).reduce()
        a = p.results()
    with eng.Plan(sub, acq, run) as q:
        q.walk().reduce()
        b = q.results()
    assert np.array_equal(a.sum_cos, b.sum_cos) and v["step"][0] == 100


@pytest.mark.parametrize("mode", [abi.STEP_AXES, abi.STEP_GAUSSIAN])
def test_step_generators(mode):
    dt = 2e-3
    sub = sb.hex_lattice(2.5, 6.0, 3, 2, D_in=1.7, D_ex=2.0, T2_in=70.0, T2_ex=50.0,
                         kappa=sb.kappa_for_probability(0.05, 1.7, dt))
    acq = wf.build_acq([wf.CosGFunc(g * 1e-6, 3.0, 0.5, 2.5, ax)
                        for g in (0.1, 0.4) for ax in ((1, 0, 0), (0, 0, 1))], 2500, dt, GAMMA)
    run = abi.RunSpec(n_walkers=2048, seed=21, boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1,
                      step_mode=mode)
    sig, _ = _run_both(f"steps_{mode}_permeable", sub, acq, run, chunks=[1, 777])
    assert sig.transmissions > 50
    # the reference's WC setting: Gaussian (or axis) steps, rejection at a pore wall
    sph = sb.single_sphere(16.0, 5.0, D_in=2.0, D_ex=2.0)
    run = abi.RunSpec(n_walkers=2048, seed=22, step_mode=mode, boundary_mode=abi.BOUNDARY_REJECTION,
                      init_mode=abi.INIT_INSIDE_PRIM)
    sig, _ = _run_both(f"steps_{mode}_rejection", sph, _pgse(1500, 1e-2), run)
    assert sig.n_intra == run.n_walkers and sig.collisions > 100


def test_trace_follows_the_oracle_step_by_step():
    """axx_trace: the chosen walkers' positions after every step equal the
    oracle's, walked one step at a time."""
    eng = _engine()
    dt = 2e-3
    sub = sb.hex_lattice(2.5, 6.0, 3, 2, D_in=1.7, D_ex=2.0, kappa=sb.kappa_for_probability(0.05, 1.7, dt))
    acq = _pgse(200, dt)
    run = abi.RunSpec(n_walkers=512, seed=31, boundary_mode=abi.BOUNDARY_PERMEABLE)
    ids = np.array([0, 1, 77, 255, 300, 511], np.uint32)
    xyz, counts = eng.trace(sub, acq, run, ids)
    st = po.walk(sub, acq, run, step_begin=0, step_end=1)
    n = run.n_walkers
    for i in range(acq.n_steps):
        if i > 0:
            st = po.walk(sub, acq, run, step_begin=i, step_end=i + 1, state=st)
        v = abi.state_view(st, n, acq.n_profiles)
        want = (v["r0"][ids] + v["disp"][ids]).astype(np.float32)
        assert np.array_equal(xyz[:, i, :], want), i
    v = abi.state_view(st, n, acq.n_profiles)
    assert np.array_equal(counts[:, 0], v["ncoll"][ids]) and np.array_equal(counts[:, 1], v["ntran"][ids])
    assert counts[:, 0].sum() > 0
    print(f"[parity] trace of {ids.size} walkers x {acq.n_steps} steps equals the oracle", flush=True)


@pytest.mark.parametrize("n_steps", [1, 2, 3])
def test_shortest_walks(n_steps):
    """One time point (placement only), one step, two steps: the walker state
    and sums equal the oracle's (with one time point the phase is the
    gradient at t = 0 on the start position; the G = 0 waveform gives N)."""
    eng = _engine()
    sub = sb.single_cylinder(20.0, 4.0, D_in=2.0, D_ex=2.0)
    acq = _pgse(n_steps, 1e-3)
    run = abi.RunSpec(n_walkers=300, seed=5)
    sig, _ = _run_both(f"short_{n_steps}", sub, acq, run)
    assert sig.sum_cos[0] == run.n_walkers and sig.sum_sin[0] == 0.0
    # the magnetization path at the same lengths
    from axonexa_amd import fields as fl
    f = fl.build_field([fl.QSC_GRE(1e-3, 1e-4, 0.002, 0.001)], n_steps, 1e-3, gamma=GAMMA,
                       origin=(10.0, 10.0, 10.0))
    with eng.Mag(sub, f, run) as m:
        m.walk()
        assert np.array_equal(m.get_state(0), po.walk_mag(sub, f, run))
