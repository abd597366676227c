"""Physics of the CPU oracle against analytic results (what pins the oracle
where no golden vectors exist: the reference ships no tests or fixtures).

  * free diffusion: the exact discrete signal for fixed-length isotropic
    steps, E = prod_k sinc(gamma dt |Q_k| L), within 3 sigma;
  * free limit exp(-b D) (continuum) within 3 sigma + discretisation;
  * restricted diffusion perpendicular to a cylinder / inside a sphere:
    Gaussian-phase approximation (Neuman; van Gelderen 1994) within
    3 sigma + 2 % of (1 - E) (the GPD's own error grows with q);
  * parallel to cylinders: z-motion is exactly free (reflection normals have
    no z component), so the sinc product holds exactly;
  * T2 weighting without permeability: exp(-(n-1) dt / T2) exactly;
  * permeable membranes with p = 4 kappa / v keep a uniform density;
  * binned and linear-scan candidate searches give identical walkers;
  * chunked (checkpointed) walks and thread counts give identical walkers.
Reference semantics: src/kernels/kernelLattice.cuh:1-104,
src/kernels/boundaryCheck.cuh:34-64, :66-113.
"""
import math

import numpy as np
import pytest

from axonexa_amd import abi
from axonexa_amd import substrates as sb
from axonexa_amd import waveforms as wf
from axonexa_amd import workloads as W
from oracle import pyoracle as po
from tests import analytic as an

GAMMA = 2.675e5


def test_free_box_matches_exact_sinc_product():
    w = W.c1_free(n_walkers=20000)
    sig, _ = po.run_phase(w.sub, w.acq, w.run, bin_width=1.0)
    exact = np.array([w and wf.free_signal_exact(w.acq, j, 2.0, cell=(20, 20, 20)).real
                      for j in range(w.acq.n_waveforms)])
    z = np.abs(sig.E_re - exact) / np.maximum(sig.sigma_re, 1e-12)
    assert np.all(z[1:] < 3.0), z
    assert abs(sig.E_re[0] - 1.0) < 1e-12
    # the continuum limit exp(-bD) differs from the exact discrete signal by
    # O(dt/delta); both agree within 3 sigma + 1e-3 here
    b = np.array([f for f in [wf.RectGrad(g * 1e-6, 5.0, 10.0).getB(GAMMA)
                              for g in np.linspace(0, 0.3, 16)]])
    assert np.all(np.abs(sig.E_re - np.exp(-b * 2.0)) < 3 * sig.sigma_re + 1e-3)


def test_cylinder_perpendicular_matches_gpd():
    w = W.c1_cylinder(n_walkers=20000)
    w.run.init_mode, w.run.init_arg = abi.INIT_INSIDE_PRIM, 0
    sig, _ = po.run_phase(w.sub, w.acq, w.run, bin_width=1.0)
    assert sig.n_intra == w.run.n_walkers
    for j, g in enumerate(np.linspace(0.0, 0.3, 16) * 1e-6):
        e = an.gpd_pgse(GAMMA, g, 5.0, 15.0, 2.0, 4.0, "cylinder")
        assert abs(sig.E_re[j] - e) <= 3 * sig.sigma_re[j] + 0.02 * (1 - e) + 1e-9, (j, sig.E_re[j], e)
    # in the Neuman (long-time) regime the GPD and Neuman's closed form agree
    assert abs(an.gpd_pgse(GAMMA, 1e-7, 40.0, 60.0, 2.0, 2.0, "cylinder")
               - an.neuman_limit(GAMMA, 1e-7, 40.0, 2.0, 2.0, "cylinder")) < 1e-3


def test_sphere_matches_gpd():
    sub = sb.single_sphere(16.0, 3.0, D_in=2.0, D_ex=2.0)
    gs = np.linspace(0.0, 0.4, 9) * 1e-6
    acq = wf.build_acq([wf.RectGrad(g, 5.0, 10.0) for g in gs], 20000, 1e-3, GAMMA)
    run = abi.RunSpec(n_walkers=8000, seed=4, init_mode=abi.INIT_INSIDE_PRIM, init_arg=0)
    sig, _ = po.run_phase(sub, acq, run, bin_width=1.0)
    for j, g in enumerate(gs):
        e = an.gpd_pgse(GAMMA, g, 5.0, 15.0, 2.0, 3.0, "sphere")
        assert abs(sig.E_re[j] - e) <= 3 * sig.sigma_re[j] + 0.02 * (1 - e) + 1e-9, (j, sig.E_re[j], e)
    assert abs(an.gpd_pgse(GAMMA, 1e-7, 40.0, 60.0, 2.0, 2.0, "sphere")
               - an.neuman_limit(GAMMA, 1e-7, 40.0, 2.0, 2.0, "sphere")) < 1e-3


def test_parallel_to_cylinders_is_free_diffusion_exactly():
    sub = sb.hex_lattice(2.5, 6.0, 2, 1, D_in=2.0, D_ex=2.0)
    gs = np.linspace(0.0, 0.3, 6) * 1e-6
    acq = wf.build_acq([wf.RectGrad(g, 5.0, 10.0, (0, 0, 1)) for g in gs], 20000, 1e-3, GAMMA)
    run = abi.RunSpec(n_walkers=10000, seed=9)
    sig, _ = po.run_phase(sub, acq, run, bin_width=1.0)
    assert sig.collisions > 0
    for j in range(1, len(gs)):
        exact = wf.free_signal_exact(acq, j, 2.0, cell=(sub.a, sub.b, sub.c)).real
        assert abs(sig.E_re[j] - exact) < 3 * sig.sigma_re[j], (j, sig.E_re[j], exact)


def test_t2_weight_without_permeability_is_exact():
    sub = sb.single_cylinder(12.0, 3.0, D_in=1.5, D_ex=2.0, T2_in=70.0, T2_ex=50.0)
    acq = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0)], 3001, 1e-2, GAMMA)
    run = abi.RunSpec(n_walkers=3000, seed=2, relaxation=1)
    sig, st = po.run_phase(sub, acq, run, bin_width=1.0)
    v = abi.state_view(st, run.n_walkers, acq.n_profiles)
    fin = np.mean(v["comp0"] != 0)
    expect = fin * math.exp(-3000 * 1e-2 / 70.0) + (1 - fin) * math.exp(-3000 * 1e-2 / 50.0)
    assert abs(sig.E_re[0] - expect) < 1e-6
    assert np.all(v["comp"] == v["comp0"])


def test_permeable_membranes_keep_uniform_density():
    dt = 1e-2
    sub = sb.hex_lattice(2.5, 6.0, 2, 1, D_in=1.0, D_ex=2.5, kappa=sb.kappa_for_probability(0.2, 1.0, dt))
    acq = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0)], 4000, dt, GAMMA)
    run = abi.RunSpec(n_walkers=20000, seed=5, boundary_mode=abi.BOUNDARY_PERMEABLE)
    st = po.walk(sub, acq, run, bin_width=1.0)
    v = abi.state_view(st, run.n_walkers, acq.n_profiles)
    f = sb.hex_intra_fraction(2.5, 6.0)
    assert v["ntran"].sum() > 20000           # well mixed
    frac = np.mean(v["comp"] != 0)
    sd = math.sqrt(f * (1 - f) / run.n_walkers)
    assert abs(frac - f) < 4 * sd, (frac, f)
    assert abs(np.mean(v["comp0"] != 0) - f) < 4 * sd


@pytest.mark.parametrize("mode", [abi.BOUNDARY_SPECULAR, abi.BOUNDARY_PERMEABLE, abi.BOUNDARY_REJECTION])
def test_binned_equals_linear_scan_and_threads(mode):
    dt = 2e-3
    sub = sb.hex_lattice(2.5, 6.0, 3, 2, D_in=1.7, D_ex=2.0, T2_in=70.0, T2_ex=50.0,
                         kappa=sb.kappa_for_probability(0.05, 1.7, dt))
    acq = wf.build_acq([wf.CosGFunc(4e-7, 3.0, 0.5, 2.5, (1, 0, 1))], 1500, dt, GAMMA)
    run = abi.RunSpec(n_walkers=1024, seed=77, boundary_mode=mode, relaxation=1)
    a = po.walk(sub, acq, run, threads=1)
    b = po.walk(sub, acq, run, threads=7, bin_width=0.9)
    c = po.walk(sub, acq, run, threads=3, bin_width=2.5)
    assert np.array_equal(a, b) and np.array_equal(a, c)


@pytest.mark.parametrize("permeable", [False, True])
def test_chunked_walk_is_bitwise_identical(permeable):
    # permeable: transmission draws come four boundary events per Philox call,
    # numbered over the whole walk, so a chunk may start inside a block of four
    kappa = sb.kappa_for_probability(0.2, 2.0, 1e-2) if permeable else 0.0
    sub = sb.single_cylinder(20.0, 4.0, D_in=2.0, D_ex=2.0, kappa=kappa)
    acq = W.c1_acq(n_steps=2000, dt=1e-2)
    mode = abi.BOUNDARY_PERMEABLE if permeable else abi.BOUNDARY_SPECULAR
    run = abi.RunSpec(n_walkers=512, seed=8, boundary_mode=mode)
    one = po.walk(sub, acq, run, bin_width=1.0)
    st = po.walk(sub, acq, run, 0, 1, bin_width=1.0)
    for a, b in ((1, 2), (2, 999), (999, 1500), (1500, 2000)):
        st = po.walk(sub, acq, run, a, b, state=st, bin_width=1.0)
    assert np.array_equal(one, st)
    v = abi.state_view(one, run.n_walkers, acq.n_profiles)
    if permeable:
        assert int(v["ntran"].sum()) > 1000


def test_step_length_and_isometry_invariants():
    # after one step (i=1) in an empty box every walker moved exactly L
    sub = sb.empty_box(10.0, 2.0)
    acq = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0)], 2, 1e-3, GAMMA)
    run = abi.RunSpec(n_walkers=4096, seed=3)
    st = po.walk(sub, acq, run)
    v = abi.state_view(st, run.n_walkers, acq.n_profiles)
    L = math.sqrt(6 * 2.0 * 1e-3)
    d = np.linalg.norm(v["disp"].astype(np.float64), axis=1)
    assert np.max(np.abs(d - L)) < 1e-6
    # with reflections the path length is still L: displacement <= L
    sub = sb.single_sphere(4.0, 1.0, D_in=2.0, D_ex=2.0)
    acq = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0)], 2, 0.05, GAMMA)
    run = abi.RunSpec(n_walkers=4096, seed=3, init_mode=abi.INIT_INSIDE_PRIM)
    st = po.walk(sub, acq, run)
    v = abi.state_view(st, run.n_walkers, acq.n_profiles)
    L = math.sqrt(6 * 2.0 * 0.05)
    d = np.linalg.norm(v["disp"].astype(np.float64), axis=1)
    assert np.all(d <= L * (1 + 1e-6)) and v["ncoll"].sum() > 100
    r = np.linalg.norm(v["pos"].astype(np.float64) - 2.0, axis=1)
    assert np.all(r < 1.0 + 1e-5)          # nobody left the sphere


def test_translation_of_the_substrate_shifts_walkers():
    # the same walk in a substrate shifted by a lattice vector fraction gives
    # the same displacements when the initial positions are shifted alike:
    # checked through the region-independent free box with one cylinder at
    # two positions related by a full cell translation (images identical)
    acq = W.c1_acq(n_steps=500, dt=1e-2)
    run = abi.RunSpec(n_walkers=256, seed=12)
    a = sb.single_cylinder(20.0, 4.0, D_in=2.0, D_ex=2.0, center=(10.0, 10.0, 0.0))
    b = sb.single_cylinder(20.0, 4.0, D_in=2.0, D_ex=2.0, center=(30.0, -10.0, 0.0))
    assert np.array_equal(po.walk(a, acq, run), po.walk(b, acq, run))


def test_parallel_planes_match_gpd():
    # two reflecting planes x = a/4 and x = 3a/4 in a periodic cell: both
    # slabs (inner, and outer across the periodic edge) have width l = a/2
    a = 16.0
    sub = abi.SubstrateSpec(a=a, b=a, c=a, D_lat=2.0, types=[abi.PRIM_PLANE] * 2,
                            centers=[[a / 4, 0, 0], [3 * a / 4, 0, 0]], radii=[1.0, 1.0],
                            normals=[[1, 0, 0], [-1, 0, 0]])
    gs = np.linspace(0.0, 0.3, 7) * 1e-6
    acq = wf.build_acq([wf.RectGrad(g, 5.0, 10.0) for g in gs], 20000, 1e-3, GAMMA)
    run = abi.RunSpec(n_walkers=8000, seed=21)
    sig, st = po.run_phase(sub, acq, run, bin_width=1.0)
    v = abi.state_view(st, run.n_walkers, acq.n_profiles)
    x0, x = v["r0"][:, 0], v["pos"][:, 0]
    inner0, inner = (x0 > a / 4) & (x0 < 3 * a / 4), (x > a / 4) & (x < 3 * a / 4)
    assert np.array_equal(inner0, inner)       # nobody crossed a plane
    assert sig.collisions > 1000
    for j, g in enumerate(gs):
        e = an.gpd_pgse_planes(GAMMA, g, 5.0, 15.0, 2.0, a / 2)
        assert abs(sig.E_re[j] - e) <= 3 * sig.sigma_re[j] + 0.02 * (1 - e) + 1e-9, (j, sig.E_re[j], e)


def test_oblique_plane_bands_hold_walkers():
    # plane x + y = a (normal (1,1,0)/sqrt 2) in a periodic square cell: its
    # images x + y = k a cut the plane into bands that no walker may leave.
    # The band is read off the unwrapped position r0 + displacement.
    a = 8.0
    sub = abi.SubstrateSpec(a=a, b=a, c=a, D_lat=2.0, types=[abi.PRIM_PLANE],
                            centers=[[a / 2, a / 2, 0.0]], radii=[0.0], normals=[[1.0, 1.0, 0.0]])
    acq = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0)], 5000, 1e-3, GAMMA)
    run = abi.RunSpec(n_walkers=4000, seed=5)
    st = po.walk(sub, acq, run)
    v = abi.state_view(st, run.n_walkers, acq.n_profiles)
    r0 = v["r0"].astype(np.float64)
    r1 = r0 + v["disp"].astype(np.float64)
    b0 = np.floor((r0[:, 0] + r0[:, 1]) / a)
    b1 = np.floor((r1[:, 0] + r1[:, 1]) / a)
    assert np.array_equal(b0, b1)
    assert int(v["ncoll"].sum()) > 1000
    # the binned scan gives the same walk
    assert np.array_equal(po.walk(sub, acq, run, bin_width=1.0), st)


def test_plane_rejection_and_invalid_init():
    sub = sb.parallel_planes(10.0, 2.5, 7.5, 2.0)
    acq = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0)], 2000, 1e-2, GAMMA)
    run = abi.RunSpec(n_walkers=2000, seed=9, boundary_mode=abi.BOUNDARY_REJECTION)
    st = po.walk(sub, acq, run)
    v = abi.state_view(st, run.n_walkers, acq.n_profiles)
    x0, x = v["r0"][:, 0], v["pos"][:, 0]
    assert np.array_equal((x0 > 2.5) & (x0 < 7.5), (x > 2.5) & (x < 7.5))
    assert int(v["ncoll"].sum()) > 100
    with pytest.raises(RuntimeError):
        po.walk(sub, acq, abi.RunSpec(n_walkers=4, seed=1, init_mode=abi.INIT_INSIDE_PRIM, init_arg=0))
    with pytest.raises(RuntimeError):
        po.walk(sub, acq, abi.RunSpec(n_walkers=4, seed=1, init_mode=abi.INIT_COMPARTMENT, init_arg=1))


def test_walker_id_shards_reassemble_the_full_run():
    # the sharding rule of a multi-device run: device g walks the id range
    # [g N/G, (g+1) N/G) via walker_offset; states concatenate bitwise and
    # the fp64 sums add up to the single run's
    sub = sb.single_cylinder(12.0, 3.0, D_in=1.5, D_ex=2.0, T2_in=60.0, T2_ex=40.0,
                             kappa=sb.kappa_for_probability(0.05, 1.5, 1e-2))
    acq = W.c1_acq(n_steps=800, dt=1e-2)
    full = abi.RunSpec(n_walkers=3000, seed=91, boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
    st = po.walk(sub, acq, full)
    sig = po.reduce(sub, acq, full, st)
    parts, sums = [], np.zeros_like(sig.sum_cos)
    for g in range(3):
        r = abi.RunSpec(n_walkers=1000, seed=91, walker_offset=1000 * g,
                        boundary_mode=abi.BOUNDARY_PERMEABLE, relaxation=1)
        s = po.walk(sub, acq, r)
        parts.append(s.reshape(-1, 1000))
        sums += po.reduce(sub, acq, r, s).sum_cos
    assert np.array_equal(np.concatenate(parts, axis=1).reshape(-1), st)
    assert np.allclose(sums, sig.sum_cos, rtol=1e-12, atol=1e-9)


def test_gaussian_steps_give_exp_minus_bD_exactly():
    # N(0, 2 D dt) per axis: the phase is Gaussian, E = exp(-b D) with the
    # b of the sampled table (waveforms.b_discrete), within 3 sigma
    gs = np.linspace(0.0, 0.25, 6) * 1e-6
    acq = wf.build_acq([wf.RectGrad(g, 5.0, 10.0) for g in gs], 20000, 1e-3, GAMMA)
    run = abi.RunSpec(n_walkers=20000, seed=8, step_mode=abi.STEP_GAUSSIAN)
    sig, st = po.run_phase(sb.empty_box(20.0, 2.0), acq, run, bin_width=1.0)
    for j in range(1, len(gs)):
        b = wf.b_discrete(acq.profiles[:, acq.wf_profile[j]], acq.wf_amplitude[j], acq.dt, GAMMA)
        assert abs(sig.E_re[j] - math.exp(-b * 2.0)) < 3 * sig.sigma_re[j], (j, sig.E_re[j])


def test_gaussian_step_moments():
    # one step: per-axis displacement N(0, 2 D dt): variance, kurtosis 3,
    # isotropy, and lengths beyond the isotropic step length
    acq = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0)], 2, 1e-3, GAMMA)
    run = abi.RunSpec(n_walkers=200000, seed=4, step_mode=abi.STEP_GAUSSIAN)
    st = po.walk(sb.empty_box(50.0, 2.0), acq, run)
    d = abi.state_view(st, run.n_walkers, 1)["disp"].astype(np.float64)
    var = 2 * 2.0 * 1e-3
    n = run.n_walkers
    assert np.all(np.abs(d.mean(axis=0)) < 5 * math.sqrt(var / n))
    assert np.all(np.abs(d.var(axis=0) / var - 1) < 5 * math.sqrt(2.0 / n))
    kurt = (d ** 4).mean(axis=0) / d.var(axis=0) ** 2
    assert np.all(np.abs(kurt - 3.0) < 5 * math.sqrt(96.0 / n))
    L = math.sqrt(6 * 2.0 * 1e-3)
    assert np.max(np.linalg.norm(d, axis=1)) > 2 * L


def test_axis_sign_steps_give_the_cosine_product():
    # +-s per axis: E = prod_k cos(gamma dt Q_k s) for a gradient along x
    g = 0.2e-6
    acq = wf.build_acq([wf.RectGrad(0.0, 5.0, 10.0), wf.RectGrad(g, 5.0, 10.0)], 20000, 1e-3, GAMMA)
    run = abi.RunSpec(n_walkers=20000, seed=6, step_mode=abi.STEP_AXES)
    sig, st = po.run_phase(sb.empty_box(20.0, 2.0), acq, run, bin_width=1.0)
    prof = acq.profiles[:, acq.wf_profile[1]].astype(np.float64) * acq.wf_amplitude[1]
    Q = np.cumsum(prof[::-1])[::-1][1:]          # Q_k = sum_{i >= k} G_i, k = 1..n-1
    s = math.sqrt(2 * 2.0 * 1e-3)
    exact = float(np.prod(np.cos(GAMMA * acq.dt * Q * s)))
    assert abs(sig.E_re[1] - exact) < 3 * sig.sigma_re[1], (sig.E_re[1], exact)
    # one step moves every axis by exactly +-s, all 8 sign patterns equally
    one = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0)], 2, 1e-3, GAMMA)
    r1 = abi.RunSpec(n_walkers=8000, seed=6, step_mode=abi.STEP_AXES)
    d = abi.state_view(po.walk(sb.empty_box(20.0, 2.0), one, r1), 8000, 1)["disp"].astype(np.float64)
    assert np.allclose(np.abs(d), s, rtol=0.0, atol=4e-6)      # a few ulps of positions ~10 um
    counts = np.bincount(((d > 0) * [1, 2, 4]).sum(axis=1), minlength=8)
    assert counts.min() > 8000 / 8 - 5 * math.sqrt(8000 / 8)


def test_gaussian_rejection_keeps_walkers_in_a_sphere():
    # the reference's WC kernel: Gaussian steps, rejection at the pore wall
    sub = sb.single_sphere(16.0, 5.0, D_in=2.0, D_ex=2.0)
    acq = wf.build_acq([wf.RectGrad(0.1e-6, 5.0, 10.0)], 5000, 1e-3, GAMMA)
    run = abi.RunSpec(n_walkers=4000, seed=2, step_mode=abi.STEP_GAUSSIAN,
                      boundary_mode=abi.BOUNDARY_REJECTION, init_mode=abi.INIT_INSIDE_PRIM)
    st = po.walk(sub, acq, run)
    v = abi.state_view(st, run.n_walkers, acq.n_profiles)
    r = np.linalg.norm(v["pos"].astype(np.float64) - 8.0, axis=1)
    assert np.all(r < 5.0) and np.all(v["comp"] == 1) and int(v["ncoll"].sum()) > 1000


def test_gaussian_steps_need_a_wide_cell():
    acq = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0)], 10, 1.0, GAMMA)      # L = 3.5 um
    with pytest.raises(po.OracleError):
        po.walk(sb.empty_box(30.0, 2.0), acq, abi.RunSpec(n_walkers=4, step_mode=abi.STEP_GAUSSIAN))


def grazing_state(sub, acq, wid, bounce_cap=0):
    """A walker placed in cylinder 0 of `sub` (centre (5, 5), R = 2) at the
    point where the step-1 direction of its own Philox words is tangent to
    the surface, the position rounded to fp32 (so within rounding of the
    surface, on either side): step 1 slides along the inside in grazing
    chords.  Returns (run, state before step 1)."""
    run = abi.RunSpec(n_walkers=1, seed=1234, walker_offset=wid, init_mode=abi.INIT_INSIDE_PRIM,
                      init_arg=0, bounce_cap=bounce_cap)
    st = po.walk(sub, acq, run, 0, 1, threads=1)
    o = po.philox([wid, 0, 0, 0], [1234, 0])
    s, c = po.sincos_turn(int(o[0]))
    n = np.array([-s, c], np.float64)
    n /= np.linalg.norm(n)
    p = np.array([5.0, 5.0]) + 2.0 * n
    v = st.view(np.float32)
    v[abi.F_X], v[abi.F_Y] = np.float32(p[0]), np.float32(p[1])
    return run, st


def grazing_case():
    sub = sb.single_cylinder(10.0, 2.0, D_in=1.7, D_ex=2.0)
    acq = wf.build_acq([wf.RectGrad(0.0, 1.0, 1.0, (1, 0, 0))], 3, 5e-4, GAMMA)
    return sub, acq


def test_grazing_slides_along_a_surface_finish_their_step():
    """Walkers that start a step within fp32 rounding of the inside of a
    cylinder with a tangent direction slide along it in short chords
    (step 0.0714 um, R = 2 um).  Walker 6 needs 69 reflections and ends
    7e-6 um outside the surface by rounding while still tracked inside;
    walkers 1, 2 and 5 slide on until the default cap (4096) ends the step,
    which is counted; a cap of 50 cuts walker 6 at 50."""
    sub, acq = grazing_case()
    for wid, cap, events, capped in ((6, 0, 69, 0), (3, 0, 37, 0), (1, 0, 4096, 1), (2, 0, 4096, 1),
                                     (5, 0, 4096, 1), (6, 50, 50, 1)):
        run, st = grazing_state(sub, acq, wid, cap)
        st = po.walk(sub, acq, run, 1, 2, state=st, threads=1)
        v = abi.state_view(st, 1, acq.n_profiles)
        assert (int(v["ncoll"][0]), int(v["ncap"][0])) == (events, capped), wid
        assert int(v["comp"][0]) == int(v["comp0"][0]) == 1       # still tracked inside its cylinder
        r = math.hypot(float(v["pos"][0][0]) - 5.0, float(v["pos"][0][1]) - 5.0)
        assert r < 2.0 + 1e-5
