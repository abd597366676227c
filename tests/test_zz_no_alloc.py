"""The allocation-free one-shot run (axx_run_phase_noalloc) against
axx_run_phase: the same kernels on the same inputs, with every device buffer
carved from the library's per-GPU scratch array instead of hipMalloc, must
give the same sums and counters bit for bit -- on two calls in a row, the
second smaller than the first (the scratch keeps the first call's contents),
and a run larger than the scratch is refused with AXX_ERR_LIMIT."""
import numpy as np
import pytest

from axonexa_amd import abi
from axonexa_amd import substrates as sb
from axonexa_amd import waveforms as wf
from axonexa_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("sum_cos", "sum_sin", "sum_cos2", "sum_sin2", "intra_cos", "intra_sin")


def _same(a, b, label):
    for k in FIELDS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), (label, k)
    assert (a.collisions, a.transmissions, a.cap_hits, a.n_intra, a.n_walkers) == \
        (b.collisions, b.transmissions, b.cap_hits, b.n_intra, b.n_walkers), label


def test_noalloc_run_equals_run_phase_twice():
    from axonexa_amd import engine
    # a permeable hexagonal lattice with T2 and eight cosine-OGSE profiles
    big = W.c2_hex(n_walkers=200000, n_steps=4001)
    # then a smaller run of another substrate: one impermeable cylinder, PGSE
    small = W.c1_cylinder(n_walkers=30000, n_steps=2001)
    for w in (big, small):
        print(f"[noalloc] {w.name} {w.run.n_walkers} walkers x {w.acq.n_steps} steps", flush=True)
        ref = engine.run_phase(w.sub, w.acq, w.run)
        got = engine.run_phase_noalloc(w.sub, w.acq, w.run)
        _same(got, ref, w.name)
        assert ref.collisions > 0


def test_noalloc_refuses_a_run_beyond_the_scratch():
    from axonexa_amd import engine
    w = W.c2_hex(n_walkers=2000000, n_steps=11)      # 344 MB of walker state > 256 MiB
    with pytest.raises(engine.AxxError) as e:
        engine.run_phase_noalloc(w.sub, w.acq, w.run)
    assert e.value.code == -4 and "scratch" in str(e.value)
    # the context stays usable: a run that fits still equals axx_run_phase
    sub = sb.single_sphere(12.0, 3.0, D_in=2.0, D_ex=2.0)
    acq = wf.build_acq([wf.RectGrad(g * 1e-6, 2.0, 1.0) for g in (0.0, 0.2)], 600, 1e-2, W.GAMMA)
    run = abi.RunSpec(n_walkers=4096, seed=5, init_mode=abi.INIT_INSIDE_PRIM, init_arg=0)
    _same(engine.run_phase_noalloc(sub, acq, run), engine.run_phase(sub, acq, run), "sphere after refusal")
