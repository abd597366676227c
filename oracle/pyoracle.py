"""ctypes wrapper of the CPU oracle (``oracle/liboracle.so``).

Test infrastructure only: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` — never by the product package.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

from axonexa_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    """The oracle library: ``oracle/liboracle.so`` (built on first use), or
    the build that the environment variable AXX_ORACLE_LIB names (the
    sanitizer build of ``make -C oracle sanitize``)."""
    global _LIB
    if _LIB is None:
        path = os.environ.get("AXX_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            if os.environ.get("AXX_ORACLE_LIB"):
                raise FileNotFoundError(path)
            build()
        _LIB = _configure(C.CDLL(path))
    return _LIB


@contextlib.contextmanager
def planted(variant: int, workdir: str):
    """The oracle rebuilt with the deliberately wrong variant ``variant``
    (ORC_PLANT in axx_oracle.c) into ``workdir`` and used by every function
    of this module inside the block: tests show their checks fail on it."""
    global _LIB
    so = os.path.join(workdir, f"liboracle_plant{int(variant)}.so")
    subprocess.run(["gcc", "-O2", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
                    f"-DORC_PLANT={int(variant)}", "-o", so, os.path.join(_HERE, "axx_oracle.c"),
                    "-shared", "-lpthread", "-lm"], check=True)
    saved = lib()
    _LIB = _configure(C.CDLL(so))
    try:
        yield
    finally:
        _LIB = saved


def _configure(L):
    u32p = C.POINTER(C.c_uint32)
    L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
    L.orc_unif.argtypes = [C.c_uint32]
    L.orc_unif.restype = C.c_float
    L.orc_sincos_turn.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_sincos_turn_n.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
    L.orc_sincos_u24.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_azimuth_table.argtypes = [C.c_void_p, C.c_void_p]
    L.orc_neg2log_unif.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    L.orc_sincos64.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
    L.orc_sqrt_cr.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    L.orc_walk_mag.argtypes = [C.POINTER(abi.Substrate), C.POINTER(abi.Field), C.POINTER(abi.Run),
                               C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint32), C.c_int, C.c_double]
    L.orc_reduce_mag.argtypes = [C.POINTER(C.c_uint32), C.c_int64, C.c_int, C.c_void_p]
    L.orc_validate.argtypes = [C.POINTER(abi.Substrate), C.POINTER(abi.Acq)]
    L.orc_walk.argtypes = [C.POINTER(abi.Substrate), C.POINTER(abi.Acq), C.POINTER(abi.Run),
                           C.c_int, C.c_int, u32p, C.c_int, C.c_double]
    L.orc_reduce.argtypes = [C.POINTER(abi.Substrate), C.POINTER(abi.Acq), C.POINTER(abi.Run),
                             u32p, C.POINTER(abi.Result)]
    L.orc_image_count.argtypes = [C.POINTER(abi.Substrate), C.POINTER(abi.Acq)]
    L.orc_axes_mask.argtypes = [C.POINTER(abi.Acq)]
    L.orc_hit_test.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                               C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                               C.POINTER(C.c_float)]
    L.orc_walk_events.argtypes = [C.POINTER(abi.Substrate), C.POINTER(abi.Acq), C.POINTER(abi.Run),
                                  C.c_int, C.c_int, u32p, C.c_int, C.c_double, C.POINTER(C.c_int64)]
    L.orc_exponents.argtypes = [C.POINTER(abi.Substrate), C.POINTER(abi.Acq), C.POINTER(abi.Run),
                                u32p, C.c_void_p]
    L.orc_set_unit_cells.argtypes = [C.c_int]
    L.orc_unit_cells.argtypes = [C.POINTER(abi.Substrate), C.POINTER(abi.Acq), C.POINTER(abi.Run)]
    return L


class OracleError(RuntimeError):
    pass


def _check(rc: int, what: str):
    if rc != 0:
        raise OracleError(f"{what}: {abi.ERRORS.get(rc, rc)} ({rc})")


def philox(ctr, key) -> np.ndarray:
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return np.array(list(o), dtype=np.uint32)


def unif(w: int) -> float:
    return float(lib().orc_unif(int(w) & 0xFFFFFFFF))


def sincos_turn(w: int):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincos_turn(int(w) & 0xFFFFFFFF, C.byref(s), C.byref(c))
    return s.value, c.value


def neg2log_unif(w) -> np.ndarray:
    """-2 ln u(w) of the Gaussian step generator, elementwise."""
    w = np.ascontiguousarray(w, dtype=np.uint32)
    out = np.empty(w.shape, np.float32)
    lib().orc_neg2log_unif(w.ctypes.data, out.ctypes.data, w.size)
    return out


def sincos_turn_n(w):
    """sin, cos of 2 pi u(w) for an array of Philox words (float32 arrays)."""
    w = np.ascontiguousarray(w, dtype=np.uint32)
    s = np.empty(w.size, np.float32)
    c = np.empty(w.size, np.float32)
    lib().orc_sincos_turn_n(w.ctypes.data, s.ctypes.data, c.ctypes.data, w.size)
    return s, c


def sincos_u24(v: int):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincos_u24(int(v) & 0xFFFFFF, C.byref(s), C.byref(c))
    return s.value, c.value


def azimuth_table():
    """The 256 (cos, sin) entries the azimuth of a step is turned from."""
    c = np.empty(256, np.float32)
    s = np.empty(256, np.float32)
    lib().orc_azimuth_table(c.ctypes.data, s.ctypes.data)
    return c, s


def sincos64(x):
    """Double-precision sine and cosine of the Bloch path, elementwise."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c = np.empty_like(x), np.empty_like(x)
    lib().orc_sincos64(x.ctypes.data, s.ctypes.data, c.ctypes.data, x.size)
    return s, c


def sqrt_cr(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().orc_sqrt_cr(x.ctypes.data, y.ctypes.data, x.size)
    return y


def walk_mag(sub, field, run, group=0, step_begin=0, step_end=None, state=None, threads=None,
             bin_width=0.0) -> np.ndarray:
    """Magnetization walk of one group of fields; returns the group's state."""
    if step_end is None:
        step_end = field.n_steps
    F = min(abi.MAG_GROUP, field.n_fields - group * abi.MAG_GROUP)
    if state is None:
        state = np.zeros(abi.mag_state_fields(F) * run.n_walkers, np.uint32)
    s, f, r = sub.struct(), field.struct(), run.struct()
    if threads is None:
        threads = min(os.cpu_count() or 1, 16)
    rc = lib().orc_walk_mag(C.byref(s), C.byref(f), C.byref(r), int(group), int(step_begin), int(step_end),
                            state.ctypes.data_as(C.POINTER(C.c_uint32)), int(threads), float(bin_width))
    _check(rc, "orc_walk_mag")
    return state


def reduce_mag(state, n_walkers: int, F: int) -> np.ndarray:
    """[F, 3] sums of Mx, My, Mz over walkers."""
    out = np.zeros(F * 3)
    lib().orc_reduce_mag(state.ctypes.data_as(C.POINTER(C.c_uint32)), int(n_walkers), int(F), out.ctypes.data)
    return out.reshape(F, 3)


def run_mag(sub, field, run, threads=None) -> abi.MagSignals:
    """Every group walked and summed (the oracle of axx_run_mag)."""
    n_groups = (field.n_fields + abi.MAG_GROUP - 1) // abi.MAG_GROUP
    sums, coll = [], (0, 0, 0)
    for g in range(n_groups):
        F = min(abi.MAG_GROUP, field.n_fields - g * abi.MAG_GROUP)
        st = walk_mag(sub, field, run, group=g, threads=threads)
        sums.append(reduce_mag(st, run.n_walkers, F))
        if g == 0:
            v = st.reshape(abi.mag_state_fields(F), run.n_walkers)
            coll = (int(v[abi.F_NCOLL].sum(dtype=np.uint64)), int(v[abi.F_NTRAN].sum(dtype=np.uint64)),
                    int(v[abi.F_NCAP].sum(dtype=np.uint64)))
    s = np.concatenate(sums)
    return abi.MagSignals(mx=s[:, 0].copy(), my=s[:, 1].copy(), mz=s[:, 2].copy(), n_walkers=run.n_walkers,
                          collisions=coll[0], transmissions=coll[1], cap_hits=coll[2])


@contextlib.contextmanager
def whole_cell():
    """Inside the block the oracle walks every substrate in its whole cell
    (no unit-cell reduction of repeating substrates)."""
    lib().orc_set_unit_cells(0)
    try:
        yield
    finally:
        lib().orc_set_unit_cells(1)


def unit_cells(sub, acq, run) -> int:
    """Unit cells of the substrate's cell the oracle walks in one of."""
    s, q, r = sub.struct(), acq.struct(), run.struct()
    return lib().orc_unit_cells(C.byref(s), C.byref(q), C.byref(r))


def validate(sub: abi.SubstrateSpec, acq: abi.AcqSpec) -> int:
    s, q = sub.struct(), acq.struct()
    return lib().orc_validate(C.byref(s), C.byref(q))


def image_count(sub: abi.SubstrateSpec, acq: abi.AcqSpec) -> int:
    s, q = sub.struct(), acq.struct()
    return lib().orc_image_count(C.byref(s), C.byref(q))


def axes_mask(acq: abi.AcqSpec) -> int:
    """Moment axes the walk keeps (bit 0 = x, 1 = y, 2 = z)."""
    q = acq.struct()
    return lib().orc_axes_mask(C.byref(q))


def hit_test(ptype, center, R, inside, s3, u3, ell):
    sa = (C.c_float * 3)(*s3)
    ua = (C.c_float * 3)(*u3)
    t = C.c_float(-1.0)
    hit = lib().orc_hit_test(int(ptype), float(center[0]), float(center[1]), float(center[2]),
                             float(R), int(inside), sa, ua, float(ell), C.byref(t))
    return bool(hit), t.value


def new_state(n_walkers: int, n_profiles: int) -> np.ndarray:
    return np.zeros(abi.state_fields(n_profiles) * n_walkers, dtype=np.uint32)


def walk(sub, acq, run, step_begin=0, step_end=None, state=None, threads=None,
         bin_width=0.0) -> np.ndarray:
    """Advance walkers over [step_begin, step_end) in place; returns the state."""
    if step_end is None:
        step_end = acq.n_steps
    if state is None:
        state = new_state(run.n_walkers, acq.n_profiles)
    if threads is None:
        threads = os.cpu_count() or 1
    s, q, r = sub.struct(), acq.struct(), run.struct()
    rc = lib().orc_walk(C.byref(s), C.byref(q), C.byref(r), int(step_begin), int(step_end),
                        state.ctypes.data_as(C.POINTER(C.c_uint32)), int(threads),
                        float(bin_width))
    _check(rc, "orc_walk")
    return state


def reduce(sub, acq, run, state) -> abi.Signals:
    s, q, r = sub.struct(), acq.struct(), run.struct()
    res, bufs = abi.alloc_result(acq.n_waveforms)
    rc = lib().orc_reduce(C.byref(s), C.byref(q), C.byref(r),
                          state.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(res))
    _check(rc, "orc_reduce")
    return abi.signals_from(res, bufs)


def run_phase(sub, acq, run, threads=None, bin_width=0.0):
    st = walk(sub, acq, run, threads=threads, bin_width=bin_width)
    return reduce(sub, acq, run, st), st


def walk_events(sub, acq, run, step_begin=0, step_end=None, state=None, threads=None, bin_width=0.0):
    """walk() that also returns the boundary events at circles and spheres:
    (from inside, transmitted from inside, from outside, transmitted from
    outside)."""
    if step_end is None:
        step_end = acq.n_steps
    if state is None:
        state = new_state(run.n_walkers, acq.n_profiles)
    if threads is None:
        threads = os.cpu_count() or 1
    ev = (C.c_int64 * 4)()
    s, q, r = sub.struct(), acq.struct(), run.struct()
    rc = lib().orc_walk_events(C.byref(s), C.byref(q), C.byref(r), int(step_begin), int(step_end),
                               state.ctypes.data_as(C.POINTER(C.c_uint32)), int(threads), float(bin_width), ev)
    _check(rc, "orc_walk_events")
    return state, tuple(int(v) for v in ev)


def exponents(sub, acq, run, state) -> np.ndarray:
    """Per-walker T2 exponent of a finished walk (signal weight exp(-x))."""
    out = np.empty(run.n_walkers, np.float64)
    s, q, r = sub.struct(), acq.struct(), run.struct()
    rc = lib().orc_exponents(C.byref(s), C.byref(q), C.byref(r), state.ctypes.data_as(C.POINTER(C.c_uint32)),
                             out.ctypes.data)
    _check(rc, "orc_exponents")
    return out
