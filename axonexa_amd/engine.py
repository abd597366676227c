"""ctypes binding of the HIP library ``libaxonexa.so`` (C ABI of
``include/axonexa.h``).

This is the product path: every call goes to the gfx950 kernels.  There is
no CPU fallback — if the library is missing or no GPU is present, the calls
raise.

Reference interface mirrored: ``phaseAcquisitionStream<F>::
runAcquisitionLattice`` (src/acquisition/phaseAcquisitionStream.cuh:126-190)
as :func:`run_phase`; its parts (allocation/upload :134-158, walk kernel
:163-169, reductions :171-183, download :185-189) as :class:`Plan`.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaxonexa.so")
_LIB = None


class AxxError(RuntimeError):
    def __init__(self, code: int, what: str, detail: str = ""):
        self.code = code
        msg = f"{what}: {abi.ERRORS.get(code, 'error')} ({code})"
        if detail:
            msg += f" — {detail}"
        super().__init__(msg)


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.axx_create.argtypes = [C.POINTER(vp), C.c_int]
    L.axx_destroy.argtypes = [vp]
    L.axx_strerror.argtypes = [C.c_int]
    L.axx_strerror.restype = C.c_char_p
    L.axx_last_error.argtypes = [vp]
    L.axx_last_error.restype = C.c_char_p
    L.axx_version.restype = C.c_char_p
    S, Q, R, Res = (C.POINTER(abi.Substrate), C.POINTER(abi.Acq), C.POINTER(abi.Run),
                    C.POINTER(abi.Result))
    L.axx_run_phase.argtypes = [vp, S, Q, R, Res]
    L.axx_run_phase_noalloc.argtypes = [vp, S, Q, R, Res]
    L.axx_trace.argtypes = [vp, S, Q, R, C.POINTER(C.c_uint32), C.c_int, C.POINTER(C.c_float),
                            C.POINTER(C.c_uint32)]
    L.axx_plan_create.argtypes = [vp, S, Q, R, C.POINTER(vp)]
    L.axx_plan_create_ex.argtypes = [vp, S, Q, R, C.POINTER(abi.Tuning), C.POINTER(vp)]
    L.axx_plan_walk.argtypes = [vp, C.c_int, C.c_int]
    L.axx_plan_reduce.argtypes = [vp]
    L.axx_plan_results.argtypes = [vp, Res]
    L.axx_plan_sync.argtypes = [vp]
    L.axx_plan_get_state.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.axx_plan_set_state.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.axx_plan_time_walk.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_float)]
    L.axx_plan_get_config.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.axx_plan_set_config.argtypes = [vp, C.c_int, C.c_int]
    L.axx_plan_set_schedule.argtypes = [vp, C.c_int]
    L.axx_plan_get_schedule.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.axx_plan_profile_count.argtypes = [vp]
    L.axx_plan_unit_cells.argtypes = [vp]
    L.axx_plan_destroy.argtypes = [vp]
    IP = C.POINTER(abi.Instance)
    L.axx_batch_create.argtypes = [vp, C.c_int, IP, C.POINTER(abi.Tuning), C.POINTER(vp)]
    L.axx_batch_run.argtypes = [vp]
    L.axx_batch_time.argtypes = [vp, C.POINTER(C.c_float)]
    L.axx_batch_sync.argtypes = [vp]
    L.axx_batch_results.argtypes = [vp, C.c_int, Res]
    L.axx_batch_get_state.argtypes = [vp, C.c_int, C.POINTER(C.c_uint32)]
    L.axx_batch_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_float)]
    L.axx_batch_set_config.argtypes = [vp, C.c_int, C.c_int]
    L.axx_batch_get_config.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.axx_batch_destroy.argtypes = [vp]
    L.axx_run_batch.argtypes = [vp, C.c_int, IP, Res]
    fp = C.POINTER(C.c_float)
    L.axx_mlp_create.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(vp)]
    L.axx_mlp_param_count.argtypes = [vp, C.POINTER(C.c_int64)]
    L.axx_mlp_set_params.argtypes = [vp, fp]
    L.axx_mlp_get_params.argtypes = [vp, fp]
    L.axx_mlp_forward.argtypes = [vp, fp, C.c_int, fp]
    L.axx_mlp_time_forward.argtypes = [vp, C.c_int, fp]
    L.axx_mlp_loss_grad.argtypes = [vp, fp, fp, C.c_int, fp, fp]
    L.axx_mlp_set_data.argtypes = [vp, fp, fp, C.c_int]
    L.axx_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_uint64, fp]
    L.axx_mlp_destroy.argtypes = [vp]
    Fp, MR = C.POINTER(abi.Field), C.POINTER(abi.MagResult)
    L.axx_mag_sample_count.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
    L.axx_mag_create.argtypes = [vp, S, Fp, R, C.POINTER(vp)]
    L.axx_mag_walk.argtypes = [vp, C.c_int, C.c_int]
    L.axx_mag_time_walk.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_float)]
    L.axx_mag_set_schedule.argtypes = [vp, C.c_int]
    L.axx_mag_results.argtypes = [vp, MR]
    L.axx_mag_get_state.argtypes = [vp, C.c_int, C.POINTER(C.c_uint32)]
    L.axx_mag_destroy.argtypes = [vp]
    L.axx_run_mag.argtypes = [vp, S, Fp, R, MR]
    MI = C.POINTER(abi.MagInstance)
    L.axx_mag_batch_create.argtypes = [vp, C.c_int, MI, C.POINTER(vp)]
    L.axx_mag_batch_run.argtypes = [vp]
    L.axx_mag_batch_time.argtypes = [vp, C.POINTER(C.c_float)]
    L.axx_mag_batch_results.argtypes = [vp, C.c_int, MR]
    L.axx_mag_batch_get_state.argtypes = [vp, C.c_int, C.POINTER(C.c_uint32)]
    L.axx_mag_batch_destroy.argtypes = [vp]
    L.axx_run_mag_batch.argtypes = [vp, C.c_int, MI, MR]
    _LIB = L
    return L


class Context:
    """One device + stream (axx_create)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        rc = lib().axx_create(C.byref(self._h), int(device))
        if rc != 0:
            raise AxxError(rc, "axx_create")

    def check(self, rc: int, what: str):
        if rc != 0:
            detail = lib().axx_last_error(self._h).decode(errors="replace")
            raise AxxError(rc, what, detail)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().axx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


_DEFAULT_CTX: Optional[Context] = None


def default_context() -> Context:
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


class Plan:
    """Device-resident run: create -> walk (chunks) -> reduce -> results."""

    def __init__(self, sub: abi.SubstrateSpec, acq: abi.AcqSpec, run: abi.RunSpec,
                 ctx: Optional[Context] = None, block: int = 0, bin_width: float = 0.0,
                 lds_budget: int = 0, whole_cell: bool = False):
        self.ctx = ctx or default_context()
        self.sub, self.acq, self.run = sub, acq, run
        self._s, self._q, self._r = sub.struct(), acq.struct(), run.struct()
        self._t = abi.Tuning(int(block), float(bin_width), int(lds_budget), int(bool(whole_cell)))
        self._h = C.c_void_p()
        rc = lib().axx_plan_create_ex(self.ctx.handle, C.byref(self._s), C.byref(self._q),
                                      C.byref(self._r), C.byref(self._t), C.byref(self._h))
        self.ctx.check(rc, "axx_plan_create_ex")

    def walk(self, step_begin: int = 0, step_end: Optional[int] = None):
        if step_end is None:
            step_end = self.acq.n_steps
        self.ctx.check(lib().axx_plan_walk(self._h, int(step_begin), int(step_end)), "axx_plan_walk")
        return self

    def time_walk(self, step_begin: int, step_end: int) -> float:
        ms = C.c_float()
        self.ctx.check(lib().axx_plan_time_walk(self._h, int(step_begin), int(step_end),
                                                C.byref(ms)), "axx_plan_time_walk")
        return ms.value

    def reduce(self):
        self.ctx.check(lib().axx_plan_reduce(self._h), "axx_plan_reduce")
        return self

    def sync(self):
        self.ctx.check(lib().axx_plan_sync(self._h), "axx_plan_sync")

    def results(self) -> abi.Signals:
        res, bufs = abi.alloc_result(self.acq.n_waveforms)
        self.ctx.check(lib().axx_plan_results(self._h, C.byref(res)), "axx_plan_results")
        return abi.signals_from(res, bufs)

    def get_state(self) -> np.ndarray:
        st = np.zeros(abi.state_fields(self.acq.n_profiles) * self.run.n_walkers, np.uint32)
        self.ctx.check(lib().axx_plan_get_state(self._h, st.ctypes.data_as(C.POINTER(C.c_uint32))),
                       "axx_plan_get_state")
        return st

    def set_state(self, st: np.ndarray):
        st = np.ascontiguousarray(st, dtype=np.uint32)
        self.ctx.check(lib().axx_plan_set_state(self._h, st.ctypes.data_as(C.POINTER(C.c_uint32))),
                       "axx_plan_set_state")

    def config(self):
        b, w, l = C.c_int(), C.c_int(), C.c_int()
        bw = C.c_double()
        self.ctx.check(lib().axx_plan_get_config(self._h, C.byref(b), C.byref(w), C.byref(bw),
                                                 C.byref(l)), "axx_plan_get_config")
        return {"block": b.value, "wpt": w.value, "bin_width": bw.value, "lds_bytes": l.value,
                "unit_cells": int(lib().axx_plan_unit_cells(self._h))}

    def set_config(self, block: int, wpt: int = 1):
        self.ctx.check(lib().axx_plan_set_config(self._h, int(block), int(wpt)), "axx_plan_set_config")

    def set_schedule(self, task_steps: int = 0):
        """Steps per walk task: 0 automatic, -1 the whole walk, > 0 that many."""
        self.ctx.check(lib().axx_plan_set_schedule(self._h, int(task_steps)), "axx_plan_set_schedule")

    def schedule(self):
        s, ls, g = C.c_int(), C.c_int(), C.c_int()
        self.ctx.check(lib().axx_plan_get_schedule(self._h, C.byref(s), C.byref(ls), C.byref(g)),
                       "axx_plan_get_schedule")
        return {"task_steps": s.value, "last_task_steps": ls.value, "last_grid": g.value}

    def close(self):
        if self._h:
            lib().axx_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_phase_noalloc(sub: abi.SubstrateSpec, acq: abi.AcqSpec, run: abi.RunSpec,
                      ctx: Optional[Context] = None) -> abi.Signals:
    """One-shot run without device allocation (axx_run_phase_noalloc): the
    run's device buffers are pieces of the library's per-GPU scratch array."""
    ctx = ctx or default_context()
    s, q, r = sub.struct(), acq.struct(), run.struct()
    res, bufs = abi.alloc_result(acq.n_waveforms)
    ctx.check(lib().axx_run_phase_noalloc(ctx.handle, C.byref(s), C.byref(q), C.byref(r), C.byref(res)),
              "axx_run_phase_noalloc")
    return abi.signals_from(res, bufs)


def run_phase(sub: abi.SubstrateSpec, acq: abi.AcqSpec, run: abi.RunSpec,
              ctx: Optional[Context] = None) -> abi.Signals:
    """One-shot run (axx_run_phase)."""
    ctx = ctx or default_context()
    s, q, r = sub.struct(), acq.struct(), run.struct()
    res, bufs = abi.alloc_result(acq.n_waveforms)
    ctx.check(lib().axx_run_phase(ctx.handle, C.byref(s), C.byref(q), C.byref(r), C.byref(res)),
              "axx_run_phase")
    return abi.signals_from(res, bufs)


def _mag_result(n_fields: int):
    bufs = {k: np.zeros(n_fields) for k in ("mx", "my", "mz")}
    r = abi.MagResult()
    for k, v in bufs.items():
        setattr(r, k, v.ctypes.data_as(C.POINTER(C.c_double)))
    return r, bufs


def _mag_signals(r, bufs) -> abi.MagSignals:
    return abi.MagSignals(n_walkers=int(r.n_walkers), collisions=int(r.collisions),
                          transmissions=int(r.transmissions), cap_hits=int(r.cap_hits),
                          **{k: v.copy() for k, v in bufs.items()})


class Mag:
    """Magnetization (Bloch) run: create -> walk (chunks) -> results
    (axx_mag_*)."""

    def __init__(self, sub: abi.SubstrateSpec, field: abi.FieldSpec, run: abi.RunSpec,
                 ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.sub, self.field, self.run = sub, field, run
        self._s, self._f, self._r = sub.struct(), field.struct(), run.struct()
        self._h = C.c_void_p()
        self.ctx.check(lib().axx_mag_create(self.ctx.handle, C.byref(self._s), C.byref(self._f),
                                            C.byref(self._r), C.byref(self._h)), "axx_mag_create")

    @property
    def n_groups(self) -> int:
        return (self.field.n_fields + abi.MAG_GROUP - 1) // abi.MAG_GROUP

    def group_fields(self, g: int) -> int:
        return min(abi.MAG_GROUP, self.field.n_fields - g * abi.MAG_GROUP)

    def walk(self, step_begin: int = 0, step_end: Optional[int] = None):
        if step_end is None:
            step_end = self.field.n_steps
        self.ctx.check(lib().axx_mag_walk(self._h, int(step_begin), int(step_end)), "axx_mag_walk")
        return self

    def set_schedule(self, task_steps: int = 0):
        """Steps per walk task of every group: 0 automatic, -1 the whole walk, > 0 that many."""
        self.ctx.check(lib().axx_mag_set_schedule(self._h, int(task_steps)), "axx_mag_set_schedule")

    def time_walk(self, step_begin: int, step_end: int) -> float:
        ms = C.c_float()
        self.ctx.check(lib().axx_mag_time_walk(self._h, int(step_begin), int(step_end), C.byref(ms)),
                       "axx_mag_time_walk")
        return ms.value

    def results(self) -> abi.MagSignals:
        res, bufs = _mag_result(self.field.n_fields)
        self.ctx.check(lib().axx_mag_results(self._h, C.byref(res)), "axx_mag_results")
        return _mag_signals(res, bufs)

    def get_state(self, group: int = 0) -> np.ndarray:
        st = np.zeros(abi.mag_state_fields(self.group_fields(group)) * self.run.n_walkers, np.uint32)
        self.ctx.check(lib().axx_mag_get_state(self._h, int(group), st.ctypes.data_as(C.POINTER(C.c_uint32))),
                       "axx_mag_get_state")
        return st

    def close(self):
        if self._h:
            lib().axx_mag_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_mag(sub: abi.SubstrateSpec, field: abi.FieldSpec, run: abi.RunSpec,
            ctx: Optional[Context] = None) -> abi.MagSignals:
    """One-shot magnetization run (axx_run_mag)."""
    ctx = ctx or default_context()
    s, f, r = sub.struct(), field.struct(), run.struct()
    res, bufs = _mag_result(field.n_fields)
    ctx.check(lib().axx_run_mag(ctx.handle, C.byref(s), C.byref(f), C.byref(r), C.byref(res)), "axx_run_mag")
    return _mag_signals(res, bufs)


def trace(sub: abi.SubstrateSpec, acq: abi.AcqSpec, run: abi.RunSpec, walker_ids,
          ctx: Optional[Context] = None):
    """Positions of the chosen walkers after every step (axx_trace):
    xyz [k, n_steps, 3] (start + displacement), counts [k, 3] (collisions,
    transmissions, bounce-cap hits)."""
    ctx = ctx or default_context()
    ids = np.ascontiguousarray(walker_ids, dtype=np.uint32)
    k = ids.size
    xyz = np.zeros((k, acq.n_steps, 3), np.float32)
    counts = np.zeros((k, 3), np.uint32)
    s, q, r = sub.struct(), acq.struct(), run.struct()
    ctx.check(lib().axx_trace(ctx.handle, C.byref(s), C.byref(q), C.byref(r),
                              ids.ctypes.data_as(C.POINTER(C.c_uint32)), k,
                              xyz.ctypes.data_as(C.POINTER(C.c_float)),
                              counts.ctypes.data_as(C.POINTER(C.c_uint32))), "axx_trace")
    return xyz, counts


class MagBatch:
    """Many (substrate, field, run) magnetization runs in one persistent
    launch (axx_mag_batch_*); each run gives exactly its own run_mag."""

    def __init__(self, items, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.items = list(items)
        self._structs = [(s.struct(), f.struct(), r.struct()) for s, f, r in self.items]
        arr = (abi.MagInstance * len(self.items))()
        for k, (s, f, r) in enumerate(self._structs):
            arr[k].substrate = C.pointer(s)
            arr[k].field = C.pointer(f)
            arr[k].run = C.pointer(r)
        self._arr = arr
        self._h = C.c_void_p()
        self.ctx.check(lib().axx_mag_batch_create(self.ctx.handle, len(self.items), arr, C.byref(self._h)),
                       "axx_mag_batch_create")

    def run(self):
        self.ctx.check(lib().axx_mag_batch_run(self._h), "axx_mag_batch_run")
        return self

    def time(self) -> float:
        ms = C.c_float()
        self.ctx.check(lib().axx_mag_batch_time(self._h, C.byref(ms)), "axx_mag_batch_time")
        return ms.value

    def results(self, i: int) -> abi.MagSignals:
        res, bufs = _mag_result(self.items[i][1].n_fields)
        self.ctx.check(lib().axx_mag_batch_results(self._h, int(i), C.byref(res)), "axx_mag_batch_results")
        return _mag_signals(res, bufs)

    def get_state(self, i: int) -> np.ndarray:
        _, f, r = self.items[i]
        st = np.zeros(abi.mag_state_fields(f.n_fields) * r.n_walkers, np.uint32)
        self.ctx.check(lib().axx_mag_batch_get_state(self._h, int(i), st.ctypes.data_as(C.POINTER(C.c_uint32))),
                       "axx_mag_batch_get_state")
        return st

    def close(self):
        if self._h:
            lib().axx_mag_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many (substrate, acquisition, run) instances in one persistent,
    graph-captured launch (axx_batch_*).  Instance i gives bitwise the walkers
    and sums of ``run_phase(*instances[i])``."""

    def __init__(self, instances, ctx: Optional[Context] = None, block: int = 0,
                 lds_budget: int = 0):
        self.ctx = ctx or default_context()
        self.instances = list(instances)
        self._structs = [(s.struct(), q.struct(), r.struct()) for s, q, r in self.instances]
        arr = (abi.Instance * len(self._structs))()
        for k, (s, q, r) in enumerate(self._structs):
            arr[k].substrate = C.pointer(s)
            arr[k].acq = C.pointer(q)
            arr[k].run = C.pointer(r)
        self._arr = arr
        self._t = abi.Tuning(int(block), 0.0, int(lds_budget), 0)
        self._h = C.c_void_p()
        rc = lib().axx_batch_create(self.ctx.handle, len(self._structs), arr, C.byref(self._t),
                                    C.byref(self._h))
        self.ctx.check(rc, "axx_batch_create")

    def __len__(self):
        return len(self.instances)

    def run(self):
        self.ctx.check(lib().axx_batch_run(self._h), "axx_batch_run")
        return self

    def time(self) -> float:
        ms = C.c_float()
        self.ctx.check(lib().axx_batch_time(self._h, C.byref(ms)), "axx_batch_time")
        return ms.value

    def sync(self):
        self.ctx.check(lib().axx_batch_sync(self._h), "axx_batch_sync")

    def results(self, i: int) -> abi.Signals:
        res, bufs = abi.alloc_result(self.instances[i][1].n_waveforms)
        self.ctx.check(lib().axx_batch_results(self._h, int(i), C.byref(res)), "axx_batch_results")
        return abi.signals_from(res, bufs)

    def get_state(self, i: int) -> np.ndarray:
        _, q, r = self.instances[i]
        st = np.zeros(abi.state_fields(q.n_profiles) * r.n_walkers, np.uint32)
        self.ctx.check(lib().axx_batch_get_state(self._h, int(i),
                                                 st.ctypes.data_as(C.POINTER(C.c_uint32))),
                       "axx_batch_get_state")
        return st

    def autotune(self, probe_steps: int = 200):
        b, w, ms = C.c_int(), C.c_int(), C.c_float()
        self.ctx.check(lib().axx_batch_autotune(self._h, int(probe_steps), C.byref(b), C.byref(w),
                                                C.byref(ms)), "axx_batch_autotune")
        return {"block": b.value, "wpt": w.value, "probe_ms": ms.value}

    def set_config(self, block: int, wpt: int = 1):
        self.ctx.check(lib().axx_batch_set_config(self._h, int(block), int(wpt)),
                       "axx_batch_set_config")

    def config(self):
        b, w, g, t = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.ctx.check(lib().axx_batch_get_config(self._h, C.byref(b), C.byref(w), C.byref(g),
                                                  C.byref(t)), "axx_batch_get_config")
        return {"block": b.value, "wpt": w.value, "grid": g.value, "tiles": t.value}

    def close(self):
        if self._h:
            lib().axx_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """BF16 MLP on MFMA with hand-written training kernels (axx_mlp_*)."""

    def __init__(self, dims, max_batch: int = 4096, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.dims = [int(d) for d in dims]
        arr = (C.c_int * len(self.dims))(*self.dims)
        self._h = C.c_void_p()
        rc = lib().axx_mlp_create(self.ctx.handle, len(self.dims) - 1, arr, int(max_batch),
                                  C.byref(self._h))
        self.ctx.check(rc, "axx_mlp_create")
        n = C.c_int64()
        self.ctx.check(lib().axx_mlp_param_count(self._h, C.byref(n)), "axx_mlp_param_count")
        self.n_params = n.value

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.n_params
        self.ctx.check(lib().axx_mlp_set_params(self._h, _fptr(p)), "axx_mlp_set_params")

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        self.ctx.check(lib().axx_mlp_get_params(self._h, _fptr(p)), "axx_mlp_get_params")
        return p

    def forward(self, x) -> np.ndarray:
        x = _f32(x)
        n = x.shape[0]
        y = np.empty((n, self.dims[-1]), np.float32)
        self.ctx.check(lib().axx_mlp_forward(self._h, _fptr(x), n, _fptr(y)), "axx_mlp_forward")
        return y

    def time_forward(self, n: int) -> float:
        ms = C.c_float()
        self.ctx.check(lib().axx_mlp_time_forward(self._h, int(n), C.byref(ms)), "axx_mlp_time_forward")
        return ms.value

    def loss_grad(self, x, t):
        x, t = _f32(x), _f32(t)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        self.ctx.check(lib().axx_mlp_loss_grad(self._h, _fptr(x), _fptr(t), x.shape[0], C.byref(loss),
                                               _fptr(g)), "axx_mlp_loss_grad")
        return loss.value, g

    def set_data(self, x, t):
        x, t = _f32(x), _f32(t)
        self.ctx.check(lib().axx_mlp_set_data(self._h, _fptr(x), _fptr(t), x.shape[0]), "axx_mlp_set_data")

    def train(self, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        trace = np.empty(steps, np.float32)
        self.ctx.check(lib().axx_mlp_train(self._h, int(steps), int(batch), float(lr), int(seed),
                                           _fptr(trace)), "axx_mlp_train")
        return trace

    def close(self):
        if self._h:
            lib().axx_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
