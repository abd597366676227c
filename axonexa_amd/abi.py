"""ctypes mirror of ``include/axonexa.h`` and helpers that pack numpy arrays
into the descriptor structs.

The structs restate, as C data, the inputs the reference passes to
``phaseAcquisitionStream<F>::runAcquisitionLattice``
(src/acquisition/phaseAcquisitionStream.cuh:126-190): the ``Lattice`` cell and
``basis[]`` primitives (src/substrate/lattice.cuh:14-23,
src/primitives/cylinderXY.cuh:16-40, Sphere.cuh:20-50), the acquisition's
gradient functors (src/acquisition/phaseAcquisition.cuh:18-38) and the
``SimuParamsPhase`` run parameters (src/params/simuParams.cuh:34-58).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

AXX_OK = 0
ERRORS = {
    -1: "invalid argument",
    -2: "HIP runtime error",
    -3: "out of memory",
    -4: "size beyond a compiled limit",
    -5: "invalid geometry (overlap or step >= half cell)",
    -6: "rejection initialisation did not finish",
    -7: "no HIP device",
    -8: "call out of order",
}

PRIM_CYLINDER_XY = 0
PRIM_SPHERE = 1
PRIM_PLANE = 2

BOUNDARY_SPECULAR = 0
BOUNDARY_PERMEABLE = 1
BOUNDARY_REJECTION = 2
STEP_ISOTROPIC, STEP_AXES, STEP_GAUSSIAN = 0, 1, 2

INIT_UNIFORM = 0
INIT_COMPARTMENT = 1
INIT_INSIDE_PRIM = 2

MAX_PROFILES = 64
BATCH_MAX_PROFILES = 32     # AXX_BATCH_MAX_PROFILES: profiles per instance of a batched run
MAX_PRIMS = 32768

F_X, F_Y, F_Z, F_DX, F_DY, F_DZ, F_X0, F_Y0, F_Z0, F_E, F_LF = range(11)
F_COMP, F_COMP0, F_LI, F_NCOLL, F_NTRAN, F_NCAP, F_STEP, F_BUD, F_MOM = range(11, 20)


def state_fields(n_profiles: int) -> int:
    return F_MOM + 3 * n_profiles


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class Substrate(C.Structure):
    _fields_ = [
        ("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
        ("D_lat", C.c_double), ("T2_lat", C.c_double),
        ("n_prims", C.c_int),
        ("type", _ip), ("center", _dp), ("radius", _dp),
        ("D", _dp), ("T2", _dp), ("kappa", _dp), ("normal", _dp),
    ]


class Acq(C.Structure):
    _fields_ = [
        ("n_steps", C.c_int), ("dt", C.c_double), ("gamma", C.c_double),
        ("n_profiles", C.c_int), ("profiles", _fp),
        ("n_waveforms", C.c_int), ("wf_profile", _ip),
        ("wf_amplitude", _dp), ("wf_dir", _dp),
    ]


class Run(C.Structure):
    _fields_ = [
        ("n_walkers", C.c_int64), ("seed", C.c_uint64), ("instance", C.c_uint32),
        ("walker_offset", C.c_uint32), ("init_mode", C.c_int), ("init_arg", C.c_int),
        ("boundary_mode", C.c_int), ("relaxation", C.c_int), ("bounce_cap", C.c_int),
        ("step_mode", C.c_int),
    ]


class Tuning(C.Structure):
    _fields_ = [("block", C.c_int), ("bin_width", C.c_double), ("lds_budget", C.c_int),
                ("whole_cell", C.c_int)]


class Instance(C.Structure):
    _fields_ = [("substrate", C.POINTER(Substrate)), ("acq", C.POINTER(Acq)), ("run", C.POINTER(Run))]


class Result(C.Structure):
    _fields_ = [
        ("sum_cos", _dp), ("sum_sin", _dp), ("sum_cos2", _dp), ("sum_sin2", _dp),
        ("intra_cos", _dp), ("intra_sin", _dp),
        ("n_walkers", C.c_int64), ("n_intra", C.c_int64),
        ("collisions", C.c_uint64), ("transmissions", C.c_uint64), ("cap_hits", C.c_uint64),
    ]


def _ptr(arr: Optional[np.ndarray], ctype):
    if arr is None:
        return C.cast(None, C.POINTER(ctype))
    return arr.ctypes.data_as(C.POINTER(ctype))


@dataclass
class SubstrateSpec:
    """A periodic cell and its primitives (Lattice + basis[] of the reference)."""
    a: float
    b: float
    c: float
    D_lat: float
    T2_lat: float = float("inf")
    types: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    centers: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    radii: np.ndarray = field(default_factory=lambda: np.zeros(0))
    D: np.ndarray = 0.0
    T2: np.ndarray = float("inf")
    kappa: np.ndarray = 0.0
    normals: Optional[np.ndarray] = None   # [n, 3] for planes

    def __post_init__(self):
        self.types = np.ascontiguousarray(self.types, dtype=np.int32).reshape(-1)
        n = self.types.size
        self.centers = np.ascontiguousarray(self.centers, dtype=np.float64).reshape(n, 3)
        self.radii = np.ascontiguousarray(self.radii, dtype=np.float64).reshape(n)
        self.D = np.ascontiguousarray(np.broadcast_to(self.D, (n,)), dtype=np.float64)
        self.T2 = np.ascontiguousarray(np.broadcast_to(self.T2, (n,)), dtype=np.float64)
        self.kappa = np.ascontiguousarray(np.broadcast_to(self.kappa, (n,)), dtype=np.float64)
        if self.normals is None:
            self.normals = np.zeros((n, 3))
        self.normals = np.ascontiguousarray(self.normals, dtype=np.float64).reshape(n, 3)

    @property
    def n_prims(self) -> int:
        return int(self.types.size)

    def struct(self) -> Substrate:
        s = Substrate()
        s.a, s.b, s.c = float(self.a), float(self.b), float(self.c)
        s.D_lat, s.T2_lat = float(self.D_lat), float(self.T2_lat)
        s.n_prims = self.n_prims
        s.type = _ptr(self.types, C.c_int)
        s.center = _ptr(self.centers, C.c_double)
        s.radius = _ptr(self.radii, C.c_double)
        s.D = _ptr(self.D, C.c_double)
        s.T2 = _ptr(self.T2, C.c_double)
        s.kappa = _ptr(self.kappa, C.c_double)
        s.normal = _ptr(self.normals, C.c_double)
        s._keep = self  # keep arrays alive with the struct
        return s


@dataclass
class AcqSpec:
    """Sampled unit profiles plus (profile, amplitude, direction) per waveform."""
    n_steps: int
    dt: float
    gamma: float
    profiles: np.ndarray            # [n_steps, P] float32
    wf_profile: np.ndarray          # [M] int32
    wf_amplitude: np.ndarray        # [M] float64
    wf_dir: np.ndarray              # [M, 3] float64

    def __post_init__(self):
        self.profiles = np.ascontiguousarray(self.profiles, dtype=np.float32)
        if self.profiles.ndim == 1:
            self.profiles = self.profiles.reshape(-1, 1)
        assert self.profiles.shape[0] == self.n_steps
        self.wf_profile = np.ascontiguousarray(self.wf_profile, dtype=np.int32).reshape(-1)
        m = self.wf_profile.size
        self.wf_amplitude = np.ascontiguousarray(self.wf_amplitude, dtype=np.float64).reshape(m)
        self.wf_dir = np.ascontiguousarray(self.wf_dir, dtype=np.float64).reshape(m, 3)

    @property
    def n_profiles(self) -> int:
        return int(self.profiles.shape[1])

    @property
    def n_waveforms(self) -> int:
        return int(self.wf_profile.size)

    def struct(self) -> Acq:
        q = Acq()
        q.n_steps, q.dt, q.gamma = int(self.n_steps), float(self.dt), float(self.gamma)
        q.n_profiles = self.n_profiles
        q.profiles = _ptr(self.profiles, C.c_float)
        q.n_waveforms = self.n_waveforms
        q.wf_profile = _ptr(self.wf_profile, C.c_int)
        q.wf_amplitude = _ptr(self.wf_amplitude, C.c_double)
        q.wf_dir = _ptr(self.wf_dir, C.c_double)
        q._keep = self
        return q


@dataclass
class RunSpec:
    n_walkers: int
    seed: int = 1234
    instance: int = 0
    walker_offset: int = 0
    init_mode: int = INIT_UNIFORM
    init_arg: int = 0
    boundary_mode: int = BOUNDARY_SPECULAR
    relaxation: int = 0
    bounce_cap: int = 0
    step_mode: int = 0           # STEP_ISOTROPIC / STEP_AXES / STEP_GAUSSIAN

    def struct(self) -> Run:
        r = Run()
        r.n_walkers = int(self.n_walkers)
        r.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        r.instance = int(self.instance) & 0xFFFFFFFF
        r.walker_offset = int(self.walker_offset) & 0xFFFFFFFF
        r.init_mode, r.init_arg = int(self.init_mode), int(self.init_arg)
        r.boundary_mode, r.relaxation = int(self.boundary_mode), int(self.relaxation)
        r.bounce_cap = int(self.bounce_cap)
        r.step_mode = int(self.step_mode)
        return r


@dataclass
class Signals:
    """Raw and normalised signal sums of one run."""
    sum_cos: np.ndarray
    sum_sin: np.ndarray
    sum_cos2: np.ndarray
    sum_sin2: np.ndarray
    intra_cos: np.ndarray
    intra_sin: np.ndarray
    n_walkers: int
    n_intra: int
    collisions: int
    transmissions: int
    cap_hits: int

    @property
    def E(self) -> np.ndarray:
        """Normalised attenuation |sum e^{i phi}| / N."""
        return np.hypot(self.sum_cos, self.sum_sin) / self.n_walkers

    @property
    def E_re(self) -> np.ndarray:
        return self.sum_cos / self.n_walkers

    @property
    def sigma_re(self) -> np.ndarray:
        """Monte-Carlo standard error of E_re."""
        n = self.n_walkers
        var = np.maximum(self.sum_cos2 / n - (self.sum_cos / n) ** 2, 0.0)
        return np.sqrt(var / n)

    @property
    def sigma_im(self) -> np.ndarray:
        n = self.n_walkers
        var = np.maximum(self.sum_sin2 / n - (self.sum_sin / n) ** 2, 0.0)
        return np.sqrt(var / n)

    @property
    def E_intra(self) -> np.ndarray:
        return np.hypot(self.intra_cos, self.intra_sin) / max(self.n_intra, 1)


def alloc_result(m: int):
    bufs = {k: np.zeros(m) for k in ("sum_cos", "sum_sin", "sum_cos2", "sum_sin2",
                                     "intra_cos", "intra_sin")}
    r = Result()
    for k, v in bufs.items():
        setattr(r, k, _ptr(v, C.c_double))
    return r, bufs


def signals_from(r: Result, bufs) -> Signals:
    return Signals(n_walkers=int(r.n_walkers), n_intra=int(r.n_intra),
                   collisions=int(r.collisions), transmissions=int(r.transmissions),
                   cap_hits=int(r.cap_hits), **{k: v.copy() for k, v in bufs.items()})


def state_view(state: np.ndarray, n_walkers: int, n_profiles: int):
    """Split a flat SoA state buffer (uint32) into named float / int views."""
    st = state.reshape(state_fields(n_profiles), n_walkers)
    f = st.view(np.float32)
    i = st.view(np.int32)
    return {
        "pos": f[F_X:F_Z + 1].T, "disp": f[F_DX:F_DZ + 1].T, "r0": f[F_X0:F_Z0 + 1].T,
        "E": f[F_E], "lf": f[F_LF], "comp": i[F_COMP], "comp0": i[F_COMP0],
        "li": i[F_LI], "ncoll": st[F_NCOLL], "ntran": st[F_NTRAN], "ncap": st[F_NCAP],
        "step": i[F_STEP], "bud": f[F_BUD],
        "mom": f[F_MOM:].reshape(n_profiles, 3, n_walkers).transpose(2, 0, 1),
    }


# ---------------------------------------------------------------- Bloch path
MAG_ROTATE, MAG_RK4, MAG_DOPRI = 0, 1, 2
MAG_GROUP = 8


class Field(C.Structure):
    _fields_ = [
        ("n_steps", C.c_int), ("dt", C.c_double), ("gamma", C.c_double),
        ("n_fields", C.c_int), ("integrator", C.c_int),
        ("substeps", _ip), ("samples", _dp),
        ("origin", C.c_double * 3), ("m0", C.c_double * 3), ("T1", _dp),
    ]


class MagResult(C.Structure):
    _fields_ = [
        ("mx", _dp), ("my", _dp), ("mz", _dp), ("n_walkers", C.c_int64),
        ("collisions", C.c_uint64), ("transmissions", C.c_uint64), ("cap_hits", C.c_uint64),
    ]


def mag_samples_per_step(integrator: int, substeps: np.ndarray) -> np.ndarray:
    s = np.asarray(substeps, dtype=np.int64)
    if integrator == MAG_ROTATE:
        return s
    if integrator == MAG_RK4:
        return 2 * s + 1
    return 6 * s


@dataclass
class FieldSpec:
    """Sampled affine fields B(r, t) = b0(t) + J(t) (r - origin) of the
    magnetization path: samples[k, j] = (b0[3], J[3][3] row major) of field j
    at the k-th sample time (axx_field of include/axonexa.h)."""
    n_steps: int
    dt: float
    gamma: float
    samples: np.ndarray               # [n_samples, n_fields, 12] float64
    integrator: int = MAG_ROTATE
    substeps: Optional[np.ndarray] = None   # [n_steps] int32 or None (= 1)
    origin: Sequence[float] = (0.0, 0.0, 0.0)
    m0: Sequence[float] = (1.0, 0.0, 0.0)
    T1: Optional[np.ndarray] = None   # [n_prims + 1] or None

    def __post_init__(self):
        self.samples = np.ascontiguousarray(self.samples, dtype=np.float64)
        assert self.samples.ndim == 3 and self.samples.shape[2] == 12
        if self.substeps is not None:
            self.substeps = np.ascontiguousarray(self.substeps, dtype=np.int32).reshape(self.n_steps)
        if self.T1 is not None:
            self.T1 = np.ascontiguousarray(self.T1, dtype=np.float64).reshape(-1)
        sub = self.substeps if self.substeps is not None else np.ones(self.n_steps, np.int64)
        need = int(mag_samples_per_step(self.integrator, sub).sum())
        assert self.samples.shape[0] == need, (self.samples.shape, need)

    @property
    def n_fields(self) -> int:
        return int(self.samples.shape[1])

    def struct(self) -> Field:
        f = Field()
        f.n_steps, f.dt, f.gamma = int(self.n_steps), float(self.dt), float(self.gamma)
        f.n_fields, f.integrator = self.n_fields, int(self.integrator)
        f.substeps = _ptr(self.substeps, C.c_int)
        f.samples = _ptr(self.samples, C.c_double)
        for a in range(3):
            f.origin[a] = float(self.origin[a])
            f.m0[a] = float(self.m0[a])
        f.T1 = _ptr(self.T1, C.c_double)
        f._keep = self
        return f


@dataclass
class MagSignals:
    """Sums over walkers of the magnetization per field."""
    mx: np.ndarray
    my: np.ndarray
    mz: np.ndarray
    n_walkers: int
    collisions: int = 0
    transmissions: int = 0
    cap_hits: int = 0

    @property
    def transverse(self) -> np.ndarray:
        """|Mxy| / N (the signal of magAcquisition::calcADC)."""
        return np.hypot(self.mx, self.my) / self.n_walkers


def mag_state_fields(n_fields_in_group: int) -> int:
    return F_MOM + 6 * n_fields_in_group


def mag_from_state(state: np.ndarray, n_walkers: int, F: int) -> np.ndarray:
    """[n_walkers, F, 3] float64 magnetization from a group's SoA state."""
    st = state.reshape(mag_state_fields(F), n_walkers)
    words = st[F_MOM:].reshape(F, 3, 2, n_walkers)
    both = (words[:, :, 0].astype(np.uint64) | (words[:, :, 1].astype(np.uint64) << np.uint64(32)))
    return both.view(np.float64).transpose(2, 0, 1)


class MagInstance(C.Structure):
    _fields_ = [("substrate", C.POINTER(Substrate)), ("field", C.POINTER(Field)), ("run", C.POINTER(Run))]
