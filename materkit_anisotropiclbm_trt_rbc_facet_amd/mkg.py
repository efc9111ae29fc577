"""ctypes binding of the C ABI in ``include/mkg.h`` (``libmkg.so``).

The library is loaded from this package directory.  There is no CPU fallback: if the
library is missing or no GPU is present, the calls fail loudly.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmkg.so"

MAX_GRAINS = 64

# flags (MaterKit/inc/mkGlobalDefs.h:140-145)
IS_OUTLET, IS_INLET, IS_FLUID, IS_INTER, IS_SOLID, IS_WALLS = -2, -1, 0, 1, 2, 3
OP_BGK, OP_TRT, OP_MRT = 0, 1, 2
PREC_FP64, PREC_FP32 = 0, 1
BC_PERIODIC, BC_BOUNCE, BC_DIRICHLET, BC_OUTFLOW, BC_PRESSURE, BC_VELOCITY = 0, 1, 2, 3, 4, 5
ANISO_CUBIC, ANISO_CBC, ANISO_HCP, ANISO_FACET = 0, 1, 2, 3
PF_KARMA_RAPPEL, PF_ALLEN_CAHN = 0, 1
FLOW, PHASE, THERMAL, SOLUTE = 1, 2, 4, 8
ALL = FLOW | PHASE | THERMAL | SOLUTE
F_POP_FLOW, F_POP_PHASE, F_POP_THERMAL, F_POP_SOLUTE = 0, 1, 2, 3
F_RHO, F_VEL, F_TEMP, F_CONC, F_PHI, F_PHI_PREV, F_FLAG, F_GRAIN, F_KAPPA = range(4, 13)
DT_F64, DT_F32, DT_I8, DT_I16 = 0, 1, 2, 3
(R_MASS, R_ENERGY, R_SOLUTE, R_SOLID, R_PHI, R_NUSSELT, R_MOMENTUM, R_UMAX, R_NONFINITE,
 R_TIP) = range(10)

PARAM_FIELDS = [
    ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
    ("flow_q", C.c_int32), ("phase_q", C.c_int32), ("thermal", C.c_int32),
    ("solute", C.c_int32), ("precision", C.c_int32), ("flow_op", C.c_int32),
    ("pf_coupling", C.c_int32), ("ngrains", C.c_int32), ("has_kappa", C.c_int32),
    ("bc_f", C.c_int32 * 6), ("bc_g", C.c_int32 * 6), ("bc_h", C.c_int32 * 6),
    ("bc_c", C.c_int32 * 6), ("flow_incomp", C.c_int32),
    ("s_nu", C.c_double), ("s_bulk", C.c_double), ("s_odd", C.c_double),
    ("s_even", C.c_double), ("rho0", C.c_double), ("force", C.c_double * 3),
    ("buoy", C.c_double * 3), ("T_ref", C.c_double), ("u_face", (C.c_double * 3) * 6),
    ("alpha_l", C.c_double), ("alpha_s", C.c_double), ("latent_T", C.c_double),
    ("T_face", C.c_double * 6),
    ("D_l", C.c_double), ("D_s", C.c_double), ("k_part", C.c_double),
    ("latent_C", C.c_double), ("q_min", C.c_double), ("at_coef", C.c_double),
    ("C_face", C.c_double * 6),
    ("pf_dt", C.c_double), ("pf_wdx2", C.c_double), ("eps4", C.c_double),
    ("lambda_", C.c_double), ("phi_s", C.c_double),
    ("aniso_kind", C.c_int32), ("pf_model", C.c_int32), ("aniso", C.c_double * 2),
    ("ref_f", C.c_double), ("ref_g", C.c_double), ("ref_h", C.c_double),
    ("ref_c", C.c_double),
    ("rot", (C.c_double * 9) * MAX_GRAINS),
    ("wall_omega", C.c_double * 3), ("wall_center", C.c_double * 3),
]


class MkgParams(C.Structure):
    """Layout-identical mirror of ``mkg_params``."""

    _fields_ = PARAM_FIELDS

    def copy(self) -> "MkgParams":
        out = type(self)()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(self))
        return out

    def set(self, **kw) -> "MkgParams":
        """Assign fields by name; arrays may be partial (leading elements are filled)."""
        for k, v in kw.items():
            if k == "lambda":
                k = "lambda_"
            cur = getattr(self, k)
            if isinstance(cur, C.Array):
                flat = np.asarray(v).reshape(-1)
                if isinstance(cur[0], C.Array):
                    inner = len(cur[0])
                    for n in range(flat.size):
                        cur[n // inner][n % inner] = float(flat[n])
                else:
                    conv = float if isinstance(cur[0], float) else int
                    for n in range(flat.size):
                        cur[n] = conv(flat[n])
            else:
                setattr(self, k, v)
        return self

    @property
    def ncell(self) -> int:
        return int(self.nx) * int(self.ny) * int(self.nz)


FVM_PHI, FVM_CONC, FVM_U, FVM_TEMP = 0, 1, 2, 3
MP_POP_FLOW, MP_POP_PHASE, MP_PHI, MP_MU, MP_PRESSURE, MP_VEL, MP_PHIU = range(7)
MP_VISC_NU_LINEAR, MP_VISC_MU_LINEAR, MP_VISC_MU_HARMONIC = range(3)
MPF_MAX = 4
MPF_POPS, MPF_PHI, MPF_PHI_PREV, MPF_THERMAL, MPF_FLAG, MPF_SOLUTE = range(6)
MPF_TEMPLATE, MPF_KARMA_RAPPEL = 0, 1

MPF_FIELDS = [
    ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("n", C.c_int32),
    ("precision", C.c_int32), ("thermal", C.c_int32),
    ("bc", C.c_int32 * 6), ("bc_h", C.c_int32 * 6),
    ("dt", C.c_double), ("wdx", C.c_double),
    ("u", (C.c_double * 3) * MPF_MAX), ("kappa", C.c_double * MPF_MAX),
    ("growth", C.c_double), ("lambda_", C.c_double), ("eps", C.c_double),
    ("alpha", C.c_double), ("latent_T", C.c_double), ("T_ref", C.c_double),
    ("T_face", C.c_double * 6), ("phi_face", C.c_double * 6),
    ("model", C.c_int32), ("coupling", C.c_int32), ("aniso_kind", C.c_int32 * MPF_MAX),
    ("eps4", C.c_double * MPF_MAX), ("aniso", (C.c_double * 2) * MPF_MAX),
    ("rot", (C.c_double * 9) * MPF_MAX),
    ("solute", C.c_int32), ("bc_c", C.c_int32 * 6),
    ("D_l", C.c_double), ("D_s", C.c_double), ("k_part", C.c_double), ("q_min", C.c_double),
    ("latent_C", C.c_double), ("at_coef", C.c_double), ("C_face", C.c_double * 6),
]


MP_FIELDS = [
    ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("precision", C.c_int32),
    ("bc", C.c_int32 * 6),
    ("rho_l", C.c_double), ("rho_g", C.c_double), ("nu_l", C.c_double), ("nu_g", C.c_double),
    ("sigma", C.c_double), ("width", C.c_double), ("mobility", C.c_double),
    ("tau_h", C.c_double), ("gravity", C.c_double * 3), ("force", C.c_double * 3),
    ("wet_faces", C.c_int32), ("visc_model", C.c_int32), ("contact_angle", C.c_double),
]


class MpParams(C.Structure):
    """Layout-identical mirror of ``mkg_mp_params`` (two-phase flow)."""

    _fields_ = MP_FIELDS

    @property
    def ncell(self) -> int:
        return int(self.nx) * int(self.ny) * int(self.nz)


def mp_params(**kw) -> "MpParams":
    """Two-phase flow parameters with the defaults of mkg_mp_params_default (host-side copy,
    usable without the library)."""
    m = MpParams()
    m.nx = m.ny = m.nz = 1
    m.rho_l, m.rho_g, m.nu_l, m.nu_g = 1.0, 1.0, 0.1, 0.1
    m.sigma, m.width, m.mobility, m.tau_h = 0.01, 4.0, 0.1, 1.0
    m.contact_angle = 90.0
    MkgParams.set(m, **kw)
    return m


class MpfParams(C.Structure):
    """Layout-identical mirror of ``mkg_mpf_params`` (multi-phase field)."""

    _fields_ = MPF_FIELDS

    copy = MkgParams.copy
    set = MkgParams.set

    @property
    def ncell(self) -> int:
        return int(self.nx) * int(self.ny) * int(self.nz)


_lib = None


class MkgError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load ``libmkg.so`` (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise MkgError(f"{LIB_PATH} not built; run materkit_anisotropiclbm_trt_rbc_facet_amd.build")
    L = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "mkg_abi_version": (C.c_int, []),
        "mkg_params_size": (C.c_size_t, []),
        "mkg_last_error": (C.c_char_p, []),
        "mkg_params_default": (C.c_int, [P(MkgParams)]),
        "mkg_params_from_xml": (C.c_int, [C.c_char_p, P(MkgParams), P(C.c_int32)]),
        "mkg_seeds_from_xml": (C.c_int, [C.c_char_p, C.c_int, vp, vp, vp]),
        "mkg_create": (C.c_int, [P(MkgParams), P(vp)]),
        "mkg_destroy": (C.c_int, [vp]),
        "mkg_get_params": (C.c_int, [vp, P(MkgParams)]),
        "mkg_set_params": (C.c_int, [vp, P(MkgParams)]),
        "mkg_init_equilibrium": (C.c_int, [vp, vp, vp, vp, vp, vp]),
        "mkg_init_seeds": (C.c_int, [vp, C.c_int, vp, vp]),
        "mkg_random_seeds": (C.c_int, [C.c_uint64, C.c_int, C.c_double, C.c_int, C.c_int,
                                       C.c_int, vp, vp]),
        "mkg_quat_to_rot": (C.c_int, [vp, vp]),
        "mkg_step": (C.c_int, [vp, C.c_int, C.c_int]),
        "mkg_sync": (C.c_int, [vp]),
        "mkg_step_count": (C.c_int64, [vp, C.c_int]),
        "mkg_get_field": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_size_t]),
        "mkg_set_field": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_size_t]),
        "mkg_reduce": (C.c_int, [vp, C.c_int, vp]),
        "mkg_reduce_noalloc": (C.c_int, [vp, C.c_int, vp]),
        "mkg_checkpoint_save": (C.c_int, [vp, C.c_char_p]),
        "mkg_checkpoint_load": (C.c_int, [vp, C.c_char_p]),
        "mkg_last_timing": (C.c_int, [vp, vp]),
        "mkg_enable_timing": (C.c_int, [vp, C.c_int]),
        "mkg_adjoint_run": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]),
        "mkg_flow_adjoint_run": (C.c_int, [vp, C.c_int, vp, vp, vp]),
        "mkg_mpf_params_size": (C.c_size_t, []),
        "mkg_mpf_params_default": (C.c_int, [P(MpfParams)]),
        "mkg_mpf_create": (C.c_int, [P(MpfParams), P(vp)]),
        "mkg_mpf_destroy": (C.c_int, [vp]),
        "mkg_mpf_init": (C.c_int, [vp, vp, vp]),
        "mkg_mpf_init_fields": (C.c_int, [vp, vp, vp, vp]),
        "mkg_aniso_eval": (C.c_int, [C.c_int, C.c_double, vp, vp, vp, vp, vp]),
        "mkg_stream_copy_rate": (C.c_int, [C.c_size_t, C.c_int, vp]),
        "mkg_mpf_step": (C.c_int, [vp, C.c_int]),
        "mkg_mpf_sync": (C.c_int, [vp]),
        "mkg_mpf_get": (C.c_int, [vp, C.c_int, vp, C.c_size_t]),
        "mkg_mpf_set": (C.c_int, [vp, C.c_int, vp, C.c_size_t]),
        "mkg_mpf_sums": (C.c_int, [vp, vp]),
        "mkg_mp_params_size": (C.c_size_t, []),
        "mkg_mp_params_default": (C.c_int, [P(MpParams)]),
        "mkg_mp_create": (C.c_int, [P(MpParams), P(vp)]),
        "mkg_mp_destroy": (C.c_int, [vp]),
        "mkg_mp_init": (C.c_int, [vp, vp, vp, vp]),
        "mkg_mp_step": (C.c_int, [vp, C.c_int]),
        "mkg_mp_sync": (C.c_int, [vp]),
        "mkg_mp_get": (C.c_int, [vp, C.c_int, vp, C.c_size_t]),
        "mkg_mp_set": (C.c_int, [vp, C.c_int, vp, C.c_size_t]),
        "mkg_mp_sums": (C.c_int, [vp, vp]),
        "mkg_mp_params_from_xml": (C.c_int, [C.c_char_p, P(MpParams), P(C.c_int32)]),
        "mkg_fvm_create": (C.c_int, [P(MkgParams), P(vp)]),
        "mkg_fvm_destroy": (C.c_int, [vp]),
        "mkg_fvm_set": (C.c_int, [vp, C.c_int, vp, C.c_size_t]),
        "mkg_fvm_step": (C.c_int, [vp, C.c_int, C.c_int]),
        "mkg_fvm_sync": (C.c_int, [vp]),
        "mkg_fvm_get": (C.c_int, [vp, C.c_int, vp, C.c_size_t]),
        "mkg_fvm_sums": (C.c_int, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.mkg_params_size() != C.sizeof(MkgParams):
        raise MkgError("mkg_params layout mismatch between libmkg.so and mkg.py")
    if L.mkg_mp_params_size() != C.sizeof(MpParams):
        raise MkgError("mkg_mp_params layout mismatch between libmkg.so and mkg.py")
    if L.mkg_mpf_params_size() != C.sizeof(MpfParams):
        raise MkgError("mkg_mpf_params layout mismatch between libmkg.so and mkg.py")
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise MkgError(lib().mkg_last_error().decode())


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_params(**kw) -> MkgParams:
    p = MkgParams()
    _check(lib().mkg_params_default(C.byref(p)))
    p.set(**kw)
    return p


def params_from_xml(path: str | Path) -> tuple[MkgParams, dict]:
    p = MkgParams()
    info = (C.c_int32 * 4)()
    _check(lib().mkg_params_from_xml(str(path).encode(), C.byref(p), info))
    return p, {"numLoops": info[0], "numFrames": info[1], "seeds": info[2]}


def seeds_from_xml(path: str | Path, max_seeds: int = MAX_GRAINS):
    centers = np.zeros((max_seeds, 3))
    radii = np.zeros(max_seeds)
    rot = np.zeros((max_seeds, 9))
    n = lib().mkg_seeds_from_xml(str(path).encode(), max_seeds, _ptr(centers), _ptr(radii),
                                  _ptr(rot))
    if n < 0:
        raise MkgError(lib().mkg_last_error().decode())
    return centers[:n].copy(), radii[:n].copy(), rot[:n].copy()


def random_seeds(seed: int, n: int, radius: float, dims) -> tuple[np.ndarray, np.ndarray]:
    centers = np.zeros((n, 3))
    quats = np.zeros((n, 4))
    _check(lib().mkg_random_seeds(seed, n, radius, int(dims[0]), int(dims[1]), int(dims[2]),
                                  _ptr(centers), _ptr(quats)))
    return centers, quats


def aniso_eval(kind: int, normal, eps4: float = 0.0, par=(0.0, 0.0), rot=None):
    """a_s and n_v = R a_s (D - n (n.D)) of one crystal at the lab-frame normal (the
    anisotropy query of mkPhaseField / mkMultiphaseField)."""
    nrm = np.ascontiguousarray(normal, dtype=np.float64)
    pp = np.ascontiguousarray(par, dtype=np.float64)
    R = None if rot is None else np.ascontiguousarray(rot, dtype=np.float64).reshape(-1)
    a = C.c_double()
    nv = np.zeros(3)
    rc = lib().mkg_aniso_eval(int(kind), float(eps4), _ptr(pp), _ptr(R), _ptr(nrm),
                              C.byref(a), _ptr(nv))
    _check(0 if rc == 1 else rc)  # 1: zero normal, isotropic
    return a.value, nv


def stream_copy_rate(nbytes: int = 1 << 30, reps: int = 10) -> float:
    """GB/s (read + write) of the library's plain copy kernel on the current device."""
    out = C.c_double()
    _check(lib().mkg_stream_copy_rate(int(nbytes), int(reps), C.byref(out)))
    return out.value


def quat_to_rot(q) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.float64)
    R = np.zeros(9)
    _check(lib().mkg_quat_to_rot(_ptr(q), _ptr(R)))
    return R.reshape(3, 3)


class Solver:
    """One device-resident case (wraps an ``mkg_handle``)."""

    def __init__(self, params: MkgParams):
        self.params = params.copy()
        h = C.c_void_p()
        _check(lib().mkg_create(C.byref(self.params), C.byref(h)))
        self._h = h
        self.N = self.params.ncell

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mkg_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state
    def init_equilibrium(self, rho=None, vel=None, T=None, C=None, phi=None) -> None:
        arrs = [None if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
                for v in (rho, vel, T, C, phi)]
        _check(lib().mkg_init_equilibrium(self._h, *[_ptr(a) for a in arrs]))

    def init_seeds(self, centers, radii) -> None:
        c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1)
        r = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
        _check(lib().mkg_init_seeds(self._h, len(r), _ptr(c), _ptr(r)))

    def set_params(self, params: MkgParams) -> None:
        _check(lib().mkg_set_params(self._h, C.byref(params)))
        self.params = params.copy()

    def step(self, nsteps: int = 1, mask: int = ALL) -> None:
        _check(lib().mkg_step(self._h, int(nsteps), int(mask)))

    def sync(self) -> None:
        _check(lib().mkg_sync(self._h))

    def step_count(self, bit: int) -> int:
        return int(lib().mkg_step_count(self._h, bit))

    def _count(self, field: int) -> int:
        p = self.params
        N = self.N
        return {F_POP_FLOW: p.flow_q * N, F_POP_PHASE: p.phase_q * N, F_POP_THERMAL: 7 * N,
                F_POP_SOLUTE: 7 * N, F_VEL: 3 * N}.get(field, N)

    def get(self, field: int, dtype=np.float64) -> np.ndarray:
        n = self._count(field)
        if field == F_FLAG:
            out = np.empty(n, dtype=np.int8)
            dt = DT_I8
        elif field == F_GRAIN:
            out = np.empty(n, dtype=np.int16)
            dt = DT_I16
        else:
            out = np.empty(n, dtype=dtype)
            dt = DT_F64 if out.dtype == np.float64 else DT_F32
        _check(lib().mkg_get_field(self._h, field, _ptr(out), dt, n))
        return out

    def set(self, field: int, arr) -> None:
        if field == F_FLAG:
            a = np.ascontiguousarray(arr, dtype=np.int8).reshape(-1)
            dt = DT_I8
        elif field == F_GRAIN:
            a = np.ascontiguousarray(arr, dtype=np.int16).reshape(-1)
            dt = DT_I16
        else:
            a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
            dt = DT_F64
        _check(lib().mkg_set_field(self._h, field, _ptr(a), dt, a.size))

    def reduce(self, which: int) -> np.ndarray:
        out = np.zeros(4)
        _check(lib().mkg_reduce(self._h, which, _ptr(out)))
        return out[:3] if which in (R_MOMENTUM, R_TIP) else out[:1]

    def reduce_noalloc(self, which: int) -> np.ndarray:
        """mkg_reduce_noalloc: the same reduction with no device allocation."""
        out = np.zeros(4)
        _check(lib().mkg_reduce_noalloc(self._h, which, _ptr(out)))
        return out[:3] if which in (R_MOMENTUM, R_TIP) else out[:1]

    def checkpoint_save(self, path) -> None:
        _check(lib().mkg_checkpoint_save(self._h, str(path).encode()))

    def checkpoint_load(self, path) -> None:
        _check(lib().mkg_checkpoint_load(self._h, str(path).encode()))

    def enable_timing(self, on: bool = True) -> None:
        _check(lib().mkg_enable_timing(self._h, 1 if on else 0))

    def last_timing(self) -> dict:
        out = np.zeros(4)
        _check(lib().mkg_last_timing(self._h, _ptr(out)))
        return {"phase_moments_ms": out[0], "fused_ms": out[1], "total_ms": out[2],
                "steps": int(out[3])}

    def adjoint_run(self, nsteps: int, ckpt: int, target, want_kappa: bool = True) -> dict:
        """Forward nsteps with checkpoints, FP64 loss 1/2 sum (phi_N - target)^2 and its
        gradient w.r.t. eps4 and (when the case has a conductivity field) kappa."""
        t = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
        if t.size != self.N:
            raise MkgError("target must have N values")
        loss = C.c_double()
        deps = C.c_double()
        dk = np.zeros(self.N) if want_kappa else None
        times = np.zeros(4)
        _check(lib().mkg_adjoint_run(self._h, int(nsteps), int(ckpt), _ptr(t), C.byref(loss),
                                     C.byref(deps), _ptr(dk), _ptr(times)))
        return {"loss": loss.value, "dL_deps4": deps.value, "dL_dkappa": dk,
                "forward_ms": times[0], "recompute_ms": times[1], "reverse_ms": times[2],
                "total_ms": times[3]}

    def flow_adjoint_run(self, nsteps: int, u_target) -> dict:
        """Forward nsteps flow-only steps, FP64 loss 1/2 sum |u_N - u_target|^2 over non-wall
        cells and its gradient with respect to the body force and the four collision rates."""
        t = np.ascontiguousarray(u_target, dtype=np.float64).reshape(-1)
        if t.size != 3 * self.N:
            raise MkgError("u_target must have 3 N values")
        loss = C.c_double()
        grad = np.zeros(7)
        _check(lib().mkg_flow_adjoint_run(self._h, int(nsteps), _ptr(t), C.byref(loss),
                                          _ptr(grad)))
        return {"loss": loss.value, "dL_dforce": grad[:3].copy(),
                "dL_drates": dict(zip(("s_nu", "s_bulk", "s_odd", "s_even"), grad[3:]))}

    def phi_now(self) -> np.ndarray:
        """phi = sum_i g_i of the current pre-collision phase populations."""
        g = self.get(F_POP_PHASE)
        return g.reshape(int(self.params.phase_q), self.N).sum(0)



def default_mpf_params(**kw) -> MpfParams:
    m = MpfParams()
    _check(lib().mkg_mpf_params_default(C.byref(m)))
    m.set(**kw)
    return m


class MpfSolver:
    """Multi-phase field on the device (wraps an ``mkg_mpf_handle``): n order parameters on
    D3Q19 plus an optional D3Q7 thermal lattice (mkFluidsUser::propagate / relaxation)."""

    def __init__(self, params: MpfParams):
        self.params = params.copy()
        h = C.c_void_p()
        _check(lib().mkg_mpf_create(C.byref(self.params), C.byref(h)))
        self._h = h
        self.N = self.params.ncell
        self.n = int(self.params.n)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mkg_mpf_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init(self, phi, T=None, C=None) -> None:
        """Equilibrium start from phi (n x N), T and the solute c (N each, None: T_ref, 0)."""
        ph = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        if ph.size != self.n * self.N:
            raise MkgError("phi must hold n x N values")
        t = None if T is None else np.ascontiguousarray(T, dtype=np.float64).reshape(-1)
        if C is None:
            _check(lib().mkg_mpf_init(self._h, _ptr(ph), _ptr(t)))
            return
        c = np.ascontiguousarray(C, dtype=np.float64).reshape(-1)
        if c.size != self.N or (t is not None and t.size != self.N):
            raise MkgError("T and C must hold N values")
        _check(lib().mkg_mpf_init_fields(self._h, _ptr(ph), _ptr(t), _ptr(c)))

    def step(self, nsteps: int = 1) -> None:
        _check(lib().mkg_mpf_step(self._h, int(nsteps)))

    def sync(self) -> None:
        _check(lib().mkg_mpf_sync(self._h))

    def _count(self, which: int) -> int:
        return {MPF_POPS: self.n * 19 * self.N, MPF_PHI: (self.n + 1) * self.N,
                MPF_PHI_PREV: (self.n + 1) * self.N, MPF_THERMAL: 7 * self.N,
                MPF_FLAG: self.N, MPF_SOLUTE: 7 * self.N}[which]

    def get(self, which: int) -> np.ndarray:
        n = self._count(which)
        out = np.empty(n, dtype=np.int8 if which == MPF_FLAG else np.float64)
        _check(lib().mkg_mpf_get(self._h, which, _ptr(out), n))
        return out

    def set(self, which: int, arr) -> None:
        dt = np.int8 if which == MPF_FLAG else np.float64
        a = np.ascontiguousarray(arr, dtype=dt).reshape(-1)
        _check(lib().mkg_mpf_set(self._h, which, _ptr(a), a.size))

    def sums(self) -> np.ndarray:
        out = np.zeros(self.n + 1)
        _check(lib().mkg_mpf_sums(self._h, _ptr(out)))
        return out


class FvmSolver:
    """Explicit finite-volume phase field + solute with anti-trapping current on the device
    (wraps an ``mkg_fvm_handle``; evolution_anisotropy / evolution_ATC1-2, or with
    ``thermal = 1`` evolution_thermal: the phase driven by T and the heat equation)."""

    def __init__(self, params: MkgParams):
        self.params = params.copy()
        h = C.c_void_p()
        _check(lib().mkg_fvm_create(C.byref(self.params), C.byref(h)))
        self._h = h
        self.N = self.params.ncell

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mkg_fvm_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set(self, which: int, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        _check(lib().mkg_fvm_set(self._h, which, _ptr(a), a.size))

    def get(self, which: int) -> np.ndarray:
        out = np.empty(self.N)
        _check(lib().mkg_fvm_get(self._h, which, _ptr(out), self.N))
        return out

    def step(self, nsteps: int = 1, evolve_phase: bool = True) -> None:
        _check(lib().mkg_fvm_step(self._h, int(nsteps), 1 if evolve_phase else 0))

    def sync(self) -> None:
        _check(lib().mkg_fvm_sync(self._h))

    def sums(self) -> np.ndarray:
        out = np.zeros(2)
        _check(lib().mkg_fvm_sums(self._h, _ptr(out)))
        return out


class MpSolver:
    """Phase-field two-phase flow on the device (wraps an ``mkg_mp_handle``;
    mkFluidsMultiphase, Liang et al. 2014 model)."""

    def __init__(self, params: MpParams):
        self.params = MpParams()
        C.memmove(C.byref(self.params), C.byref(params), C.sizeof(params))
        h = C.c_void_p()
        _check(lib().mkg_mp_create(C.byref(self.params), C.byref(h)))
        self._h = h
        self.N = self.params.ncell

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mkg_mp_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init(self, phi, u=None, p=None) -> None:
        arrs = [None if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
                for v in (phi, u, p)]
        _check(lib().mkg_mp_init(self._h, *[_ptr(a) for a in arrs]))

    def step(self, nsteps: int = 1) -> None:
        _check(lib().mkg_mp_step(self._h, int(nsteps)))

    def sync(self) -> None:
        _check(lib().mkg_mp_sync(self._h))

    def get(self, which: int) -> np.ndarray:
        n = {MP_POP_FLOW: 19 * self.N, MP_POP_PHASE: 19 * self.N, MP_VEL: 3 * self.N}.get(which, self.N)
        out = np.empty(n)
        _check(lib().mkg_mp_get(self._h, which, _ptr(out), n))
        return out

    def set(self, which: int, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        _check(lib().mkg_mp_set(self._h, which, _ptr(a), a.size))

    def phi_sum(self) -> float:
        out = np.zeros(1)
        _check(lib().mkg_mp_sums(self._h, _ptr(out)))
        return float(out[0])


def mp_params_from_xml(path) -> tuple:
    """Two-phase parameters of a case file's <multiphases> block and its run info."""
    m = MpParams()
    info = (C.c_int32 * 4)()
    _check(lib().mkg_mp_params_from_xml(str(path).encode(), C.byref(m), info))
    return m, {"numLoops": info[0], "numFrames": info[1], "seeds": info[2],
               "init": "drops" if info[3] else "layered"}
