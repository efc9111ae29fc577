"""Python binding of the FP64 CPU oracle (``liborc.so``).

TEST INFRASTRUCTURE ONLY: used by ``tests/``, the smoke check and the CPU-baseline leg of
``bench.py`` as the checker.  The product library never imports it.

The oracle is a plain-C restatement of the MaterKit step (see ``mk_oracle.c`` for the
reference file:line each part follows).  State lives in numpy arrays of full-value FP64
populations, laid out as the product's canonical get/set format: ``f[i*N + cell]`` with
``cell = (x*ny + y)*nz + z``.
"""
from __future__ import annotations

import ctypes as ct
import subprocess
from contextlib import contextmanager
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liborc.so"

OP_BGK, OP_TRT, OP_MRT = 0, 1, 2

_lib = None


def build(force: bool = False) -> Path:
    """Compile ``liborc.so`` with the committed Makefile."""
    srcs = ("mk_oracle.c", "mk_aos.c", "mk_oracle.h")
    if force or not LIB.exists() or any(LIB.stat().st_mtime < (HERE / f).stat().st_mtime
                                        for f in srcs):
        subprocess.run(["make", "-s", "-C", str(HERE)] + (["-B"] if force else []), check=True)
    return LIB


def lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        build()
        _lib = _bind(ct.CDLL(str(LIB)))
    return _lib


# Planted mistakes of mk_oracle.c (ORC_MUTANT values; the file lists what each one changes).
MUTANTS = {"thermal_cs2": 1, "latent_sign": 2, "latent_factor": 3, "buoyancy_factor": 4,
           "buoyancy_sign": 5, "solute_cs2": 6, "rejection_sign": 7, "velocity_face_sign": 8}
_mutant_libs: dict = {}


def _build_mutant(name: str) -> ct.CDLL:
    if name not in _mutant_libs:
        import tempfile

        out = Path(tempfile.mkdtemp(prefix="orc_mutant_")) / f"liborc_{name}.so"
        cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
               f"-DORC_MUTANT={MUTANTS[name]}", "-o", str(out), str(HERE / "mk_oracle.c"),
               str(HERE / "mk_aos.c"), "-lm"]
        subprocess.run(cmd, check=True)
        _mutant_libs[name] = _bind(ct.CDLL(str(out)))
    return _mutant_libs[name]


@contextmanager
def planted(name: str):
    """Run the oracle with one planted mistake (a copy built with ORC_MUTANT set) inside the
    block: used by the tests to show that a pin fails on the mistake it is meant to catch."""
    global _lib
    saved = lib()
    _lib = _build_mutant(name)
    try:
        yield
    finally:
        _lib = saved


def _bind(L: ct.CDLL) -> ct.CDLL:
    """Declare the argument and result types of the oracle's entry points."""
    vp = ct.c_void_p
    L.orc_params_size.restype = ct.c_size_t
    L.orc_step.argtypes = [vp, vp, ct.c_int, ct.c_int, ct.c_int]
    L.orc_step.restype = ct.c_int
    L.orc_aos_step.argtypes = [vp, vp, ct.c_int, ct.c_int, ct.POINTER(ct.c_double)]
    L.orc_aos_step.restype = ct.c_int
    L.orc_moments.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.orc_init_equilibrium.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.orc_init_seeds.argtypes = [vp, vp, ct.c_int, vp, vp]
    L.orc_equilibrium.argtypes = [ct.c_int, ct.c_double, vp, vp]
    L.orc_equilibrium_he.argtypes = [ct.c_int, ct.c_double, ct.c_double, vp, vp]
    L.orc_collide_cell.argtypes = [vp, ct.c_int, vp, vp, vp, ct.c_double]
    L.orc_phase_source.argtypes = [vp, vp, ct.c_double, ct.c_int, vp, vp]
    L.orc_face_flux.argtypes = [vp, vp, vp, ct.c_int]
    L.orc_face_flux.restype = ct.c_double
    L.orc_aniso.argtypes = [vp, vp, vp]
    L.orc_aniso.restype = ct.c_double
    L.orc_lattice.argtypes = [ct.c_int, vp, vp, vp, vp]
    L.orc_lattice.restype = ct.c_int
    L.orc_splitmix64.argtypes = [ct.POINTER(ct.c_uint64)]
    L.orc_splitmix64.restype = ct.c_uint64
    L.orc_uniform.argtypes = [ct.POINTER(ct.c_uint64)]
    L.orc_uniform.restype = ct.c_double
    L.orc_cell_uniform.argtypes = [ct.c_uint64, ct.c_uint64]
    L.orc_cell_uniform.restype = ct.c_double
    L.orc_mp_init.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.orc_mp_step.argtypes = [vp, vp, vp, vp, ct.c_int]
    L.orc_mp_step.restype = ct.c_int
    L.orc_mp_macro.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.orc_fvm_step.argtypes = [vp, vp, vp, ct.c_int, ct.c_int]
    L.orc_fvm_step.restype = ct.c_int
    L.orc_fvm_thermal_step.argtypes = [vp, vp, vp, ct.c_int, ct.c_int]
    L.orc_fvm_thermal_step.restype = ct.c_int
    L.orc_mpf_init.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.orc_mpf_step.argtypes = [vp, vp, vp, vp, vp, vp, vp, ct.c_int]
    L.orc_mpf_step.restype = ct.c_int
    L.orc_flow_adjoint.argtypes = [vp, vp, vp, ct.c_int, vp, vp, vp]
    L.orc_flow_adjoint.restype = ct.c_int
    return L


class _State(ct.Structure):
    _fields_ = [("f", ct.c_void_p), ("g", ct.c_void_p), ("h", ct.c_void_p), ("c", ct.c_void_p),
                ("phi", ct.c_void_p), ("phi_prev", ct.c_void_p), ("flag", ct.c_void_p),
                ("grain", ct.c_void_p), ("kappa", ct.c_void_p)]


def _p(a):
    return None if a is None else a.ctypes.data_as(ct.c_void_p)


def lattice(q: int):
    e = np.zeros((q, 3), dtype=np.int32)
    w = np.zeros(q)
    inv = np.zeros(q, dtype=np.int32)
    cs2 = ct.c_double()
    assert lib().orc_lattice(q, _p(e), _p(w), _p(inv), ct.byref(cs2)) == 0
    return e, w, inv, cs2.value


@dataclass
class State:
    """Oracle state for one parameter block (params is an ``MkgParams``-layout struct)."""

    params: object
    f: np.ndarray | None = None
    g: np.ndarray | None = None
    h: np.ndarray | None = None
    c: np.ndarray | None = None
    phi: np.ndarray | None = None
    phi_prev: np.ndarray | None = None
    flag: np.ndarray = field(default=None)
    grain: np.ndarray = field(default=None)
    kappa: np.ndarray | None = None

    @classmethod
    def empty(cls, params) -> "State":
        if lib().orc_params_size() != ct.sizeof(params):
            raise RuntimeError("oracle parameter layout differs from the caller's")
        N = int(params.nx) * int(params.ny) * int(params.nz)
        s = cls(params=params)
        if params.flow_q:
            s.f = np.zeros(params.flow_q * N)
        if params.phase_q:
            s.g = np.zeros(params.phase_q * N)
            s.phi = np.zeros(N)
            s.phi_prev = np.zeros(N)
        if params.thermal:
            s.h = np.zeros(7 * N)
        if params.solute:
            s.c = np.zeros(7 * N)
        s.flag = np.zeros(N, dtype=np.int8)
        s.grain = np.zeros(N, dtype=np.int16)
        if params.has_kappa:
            s.kappa = np.ones(N)
        return s

    @property
    def N(self) -> int:
        p = self.params
        return int(p.nx) * int(p.ny) * int(p.nz)

    def _c(self) -> _State:
        return _State(_p(self.f), _p(self.g), _p(self.h), _p(self.c), _p(self.phi),
                      _p(self.phi_prev), _p(self.flag), _p(self.grain), _p(self.kappa))

    def init_equilibrium(self, rho=None, vel=None, T=None, C=None, phi=None) -> None:
        arrs = [None if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
                for v in (rho, vel, T, C, phi)]
        st = self._c()
        lib().orc_init_equilibrium(ct.byref(self.params), ct.byref(st), *[_p(a) for a in arrs])

    def init_seeds(self, centers, radii) -> None:
        c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1)
        r = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
        st = self._c()
        lib().orc_init_seeds(ct.byref(self.params), ct.byref(st), len(r), _p(c), _p(r))

    def step(self, nsteps: int = 1, mask: int = 15, op: int = OP_BGK) -> None:
        st = self._c()
        rc = lib().orc_step(ct.byref(self.params), ct.byref(st), int(mask), int(op), int(nsteps))
        if rc != 0:
            raise RuntimeError("oracle rejected the parameter block")

    def aos_step(self, nsteps: int = 1, op: int = OP_BGK) -> float:
        """The same steps (all lattices) on the reference-structured AoS path (mk_aos.c);
        returns the seconds of the step loop."""
        st = self._c()
        sec = ct.c_double()
        rc = lib().orc_aos_step(ct.byref(self.params), ct.byref(st), int(op), int(nsteps),
                                ct.byref(sec))
        if rc != 0:
            raise RuntimeError("the AoS path covers periodic flow + phase + thermal cases only")
        return sec.value

    def moments(self) -> dict:
        N = self.N
        p = self.params
        rho = np.zeros(N)
        vel = np.zeros(3 * N)
        T = np.zeros(N)
        Cc = np.zeros(N)
        ph = np.zeros(N)
        st = self._c()
        lib().orc_moments(ct.byref(p), ct.byref(st), _p(rho), _p(vel), _p(T), _p(Cc), _p(ph))
        out = {"vel": vel}
        if p.flow_q:
            out["rho"] = rho
        if p.thermal:
            out["T"] = T
        if p.solute:
            out["C"] = Cc
        if p.phase_q:
            out["phi_sum"] = ph
        return out

    def copy(self) -> "State":
        s = State(params=self.params)
        for k in ("f", "g", "h", "c", "phi", "phi_prev", "flag", "grain", "kappa"):
            v = getattr(self, k)
            setattr(s, k, None if v is None else v.copy())
        return s


def equilibrium(q: int, rho: float, u) -> np.ndarray:
    out = np.zeros(q)
    u = np.ascontiguousarray(u, dtype=np.float64)
    lib().orc_equilibrium(q, rho, _p(u), _p(out))
    return out


def equilibrium_he(q: int, rho: float, rho0: float, u) -> np.ndarray:
    out = np.zeros(q)
    u = np.ascontiguousarray(u, dtype=np.float64)
    lib().orc_equilibrium_he(q, rho, rho0, _p(u), _p(out))
    return out


def collide_cell(params, op: int, f, F, ns: float = 0.0) -> np.ndarray:
    f = np.ascontiguousarray(f, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.float64)
    out = np.zeros_like(f)
    lib().orc_collide_cell(ct.byref(params), op, _p(f), _p(F), _p(out), ns)
    return out


def phase_source(params, phi27, upf: float, grain: int = 0):
    phi27 = np.ascontiguousarray(phi27, dtype=np.float64).reshape(-1)
    S = ct.c_double()
    a = ct.c_double()
    lib().orc_phase_source(ct.byref(params), _p(phi27), upf, grain, ct.byref(S), ct.byref(a))
    return S.value, a.value


def face_flux(params, g, axis: int, R=None) -> float:
    g = np.ascontiguousarray(g, dtype=np.float64)
    Rm = None if R is None else np.ascontiguousarray(R, dtype=np.float64).reshape(-1)
    return lib().orc_face_flux(ct.byref(params), _p(Rm), _p(g), int(axis))



def aniso(params, n):
    """a_s(n) and d a_s / d n of the parameter block's anisotropy form at the crystal-frame n."""
    n = np.ascontiguousarray(n, dtype=np.float64)
    D = np.zeros(3)
    a = lib().orc_aniso(ct.byref(params), _p(n), _p(D))
    return a, D

def splitmix_stream(seed: int, n: int) -> np.ndarray:
    s = ct.c_uint64(seed)
    return np.array([lib().orc_uniform(ct.byref(s)) for _ in range(n)])


def cell_uniform(seed: int, cells) -> np.ndarray:
    return np.array([lib().orc_cell_uniform(seed, int(c)) for c in cells])


@dataclass
class MpfState:
    """Oracle state of the multi-phase field (params is an ``MpfParams``-layout struct):
    g n x 19 x N, phi / phi_prev (n + 1) x N, h 7 x N (thermal), c 7 x N (solute), flag N."""

    params: object
    g: np.ndarray = None
    phi: np.ndarray = None
    phi_prev: np.ndarray = None
    h: np.ndarray | None = None
    c: np.ndarray | None = None
    flag: np.ndarray = None

    @classmethod
    def start(cls, params, phij, T=None, C=None) -> "MpfState":
        N = int(params.nx) * int(params.ny) * int(params.nz)
        n = int(params.n)
        s = cls(params=params, g=np.zeros(n * 19 * N), phi=np.zeros((n + 1) * N),
                phi_prev=np.zeros((n + 1) * N), flag=np.zeros(N, dtype=np.int8))
        if params.thermal:
            s.h = np.zeros(7 * N)
        if params.solute:
            s.c = np.zeros(7 * N)
        ph = np.ascontiguousarray(phij, dtype=np.float64).reshape(-1)
        t = None if T is None else np.ascontiguousarray(T, dtype=np.float64).reshape(-1)
        cc = None if C is None else np.ascontiguousarray(C, dtype=np.float64).reshape(-1)
        lib().orc_mpf_init(ct.byref(params), _p(ph), _p(t), _p(cc), _p(s.g), _p(s.phi),
                           _p(s.phi_prev), _p(s.h), _p(s.c))
        return s

    def step(self, nsteps: int = 1) -> None:
        rc = lib().orc_mpf_step(ct.byref(self.params), _p(self.g), _p(self.phi),
                                _p(self.phi_prev), _p(self.h), _p(self.c), _p(self.flag),
                                int(nsteps))
        if rc != 0:
            raise RuntimeError("oracle rejected the multi-phase parameter block")


def fvm_step(params, phi, c, nsteps: int = 1, evolve_phase: bool = True) -> None:
    """Explicit finite-volume phase field + solute (in place on float64 arrays of N values)."""
    assert phi.dtype == np.float64 and c.dtype == np.float64
    assert phi.flags.c_contiguous and c.flags.c_contiguous
    rc = lib().orc_fvm_step(ct.byref(params), _p(phi), _p(c), int(nsteps), 1 if evolve_phase else 0)
    if rc != 0:
        raise RuntimeError("oracle rejected the finite-volume parameter block")


def fvm_thermal_step(params, phi, T, nsteps: int = 1, evolve_phase: bool = True) -> None:
    """Explicit finite-volume phase field driven by T and the heat equation with the latent
    source (in place on float64 arrays of N values)."""
    assert phi.dtype == np.float64 and T.dtype == np.float64
    assert phi.flags.c_contiguous and T.flags.c_contiguous
    rc = lib().orc_fvm_thermal_step(ct.byref(params), _p(phi), _p(T), int(nsteps),
                                    1 if evolve_phase else 0)
    if rc != 0:
        raise RuntimeError("oracle rejected the finite-volume parameter block")


@dataclass
class MpState:
    """Oracle state of the two-phase flow (params is an ``MpParams``-layout struct):
    f, h 19 x N full-value populations, phiu 3 x N (phi u of the last step, cell-interleaved)."""

    params: object
    f: np.ndarray = None
    h: np.ndarray = None
    phiu: np.ndarray = None

    @classmethod
    def start(cls, params, phi, u=None, p=None) -> "MpState":
        N = int(params.nx) * int(params.ny) * int(params.nz)
        s = cls(params=params, f=np.zeros(19 * N), h=np.zeros(19 * N), phiu=np.zeros(3 * N))
        ph = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        uu = None if u is None else np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        pp = None if p is None else np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        lib().orc_mp_init(ct.byref(params), _p(ph), _p(uu), _p(pp), _p(s.f), _p(s.h),
                          _p(s.phiu))
        return s

    def step(self, nsteps: int = 1) -> None:
        rc = lib().orc_mp_step(ct.byref(self.params), _p(self.f), _p(self.h), _p(self.phiu),
                               int(nsteps))
        if rc != 0:
            raise RuntimeError("oracle rejected the two-phase parameter block")

    def macro(self) -> dict:
        N = self.f.size // 19
        phi, mu, p, u = np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(3 * N)
        lib().orc_mp_macro(ct.byref(self.params), _p(self.f), _p(self.h), _p(phi), _p(mu),
                           _p(p), _p(u))
        return {"phi": phi, "mu": mu, "p": p, "u": u.reshape(N, 3)}


def flow_adjoint(params, flag, f0, nsteps: int, u_target):
    """Discrete adjoint of nsteps flow-only steps (orc_flow_adjoint): returns the loss
    1/2 sum |u_N - u_target|^2 over non-wall cells and the gradient with respect to
    (F_x, F_y, F_z, s_nu, s_bulk, s_odd, s_even)."""
    fl = np.ascontiguousarray(flag, dtype=np.int8)
    f = np.ascontiguousarray(f0, dtype=np.float64)
    ut = np.ascontiguousarray(u_target, dtype=np.float64).reshape(-1)
    loss = ct.c_double()
    grad = np.zeros(7)
    rc = lib().orc_flow_adjoint(ct.byref(params), _p(fl), _p(f), int(nsteps), _p(ut),
                                ct.byref(loss), _p(grad))
    if rc != 0:
        raise RuntimeError("oracle rejected the flow-adjoint configuration")
    return loss.value, grad
