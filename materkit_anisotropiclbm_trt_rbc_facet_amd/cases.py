"""Case set-up: parameters from the BoltzTechs XML files under ``cases/`` plus the initial
fields each benchmark configuration prescribes.

``setup(target, case)`` works on anything with ``init_equilibrium`` / ``init_seeds`` (the
device ``Solver`` or any object with the same two methods), so every consumer starts from
identical initial fields.
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import mkg

REPO = Path(__file__).resolve().parent.parent
CASES_DIR = REPO / "cases"

NAMES = {
    "C1": "C1_poiseuille.xml",
    "C2": "C2_rayleigh_benard.xml",
    "C3": "C3_free_dendrite.xml",
    "C4": "C4_solutal_multidendrite.xml",
    "C5": "C5_adjoint_target.xml",
    "fused256": "fused_flow_thermal_dendrite_256.xml",
}


@dataclass
class Case:
    name: str
    path: Path
    params: mkg.MkgParams
    info: dict
    seeds: tuple | None = None          # (centers, radii, rotations)
    perturbation: tuple | None = None   # (amplitude, rng seed)
    conductivity: tuple | None = None   # (sigma, correlation length, rng seed)
    extra: dict = field(default_factory=dict)


def _solver_tokens(path: Path, tag: str):
    root = ET.parse(path).getroot()
    node = root.find(f"solver/{tag}")
    if node is None or node.text is None:
        return None
    return [float(t) for t in node.text.split()]


def load(name: str, **overrides) -> Case:
    """Load a case by short name (``C1`` .. ``C5``, ``fused256``) or path; keyword overrides
    are applied to the parameter block (e.g. ``nx=32, precision=mkg.PREC_FP64``)."""
    path = Path(NAMES.get(name, name))
    if not path.is_absolute():
        path = CASES_DIR / path
    p, info = mkg.params_from_xml(path)
    case = Case(name=name, path=path, params=p, info=info)
    if info["seeds"] > 0 and p.phase_q:
        c, r, R = mkg.seeds_from_xml(path)
        case.seeds = (c, r, R)
        if p.ngrains:
            p.set(rot=R.reshape(-1))
    pert = _solver_tokens(path, "perturbation")
    if pert:
        case.perturbation = (pert[0], int(pert[1]))
    cond = _solver_tokens(path, "conductivityField")
    if cond:
        case.conductivity = (cond[0], cond[1], int(cond[2]))
        p.has_kappa = 1
    if overrides:
        scale = {}
        for k in ("nx", "ny", "nz"):
            if k in overrides:
                scale[k] = overrides[k] / getattr(p, k)
        p.set(**overrides)
        if case.seeds is not None and scale:
            c, r, R = case.seeds
            c = c.copy()
            for a, k in enumerate(("nx", "ny", "nz")):
                if k in scale:
                    c[:, a] = (c[:, a] + 0.5) * scale[k] - 0.5
            case.seeds = (c, r, R)
    return case


def cell_uniform(seed: int, n: int) -> np.ndarray:
    """Counter-based uniform [0,1) per cell index: splitmix64 of seed ^ (cell * K)."""
    k = np.uint64(0xD1B54A32D192ED03)
    with np.errstate(over="ignore"):
        s = np.uint64(seed) ^ (np.arange(n, dtype=np.uint64) * k)
        z = s + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def lognormal_field(shape, sigma: float, corr: float, seed: int) -> np.ndarray:
    """Log-normal multiplier exp(sigma * g) with g a unit-variance Gaussian field of
    correlation length ``corr`` cells (white noise filtered by a periodic Gaussian kernel)."""
    n = int(np.prod(shape))
    u1 = cell_uniform(seed, n)
    u2 = cell_uniform(seed ^ 0x5DEECE66D, n)
    white = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    white = white.reshape(shape)
    k = [np.fft.fftfreq(s) for s in shape]
    kx, ky, kz = np.meshgrid(*k, indexing="ij")
    filt = np.exp(-0.5 * (2 * np.pi * corr) ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    g = np.real(np.fft.ifftn(np.fft.fftn(white) * filt))
    g = (g - g.mean()) / (g.std() + 1e-300)
    return np.exp(sigma * g)


def initial_fields(case: Case) -> dict:
    """Macroscopic initial fields of a case (None = library default)."""
    p = case.params
    nx, ny, nz = int(p.nx), int(p.ny), int(p.nz)
    N = nx * ny * nz
    out = {"rho": None, "vel": None, "T": None, "C": None, "phi": None}
    if p.thermal:
        if case.perturbation is not None:
            Th, Tc = p.T_face[4], p.T_face[5]
            z = np.arange(nz)
            prof = Th - (Th - Tc) * (z + 0.5) / nz
            T = np.broadcast_to(prof, (nx, ny, nz)).reshape(-1).copy()
            amp, seed = case.perturbation
            T += amp * (2.0 * cell_uniform(seed, N) - 1.0)
            out["T"] = T
        else:
            out["T"] = np.full(N, p.T_ref)
    return out


def setup(target, case: Case) -> None:
    """Initial state on ``target`` (device Solver or an object with the same methods)."""
    p = case.params
    f = initial_fields(case)
    if p.phase_q and case.seeds is not None:
        c, r, _ = case.seeds
        target.init_seeds(c, r)
    if p.solute:
        # c = A(phi) U0 with the U0 of the case (composition/iniValue)
        U0 = p.ref_c
        if p.phase_q and case.seeds is not None:
            phi = seed_phi(case)
            A = 0.5 * ((1 + p.k_part) - (1 - p.k_part) * phi)
            f["C"] = A * U0
        else:
            f["C"] = np.full(p.ncell, U0)
    target.init_equilibrium(rho=f["rho"], vel=f["vel"], T=f["T"], C=f["C"], phi=None)


def seed_phi(case: Case) -> np.ndarray:
    """phi of the seed geometry (same formula as the device seed kernel)."""
    p = case.params
    nx, ny, nz = int(p.nx), int(p.ny), int(p.nz)
    c, r, _ = case.seeds
    width = math.sqrt(2.0) * math.sqrt(p.pf_wdx2)
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    best = np.full((nx, ny, nz), np.inf)
    for k in range(len(r)):
        d = np.sqrt((X - c[k, 0]) ** 2 + (Y - c[k, 1]) ** 2 + (Z - c[k, 2]) ** 2) - r[k]
        best = np.minimum(best, d)
    return (-np.tanh(best / width)).reshape(-1)


# ---------------------------------------------------------------- multi-phase field
NAMES["MPF"] = "multi_phase_field.xml"


def mpf_seed_fields(dims, centers, radii, wdx: float) -> np.ndarray:
    """One order parameter per seed: phi_j = -tanh((|x - c_j| - r_j) / (sqrt 2 W0/dx)), the
    equilibrium profile of the conservative Allen-Cahn equation (+1 inside).  Returns
    n x N values, cell index (x*ny + y)*nz + z."""
    nx, ny, nz = (int(d) for d in dims)
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    width = math.sqrt(2.0) * wdx
    out = []
    for c, r in zip(np.asarray(centers, dtype=float).reshape(-1, 3), np.asarray(radii).reshape(-1)):
        d = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - r
        out.append(-np.tanh(d / width).reshape(-1))
    return np.stack(out)


def load_mpf(name: str = "MPF", precision: int = mkg.PREC_FP64, **overrides):
    """Multi-phase-field case from a BoltzTechs XML file (run_multiPhaseField): grid, dx, dt,
    W0 and the seeds (one order parameter per seed).  Returns (MpfParams, phi n x N).  The
    order parameters' velocities and interaction strengths are those of the
    mkFluidsUser::relaxation template (u_1 = 0.02 in x, kappa_0 = 0, kappa_j = 1)."""
    path = Path(NAMES.get(name, name))
    if not path.is_absolute():
        path = CASES_DIR / path
    p, info = mkg.params_from_xml(path)
    centers, radii, _ = mkg.seeds_from_xml(path)
    n = len(radii)
    if not 1 <= n <= mkg.MPF_MAX:
        raise mkg.MkgError(f"{path}: {n} seeds; the multi-phase field holds 1..{mkg.MPF_MAX}")
    m = mkg.default_mpf_params(nx=p.nx, ny=p.ny, nz=p.nz, n=n, precision=precision,
                               dt=p.pf_dt, wdx=math.sqrt(p.pf_wdx2))
    m.thermal = 1 if p.thermal else 0
    if m.thermal:
        m.alpha = p.alpha_l
        m.latent_T = p.latent_T
        m.T_ref = p.T_ref
    m.set(**overrides)
    phi = mpf_seed_fields((m.nx, m.ny, m.nz), centers, radii, m.wdx)
    return m, phi
