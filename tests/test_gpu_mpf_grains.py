"""Karma-Rappel grains of the multi-phase field on the device vs the FP64 oracle: per-grain
anisotropy form and orientation, growth driven by T (thermal lattice with latent heat) or U
(solute lattice with the anti-trapping current), pairwise interaction (mkMultiphaseField,
MaterKit/inc/mkMultiphaseField.h:31-45).

FP64 storage agrees to 1e-12 relative to the field magnitude, FP32 deviation storage to 1e-5.
Odd and even step counts cover both AA layouts; the z stage of the O steps runs at nz = 64
and nz = 256; walls and non-periodic faces take the face-rule body.
"""
import numpy as np
import pytest

from helpers import grain_fields, grain_params, quat_rot
from materkit_anisotropiclbm_trt_rbc_facet_amd import mkg
from oracle import oracle as O

pytestmark = pytest.mark.gpu

R1 = quat_rot([0.9, 0.2, -0.3, 0.25])
R2 = quat_rot([0.3, -0.7, 0.1, 0.6])
R3 = quat_rot([0.5, 0.5, 0.5, -0.5])

CASES = {
    # name: (dims, forms, rotations, coupling, thermal lattice beside the solute, walls)
    "thermal_two": ((12, 11, 10), ["hcp", "cubic"], [R1, R2], 1, False, False),
    "solute_two": ((12, 10, 9), ["cbc", "facet"], [R2, R1], 2, False, False),
    "solute_thermal_three": ((11, 10, 9), ["cubic", "hcp", "facet"], [R1, R2, R3], 2, True,
                             False),
    "one_grain_walls": ((12, 11, 8), ["hcp"], [R3], 1, False, True),
    "z64_thermal": ((5, 4, 64), ["cubic", "hcp"], [R1, R3], 1, False, False),
    "z256_solute_thermal": ((3, 3, 256), ["facet", "cbc"], [R2, R1], 2, True, False),
}


def start(name, prec, seed=3):
    dims, forms, rots, coupling, both, walls = CASES[name]
    n = len(forms)
    m = grain_params(dims, n, forms, rots, coupling=coupling, prec=prec)
    if both:
        m.thermal = 1
        m.set(bc_h=[0, 0, 0, 0, 2, 2], T_face=[0, 0, 0, 0, -0.5, -0.6])
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    centers = [[nx * (0.3 + 0.4 * rng.random()), ny * (0.3 + 0.4 * rng.random()),
                nz * (0.3 + 0.4 * rng.random())] for _ in range(n)]
    radii = [max(1.5, min(nx, ny, nz) / 4.0)] * n
    phi, T, C = grain_fields(m, centers, radii, seed=seed)
    if m.thermal and T is None:
        T = -0.55 + 0.01 * rng.standard_normal(m.ncell)
    orc = O.MpfState.start(m, phi, T, C)
    # populations off equilibrium, the noise weighted to the interfaces
    wgt = (1.0 - phi ** 2)[:, None, :]
    orc.g += (5e-5 * rng.standard_normal((n, 19, m.ncell)) * (0.05 + wgt)).reshape(-1)
    if orc.h is not None:
        orc.h += 1e-4 * rng.standard_normal(orc.h.shape)
    if orc.c is not None:
        orc.c += 1e-4 * rng.standard_normal(orc.c.shape)
    if walls:
        f3 = orc.flag.reshape(dims)
        f3[nx // 3: nx // 3 + 2, ny // 3: ny // 3 + 3, : nz // 2] = mkg.IS_WALLS
    dev = mkg.MpfSolver(m)
    dev.init(phi, T, C)
    dev.set(mkg.MPF_FLAG, orc.flag)
    dev.set(mkg.MPF_POPS, orc.g)
    if orc.h is not None:
        dev.set(mkg.MPF_THERMAL, orc.h)
    if orc.c is not None:
        dev.set(mkg.MPF_SOLUTE, orc.c)
    dev.set(mkg.MPF_PHI, orc.phi)
    dev.set(mkg.MPF_PHI_PREV, orc.phi_prev)
    return m, dev, orc


def errors(dev, orc):
    out = {"g": float(np.abs(dev.get(mkg.MPF_POPS) - orc.g).max() / np.abs(orc.g).max()),
           "phi": float(np.abs(dev.get(mkg.MPF_PHI) - orc.phi).max())}
    if orc.h is not None:
        out["h"] = float(np.abs(dev.get(mkg.MPF_THERMAL) - orc.h).max() / np.abs(orc.h).max())
    if orc.c is not None:
        out["c"] = float(np.abs(dev.get(mkg.MPF_SOLUTE) - orc.c).max() / np.abs(orc.c).max())
    return out


@pytest.mark.parametrize("prec,tol", [(mkg.PREC_FP64, 1e-12), (mkg.PREC_FP32, 1e-5)],
                         ids=["fp64", "fp32"])
@pytest.mark.parametrize("steps", [5, 6])
@pytest.mark.parametrize("name", list(CASES))
def test_grains_match_oracle(name, steps, prec, tol):
    print(f"[grains] {name} steps {steps} prec {prec}", flush=True)
    m, dev, orc = start(name, prec)
    dev.step(steps)
    orc.step(steps)
    e = errors(dev, orc)
    dev.close()
    print(f"[grains] {name} {e}", flush=True)
    assert all(v < tol for v in e.values()), e


def test_grains_long_run_against_fp64_oracle():
    """Three separated grains (cubic, hcp, facet; rotated) growing for 400 steps into a melt
    held below the liquidus (U = -0.3), with latent heat on the thermal lattice, until they
    impinge.  FP64 storage stays within 1e-12 of the oracle; FP32 deviation storage keeps every
    grain's phi within 3e-5 and the combined phi within 1e-4 (the FP32 rounding accumulates
    about 4e-8 per step)."""
    dims = (24, 22, 16)
    ms = []
    for prec in (mkg.PREC_FP64, mkg.PREC_FP32):
        m = grain_params(dims, 3, ["cubic", "hcp", "facet"], [R1, R2, R3], coupling=2,
                         prec=prec)
        m.thermal = 1
        m.set(bc_h=[0, 0, 0, 0, 2, 2], T_face=[0, 0, 0, 0, -0.5, -0.6])
        ms.append(m)
    phi, _, C = grain_fields(ms[0], [[6.0, 6.0, 8.0], [17.5, 7.0, 8.0], [12.0, 16.5, 8.0]],
                             [4.5, 4.5, 4.5])
    T = np.full(ms[0].ncell, -0.55)
    orc = O.MpfState.start(ms[0], phi, T, C)
    orc.step(400)
    N = ms[0].ncell
    solid0, grown = (phi > 0).sum(), (orc.phi[:3 * N] > 0).sum()
    err = []
    for m in ms:
        dev = mkg.MpfSolver(m)
        dev.init(phi, T, C)
        dev.step(400)
        e = np.abs(dev.get(mkg.MPF_PHI) - orc.phi).reshape(4, N).max(1)
        dev.close()
        err.append(e)
    print(f"[grains] 400 steps, phi max error per grain and combined: fp64 {err[0]}, "
          f"fp32 {err[1]}; solid cells {solid0} -> {grown}", flush=True)
    assert grown > 1.2 * solid0
    assert err[0].max() < 1e-12
    assert err[1][:3].max() < 3e-5 and err[1][3] < 1e-4


def test_grain_parameters_are_checked():
    m = grain_params((4, 4, 4), 2, ["cubic", "hcp"], [R1, R2], coupling=1)
    m.aniso_kind[1] = 7
    with pytest.raises(mkg.MkgError, match="aniso_kind"):
        mkg.MpfSolver(m)
    m = grain_params((4, 4, 4), 2, ["cubic", "hcp"], [R1, 2.0 * R2], coupling=1)
    with pytest.raises(mkg.MkgError, match="orthogonal"):
        mkg.MpfSolver(m)
    m = grain_params((4, 4, 4), 2, ["cubic", "hcp"], [R1, R2], coupling=2)
    m.solute = 0
    with pytest.raises(mkg.MkgError, match="coupling 2"):
        mkg.MpfSolver(m)
    m = grain_params((4, 4, 4), 2, ["cubic", "hcp"], [R1, R2], coupling=2)
    m.model = mkg.MPF_TEMPLATE
    with pytest.raises(mkg.MkgError, match="solute lattice"):
        mkg.MpfSolver(m)
