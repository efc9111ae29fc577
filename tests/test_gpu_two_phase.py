"""Phase-field two-phase flow on the device (HIP, AA pattern) vs the FP64 oracle (two-lattice
pull): FP64 to 1e-12 and FP32 to 1e-5 of each lattice's magnitude after the same steps, odd
and even step counts, periodic faces and no-slip walls, density and viscosity contrast,
gravity; wetting walls (contact angle) on faces, edges and corners, and the equilibrium
contact angle of a drop on the device; the three viscosity models."""
import math

import numpy as np
import pytest

from helpers import contact_angle, wetting_drop
from materkit_anisotropiclbm_trt_rbc_facet_amd import mkg
from oracle import oracle as O

pytestmark = pytest.mark.gpu

P, BB = mkg.BC_PERIODIC, mkg.BC_BOUNCE

CASES = {
    "periodic_3d": dict(dims=(12, 10, 9), bc=[P] * 6),
    "walls_y": dict(dims=(10, 12, 8), bc=[P, P, BB, BB, P, P]),
    "walls_xz_2d": dict(dims=(16, 14, 1), bc=[BB, BB, P, P, P, P]),
    # wetting walls: the drop sits on them, so the ghost term phi (1 - phi) is not small
    "wet_3d": dict(dims=(10, 12, 8), bc=[P, P, BB, BB, BB, BB], wet=0b011100, angle=50.0,
                   center=(4.5, 1.0, 1.0)),
    "wet_corner_2d": dict(dims=(16, 14, 1), bc=[BB, BB, BB, BB, P, P], wet=0b001111,
                          angle=130.0, center=(1.0, 1.0, 0.0)),
    # viscosity from mu = rho nu linear or harmonic in phi
    "visc_mu_linear": dict(dims=(10, 12, 8), bc=[P, P, BB, BB, P, P], visc=1),
    "visc_mu_harmonic": dict(dims=(10, 12, 8), bc=[P, P, BB, BB, P, P], visc=2),
    # nz = 64: O steps take the z shift through LDS (the block holds whole z rows)
    "z64_periodic": dict(dims=(5, 3, 64), bc=[P] * 6),  # last block part-filled
    "z64_wet_walls": dict(dims=(4, 3, 64), bc=[P, P, BB, BB, BB, BB], wet=0b110000, angle=70.0,
                          center=(1.5, 1.0, 3.0)),
}


def make(name, prec):
    cfg = CASES[name]
    d = cfg["dims"]
    m = mkg.mp_params(nx=d[0], ny=d[1], nz=d[2], precision=prec, rho_l=3.0, rho_g=1.0,
                      nu_l=0.12, nu_g=0.05, sigma=0.005, width=4.0, mobility=0.08, tau_h=0.9)
    m.gravity[1] = -2e-5
    m.force[0] = 1e-6
    for f in range(6):
        m.bc[f] = cfg["bc"][f]
    m.wet_faces = cfg.get("wet", 0)
    m.contact_angle = cfg.get("angle", 90.0)
    m.visc_model = cfg.get("visc", 0)
    return m


def start(m, seed=6, center=None):
    rng = np.random.default_rng(seed)
    d = (int(m.nx), int(m.ny), int(m.nz))
    c = center or (d[0] / 2.2, d[1] / 1.9, d[2] / 2.0)
    X, Y, Z = np.meshgrid(*[np.arange(n) for n in d], indexing="ij")
    r = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
    phi = (0.5 + 0.5 * np.tanh(2 * (min(d[0], d[1]) / 4.0 - r) / 4.0)).reshape(-1)
    u = 0.01 * rng.standard_normal((phi.size, 3))
    p = 1e-4 * rng.standard_normal(phi.size)
    return phi, u, p


@pytest.mark.parametrize("steps", [5, 6])
@pytest.mark.parametrize("prec,tol", [(mkg.PREC_FP64, 1e-12), (mkg.PREC_FP32, 1e-5)],
                         ids=["fp64", "fp32"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_two_phase_matches_oracle(name, prec, tol, steps):
    m = make(name, prec)
    phi, u, p = start(m, center=CASES[name].get("center"))
    print(f"two-phase {name} prec {prec} steps {steps}", flush=True)
    orc = O.MpState.start(make(name, mkg.PREC_FP64), phi, u=u, p=p)
    dev = mkg.MpSolver(m)
    # the device starts from the oracle's populations (its own init is checked below)
    dev.set(mkg.MP_POP_FLOW, orc.f)
    dev.set(mkg.MP_POP_PHASE, orc.h)
    dev.set(mkg.MP_PHIU, orc.phiu)
    dev.step(steps)
    orc.step(steps)
    gf, gh = dev.get(mkg.MP_POP_FLOW), dev.get(mkg.MP_POP_PHASE)
    mac = orc.macro()
    gp, gu = dev.get(mkg.MP_PRESSURE), dev.get(mkg.MP_VEL).reshape(-1, 3)
    gphi = dev.get(mkg.MP_PHI)
    s = dev.phi_sum()
    dev.close()
    assert np.abs(gf - orc.f).max() < tol * np.abs(orc.f).max()
    assert np.abs(gh - orc.h).max() < tol * np.abs(orc.h).max()
    assert np.abs(gphi - mac["phi"]).max() < tol
    assert np.abs(gu - mac["u"]).max() < tol * max(np.abs(mac["u"]).max(), 1e-3) * 10
    assert np.abs(gp - mac["p"]).max() < tol * np.abs(mac["p"]).max() * 10
    assert abs(s - mac["phi"].sum()) < tol * phi.size


@pytest.mark.parametrize("prec,tol", [(mkg.PREC_FP64, 1e-12), (mkg.PREC_FP32, 1e-5)],
                         ids=["fp64", "fp32"])
def test_two_phase_device_start_matches_oracle_start(prec, tol):
    m = make("periodic_3d", prec)
    phi, u, p = start(m)
    orc = O.MpState.start(make("periodic_3d", mkg.PREC_FP64), phi, u=u, p=p)
    dev = mkg.MpSolver(m)
    dev.init(phi, u, p)
    gf, gh = dev.get(mkg.MP_POP_FLOW), dev.get(mkg.MP_POP_PHASE)
    dev.close()
    assert np.abs(gf - orc.f).max() < tol * np.abs(orc.f).max()
    assert np.abs(gh - orc.h).max() < tol * np.abs(orc.h).max()


def test_two_phase_static_drop_on_device():
    """The static-drop Laplace check at FP32 on the device: dp R / sigma within 5 %."""
    n, sigma, R = 48, 0.01, 12.0
    m = mkg.mp_params(nx=n, ny=n, nz=1, precision=mkg.PREC_FP32, sigma=sigma, width=4.0,
                      mobility=0.1, tau_h=1.0)
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt((X - n / 2 + 0.5) ** 2 + (Y - n / 2 + 0.5) ** 2)
    phi = (0.5 + 0.5 * np.tanh(2 * (R - r) / 4.0)).reshape(-1)
    dev = mkg.MpSolver(m)
    dev.init(phi)
    dev.step(5000)
    p = dev.get(mkg.MP_PRESSURE)
    ph = dev.get(mkg.MP_PHI)
    dev.close()
    dp = p[ph > 0.99].mean() - p[ph < 0.01].mean()
    req = math.sqrt(ph.sum() / math.pi)
    assert abs(dp * req / sigma - 1.0) < 0.05


def test_two_phase_rejects_bad_parameters():
    m = make("periodic_3d", mkg.PREC_FP64)
    m.bc[0] = mkg.BC_DIRICHLET
    with pytest.raises(mkg.MkgError, match="periodic or bounce"):
        mkg.MpSolver(m)
    m = make("periodic_3d", mkg.PREC_FP64)
    m.tau_h = 0.4
    with pytest.raises(mkg.MkgError, match="tau_h"):
        mkg.MpSolver(m)


def test_two_phase_rejects_bad_wetting_parameters():
    m = make("walls_y", mkg.PREC_FP64)
    m.wet_faces = 1 << 0  # x faces are periodic
    with pytest.raises(mkg.MkgError, match="wetting face must be a bounce"):
        mkg.MpSolver(m)
    m = make("walls_y", mkg.PREC_FP64)
    m.wet_faces, m.contact_angle = 1 << 2, 180.0
    with pytest.raises(mkg.MkgError, match="contact_angle"):
        mkg.MpSolver(m)
    m.wet_faces, m.contact_angle = 1 << 6, 60.0
    with pytest.raises(mkg.MkgError, match="bits 0..5"):
        mkg.MpSolver(m)
    m.wet_faces, m.visc_model = 0, 3
    with pytest.raises(mkg.MkgError, match="visc_model"):
        mkg.MpSolver(m)


@pytest.mark.parametrize("theta", [60.0, 120.0])
def test_wetting_drop_on_device_settles_at_the_contact_angle(theta):
    """The oracle's contact-angle pin on the device: FP64 fields agree with the oracle after
    4000 steps and the FP32 drop settles within 3 degrees of theta."""
    print(f"wetting drop theta {theta}", flush=True)
    m64, phi = wetting_drop(theta, prec=mkg.PREC_FP64)
    orc = O.MpState.start(m64, phi)
    orc.step(4000)
    ref = orc.macro()["phi"]
    dev = mkg.MpSolver(m64)
    dev.init(phi)
    dev.step(4000)
    got64 = dev.get(mkg.MP_PHI)
    dev.close()
    assert np.abs(got64 - ref).max() < 1e-9
    m32, _ = wetting_drop(theta, prec=mkg.PREC_FP32)
    dev = mkg.MpSolver(m32)
    dev.init(phi)
    dev.step(4000)
    got32 = dev.get(mkg.MP_PHI)
    dev.close()
    angle, _ = contact_angle(got32.reshape(100, 40))
    assert abs(angle - theta) < 3.0, angle
    assert abs(angle - contact_angle(ref.reshape(100, 40))[0]) < 0.05


@pytest.mark.parametrize("name", ["z64_periodic", "z64_wet_walls"])
def test_two_phase_z_stage_agrees_with_direct_o_steps(name, monkeypatch):
    """The z stage changes only how populations move: FP32 results with and without it agree
    to FP32 rounding."""
    m = make(name, mkg.PREC_FP32)
    phi, u, p = start(m, center=CASES[name].get("center"))
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("MKG_ZSTAGE", flag)
        dev = mkg.MpSolver(m)
        dev.init(phi, u, p)
        dev.step(7)
        out.append((dev.get(mkg.MP_POP_FLOW), dev.get(mkg.MP_POP_PHASE)))
        dev.close()
    for k in range(2):
        scale = np.abs(out[1][k]).max()
        assert np.abs(out[0][k] - out[1][k]).max() <= 1e-6 * scale
