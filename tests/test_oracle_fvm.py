"""Pins of the oracle's explicit finite-volume solvers (evolution_anisotropy,
MaterKit/inc/mkPhaseField.h:51; evolution_ATC1/2, MaterKit/inc/mkSolutalField.h:44-46):

- a Fourier mode of U in a liquid decays by exactly (1 - 4 D sin^2(pi m / n))^steps per step
  (explicit Euler on the 7-point stencil);
- c - phi/2 is conserved to round-off with periodic and zero-flux faces;
- a uniform phase field follows phi' = phi + dt (phi - phi^3 - lambda U (1 - phi^2)^2);
- a small phase mode around phi = 0 grows by 1 + dt (1 - 4 W^2/dx^2 sin^2(pi m / n)) (the
  W^2/dx^2 scaling and the Laplacian of the face-flux divergence at eps4 = 0);
- the anti-trapping current makes U continuous across a growing planar alloy front and the
  solid composition about k times the liquid's, against more trapping without it.
"""
import math

import numpy as np
import pytest

from materkit_anisotropiclbm_trt_rbc_facet_amd import mkg
from oracle import oracle as O


def params(dims, **kw):
    p = mkg.default_params(nx=dims[0], ny=dims[1], nz=dims[2], phase_q=19, solute=1,
                           pf_coupling=2, eps4=0.0, pf_dt=0.008, pf_wdx2=6.25, D_l=0.1, D_s=0.0,
                           k_part=0.15, q_min=1e-3, at_coef=0.0, phi_s=0.99, **{"lambda": 0.0})
    p.set(**kw)
    return p


def A(k, phi):
    return 0.5 * ((1 + k) - (1 - k) * phi)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fourier_mode_of_U_decays_exactly(axis):
    n, m, D, steps = 32, 3, 0.1, 50
    dims = [2, 2, 2]
    dims[axis] = n
    p = params(dims, D_l=D)
    shape = tuple(dims)
    x = np.arange(n)
    mode = np.cos(2 * np.pi * m * x / n)
    U0 = np.broadcast_to(np.moveaxis(np.broadcast_to(mode, shape[:axis] + shape[axis + 1:] + (n,)),
                                     -1, axis), shape).reshape(-1).copy()
    phi = np.full(U0.size, -1.0)
    c = U0 * A(0.15, -1.0)
    O.fvm_step(p, phi, c, steps, evolve_phase=False)
    g = (1 - 4 * D * math.sin(math.pi * m / n) ** 2) ** steps
    np.testing.assert_allclose(c / A(0.15, -1.0), g * U0, rtol=0, atol=1e-13)


@pytest.mark.parametrize("bcs", ["periodic", "zero_flux"])
def test_c_minus_half_phi_is_conserved(bcs):
    dims = (14, 12, 10)
    p = params(dims, eps4=0.05, at_coef=0.5 / math.sqrt(2), **{"lambda": 3.19})
    if bcs == "zero_flux":
        p.set(bc_g=[1, 1, 0, 0, 1, 1], bc_c=[1, 1, 0, 0, 1, 1])
    rng = np.random.default_rng(2)
    X, Y, Z = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    r = np.sqrt((X - 7) ** 2 + (Y - 5.5) ** 2 + (Z - 4.5) ** 2)
    phi = np.ascontiguousarray(-np.tanh((r - 3.5) / (math.sqrt(2) * 2.5)).reshape(-1))
    U = -0.5 + 0.05 * rng.standard_normal(phi.size)
    c = U * A(0.15, phi)
    inv0 = (c - 0.5 * phi).sum()
    O.fvm_step(p, phi, c, 60)
    assert abs((c - 0.5 * phi).sum() - inv0) < 1e-12 * phi.size
    assert np.abs(phi).max() < 1.2


def test_uniform_phase_follows_the_double_well_and_coupling():
    dims = (3, 3, 3)
    lam = 3.19
    for phi0, U in ((0.3, -0.4), (-0.7, 0.2), (0.95, -0.55)):
        p = params(dims, **{"lambda": lam})
        phi = np.full(27, phi0)
        c = np.full(27, U * A(0.15, phi0))
        O.fvm_step(p, phi, c, 1)
        want = phi0 + 0.008 * (phi0 - phi0 ** 3 - lam * U * (1 - phi0 ** 2) ** 2)
        np.testing.assert_allclose(phi, want, rtol=0, atol=1e-15)


def test_small_phase_mode_grows_with_the_discrete_laplacian():
    n, m, eps = 24, 2, 1e-7
    p = params((n, 1, 1))
    x = np.arange(n)
    phi = eps * np.cos(2 * np.pi * m * x / n)
    c = np.zeros(n)
    O.fvm_step(p, phi, c, 1)
    g = 1 + 0.008 * (1 - 4 * 6.25 * math.sin(math.pi * m / n) ** 2)
    np.testing.assert_allclose(phi, g * eps * np.cos(2 * np.pi * m * x / n), rtol=0, atol=1e-19)


def _fvm_front(at_coef, steps=12000):
    nz, W, k = 256, 2.5, 0.15
    p = params((1, 1, nz), at_coef=at_coef, **{"lambda": 3.19})
    z = np.arange(nz) + 0.5 - nz / 2
    phi = np.ascontiguousarray(-np.tanh((np.abs(z) - 16) / (math.sqrt(2) * W)))
    c = -0.5 * A(k, phi)
    O.fvm_step(p, phi, c, steps)
    U = c / A(k, phi)
    i = nz // 2 + np.where((phi[nz // 2:-1] > 0) & (phi[nz // 2 + 1:] <= 0))[0][0]
    front = z[i] + phi[i] / (phi[i] - phi[i + 1])
    band = (z > 16 + 3 * W) & (z < front - 3 * W)
    assert band.sum() >= 15
    comp = c + A(k, phi) / (1 - k)
    ratio = comp[band].mean() / np.interp(front + 3 * W, z, comp)
    return U[band].mean(), np.interp(front, z, U), ratio


def test_fvm_anti_trapping_current_removes_solute_trapping():
    at = 0.5 / math.sqrt(2.0)
    u_at, u_front_at, r_at = _fvm_front(at)
    u_none, u_front_none, r_none = _fvm_front(0.0)
    assert abs(u_at - u_front_at) < 0.05
    assert u_none - u_front_none > 0.1
    assert u_at < u_none
    assert abs(r_at - 0.15) < 0.02 and r_none > r_at + 0.01


# ---- finite-volume phase field driven by T (mkSolutalField::evolution_thermal,
#      MaterKit/inc/mkSolutalField.h:47-48: "potential is temperature")

def tparams(dims, **kw):
    p = mkg.default_params(nx=dims[0], ny=dims[1], nz=dims[2], phase_q=19, thermal=1,
                           pf_coupling=1, eps4=0.0, pf_dt=0.008, pf_wdx2=6.25, alpha_l=0.1,
                           alpha_s=0.1, latent_T=0.5, phi_s=0.99, **{"lambda": 0.0})
    p.set(**kw)
    return p


@pytest.mark.parametrize("axis", [0, 2])
def test_thermal_fourier_mode_decays_exactly(axis):
    """With phi held fixed in the liquid, a mode of T decays by (1 - 4 alpha sin^2(pi m/n)) per
    step: the face diffusivity is alpha, not 2 alpha or alpha/2."""
    n, m, al, steps = 32, 3, 0.12, 40
    dims = [2, 2, 2]
    dims[axis] = n
    p = tparams(dims, alpha_l=al, alpha_s=0.03)
    shape = tuple(dims)
    mode = np.cos(2 * np.pi * m * np.arange(n) / n)
    T0 = np.broadcast_to(np.moveaxis(np.broadcast_to(mode, shape[:axis] + shape[axis + 1:] + (n,)),
                                     -1, axis), shape).reshape(-1).copy()
    phi = np.full(T0.size, -1.0)
    T = T0.copy()
    O.fvm_thermal_step(p, phi, T, steps, evolve_phase=False)
    g = (1 - 4 * al * math.sin(math.pi * m / n) ** 2) ** steps
    np.testing.assert_allclose(T, g * T0, rtol=0, atol=1e-13)
    # in the solid the rate is alpha_s
    phi[:] = 1.0
    T = T0.copy()
    O.fvm_thermal_step(p, phi, T, steps, evolve_phase=False)
    g = (1 - 4 * 0.03 * math.sin(math.pi * m / n) ** 2) ** steps
    np.testing.assert_allclose(T, g * T0, rtol=0, atol=1e-13)


@pytest.mark.parametrize("bcs", ["periodic", "zero_flux"])
def test_T_minus_latent_phi_is_conserved_while_a_seed_grows(bcs):
    """T - L_T phi is conserved to round-off: the latent source is +L_T (phi' - phi), and a
    growing seed (phi rising) heats the melt by the same amount."""
    dims = (14, 12, 10)
    p = tparams(dims, eps4=0.05, alpha_s=0.06, latent_T=0.7, **{"lambda": 3.19})
    if bcs == "zero_flux":
        p.set(bc_g=[1, 1, 0, 0, 1, 1], bc_h=[1, 1, 0, 0, 1, 1])
    X, Y, Z = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    r = np.sqrt((X - 7) ** 2 + (Y - 5.5) ** 2 + (Z - 4.5) ** 2)
    phi = np.ascontiguousarray(-np.tanh((r - 3.5) / (math.sqrt(2) * 2.5)).reshape(-1))
    T = np.full(phi.size, -0.55)
    inv0 = (T - 0.7 * phi).sum()
    phi_sum0, T_sum0 = phi.sum(), T.sum()
    O.fvm_thermal_step(p, phi, T, 60)
    assert abs((T - 0.7 * phi).sum() - inv0) < 1e-12 * phi.size
    assert phi.sum() > phi_sum0 + 1.0  # the seed grows into the undercooled melt ...
    assert T.sum() > T_sum0 + 0.7      # ... and releases its latent heat


def test_thermal_dirichlet_faces_hold_the_exact_linear_profile():
    """Between Dirichlet faces T_lo, T_hi (half a cell outside the first and last centres) the
    linear profile through the face values is a fixed point of the update."""
    n = 20
    p = tparams((n, 1, 1), alpha_l=0.15, bc_h=[2, 2, 0, 0, 0, 0], bc_g=[1, 1, 0, 0, 0, 0],
                T_face=[-0.2, -0.8, 0, 0, 0, 0])
    x = np.arange(n) + 0.5
    T = -0.2 + (-0.8 + 0.2) * x / n
    phi = np.full(n, -1.0)
    T1 = T.copy()
    O.fvm_thermal_step(p, phi, T1, 5, evolve_phase=False)
    np.testing.assert_allclose(T1, T, rtol=0, atol=1e-15)
    # a flat start relaxes to it
    T2 = np.full(n, -0.5)
    O.fvm_thermal_step(p, phi, T2, 20000, evolve_phase=False)
    np.testing.assert_allclose(T2, T, rtol=0, atol=1e-10)


def test_uniform_phase_follows_the_double_well_and_T_coupling():
    """u = T drives the phase (not T / A(phi) as for the solute)."""
    dims = (3, 3, 3)
    lam = 3.19
    for phi0, T0 in ((0.3, -0.4), (-0.7, 0.2), (0.95, -0.55)):
        p = tparams(dims, **{"lambda": lam})
        phi = np.full(27, phi0)
        T = np.full(27, T0)
        O.fvm_thermal_step(p, phi, T, 1)
        want = phi0 + 0.008 * (phi0 - phi0 ** 3 - lam * T0 * (1 - phi0 ** 2) ** 2)
        np.testing.assert_allclose(phi, want, rtol=0, atol=1e-15)
        np.testing.assert_allclose(T, T0 + 0.5 * (want - phi0), rtol=0, atol=1e-15)
