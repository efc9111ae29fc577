"""Block order of the launches (MKG_XCD_CHUNK): the remap that gives each XCD runs of
consecutive blocks only permutes which hardware block takes which 256 cells, so every handle
computes the same fields with runs of 16 (the FP32 default) and in hardware order
(MKG_XCD_CHUNK=0, the FP64 default), in both precisions.

The grids hold a multiple of 128 blocks (32 x 32 x 64 and 32 x 32 x 32 cells; 64 x 64 x 64
for the finite-volume march kernels, whose blocks cover 8 x columns of 4 whole z rows), the
condition under which the remap with runs of 16 is active.  Per-cell fields must agree bit for bit;
sums over blocks (adjoint gradients with respect to scalars) are added in another order and
agree to rounding."""
import math

import numpy as np
import pytest

from helpers import flow_adjoint_case, load_device, random_state
from materkit_anisotropiclbm_trt_rbc_facet_amd import mkg

pytestmark = pytest.mark.gpu

ORDERS = ("0", "16")


def both_orders(monkeypatch, run):
    out = []
    for k in ORDERS:
        monkeypatch.setenv("MKG_XCD_CHUNK", k)
        out.append(run())
    monkeypatch.delenv("MKG_XCD_CHUNK")
    return out


def fused_params(dims, prec, walls, solute=False):
    kw = dict(solute=1, pf_coupling=2, D_l=0.1, D_s=0.01, k_part=0.15, q_min=1e-3,
              at_coef=0.5 / math.sqrt(2.0)) if solute else \
        dict(thermal=1, alpha_l=0.1, alpha_s=0.06, latent_T=0.5, pf_coupling=1, T_ref=-0.55)
    p = mkg.default_params(nx=dims[0], ny=dims[1], nz=dims[2], flow_q=19, phase_q=19,
                           precision=prec, eps4=0.05, pf_dt=0.008, pf_wdx2=6.25,
                           **{"lambda": 3.19}, **kw)
    om = 1.0 / 0.8
    p.s_nu = p.s_bulk = p.s_even = om
    p.s_odd = 1.0 / ((3.0 / 16.0) / (1.0 / om - 0.5) + 0.5)
    p.buoy[2] = 1e-3
    if walls:
        p.set(bc_f=[1, 3, 0, 0, 1, 1], bc_g=[1, 1, 0, 0, 1, 1], bc_h=[2, 2, 0, 0, 1, 1],
              bc_c=[1, 1, 0, 0, 1, 1], T_face=[-0.5, -0.6, 0, 0, 0, 0])
    return p


@pytest.mark.parametrize("dims,prec,walls,solute", [
    ((32, 32, 64), mkg.PREC_FP32, False, False),  # all-periodic body, z stage
    ((32, 32, 64), mkg.PREC_FP64, True, False),   # face rules and wall cells, z stage
    ((32, 32, 32), mkg.PREC_FP32, True, True),    # solute lattice, direct O steps
], ids=["fp32_periodic_z64", "fp64_walls_z64", "fp32_solute_walls_z32"])
def test_fused_step_is_independent_of_block_order(monkeypatch, dims, prec, walls, solute):
    p = fused_params(dims, prec, walls, solute)
    st = random_state(p, seed=3, with_walls=walls)
    print(f"[block order] fused {dims} prec {prec} walls {walls}", flush=True)

    def run():
        d = load_device(p, st)
        d.step(9)
        r = [d.get(f) for f in (mkg.F_POP_FLOW, mkg.F_POP_PHASE,
                                mkg.F_POP_SOLUTE if solute else mkg.F_POP_THERMAL,
                                mkg.F_PHI, mkg.F_FLAG)]
        d.close()
        return r

    a, b = both_orders(monkeypatch, run)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_two_phase_step_is_independent_of_block_order(monkeypatch):
    d = (32, 32, 64)
    m = mkg.mp_params(nx=d[0], ny=d[1], nz=d[2], precision=mkg.PREC_FP32, rho_l=3.0, rho_g=1.0,
                      nu_l=0.12, nu_g=0.05, sigma=0.005, width=4.0, mobility=0.08, tau_h=0.9)
    m.gravity[1] = -2e-5
    X, Y, Z = np.meshgrid(*[np.arange(n) for n in d], indexing="ij")
    r = np.sqrt((X - 15.5) ** 2 + (Y - 14.0) ** 2 + (Z - 30.0) ** 2)
    phi = (0.5 + 0.5 * np.tanh(2 * (9.0 - r) / 4.0)).reshape(-1)
    u = 0.01 * np.random.default_rng(6).standard_normal((phi.size, 3))
    print("[block order] two-phase", flush=True)

    def run():
        dev = mkg.MpSolver(m)
        dev.init(phi, u)
        dev.step(9)
        out = [dev.get(k) for k in (mkg.MP_POP_FLOW, mkg.MP_POP_PHASE, mkg.MP_PHI, mkg.MP_MU)]
        dev.close()
        return out

    a, b = both_orders(monkeypatch, run)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_multi_phase_field_step_is_independent_of_block_order(monkeypatch):
    m = mkg.default_mpf_params(nx=32, ny=32, nz=64, n=2, precision=mkg.PREC_FP32, thermal=1,
                               growth=0.2, lambda_=3.19, alpha=0.1)
    m.u[1][0] = 0.02
    phi = np.stack([np.sin(0.3 * np.arange(m.ncell) + k) * 0.9 for k in range(2)])
    print("[block order] multi-phase field", flush=True)

    def run():
        dev = mkg.MpfSolver(m)
        dev.init(phi)
        dev.step(9)
        out = [dev.get(mkg.MPF_POPS), dev.get(mkg.MPF_PHI), dev.get(mkg.MPF_THERMAL)]
        dev.close()
        return out

    a, b = both_orders(monkeypatch, run)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("thermal", [False, True], ids=["solute", "thermal"])
def test_finite_volume_step_is_independent_of_block_order(monkeypatch, thermal):
    d = (64, 64, 64)
    kw = dict(thermal=1, pf_coupling=1, alpha_l=0.1, alpha_s=0.06, latent_T=0.5) if thermal \
        else dict(solute=1, pf_coupling=2, D_l=0.1, D_s=0.01, k_part=0.15, q_min=1e-3,
                  at_coef=0.5 / math.sqrt(2.0))
    p = mkg.default_params(nx=d[0], ny=d[1], nz=d[2], phase_q=19, eps4=0.05, pf_dt=0.008,
                           pf_wdx2=6.25, phi_s=0.99, precision=mkg.PREC_FP64,
                           **{"lambda": 3.19}, **kw)
    rng = np.random.default_rng(4)
    X, Y, Z = np.meshgrid(*[np.arange(n) for n in d], indexing="ij")
    r = np.sqrt((X - 31.2) ** 2 + (Y - 32.8) ** 2 + (Z - 32.0) ** 2)
    phi = np.clip(-np.tanh((r - 14.0) / (math.sqrt(2) * 2.5)).reshape(-1)
                  + 0.02 * rng.standard_normal(r.size), -1, 1)
    second = -0.5 + 0.05 * rng.standard_normal(phi.size)
    which = mkg.FVM_TEMP if thermal else mkg.FVM_CONC
    print(f"[block order] finite volume thermal {thermal}", flush=True)

    def run():
        dev = mkg.FvmSolver(p)
        dev.set(mkg.FVM_PHI, phi)
        dev.set(which, second)
        dev.step(9)
        out = [dev.get(mkg.FVM_PHI), dev.get(which)]
        dev.close()
        return out

    a, b = both_orders(monkeypatch, run)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_adjoints_are_independent_of_block_order(monkeypatch):
    p = fused_params((32, 32, 64), mkg.PREC_FP64, walls=True)
    p.flow_q = 0
    p.has_kappa = 1
    st = random_state(p, seed=5)
    target = st["phi"] * 0.9
    print("[block order] phase-field adjoint", flush=True)

    def run_pf():
        d = load_device(p, st)
        r = d.adjoint_run(8, 3, target)
        d.close()
        return r

    a, b = both_orders(monkeypatch, run_pf)
    assert abs(a["dL_deps4"] - b["dL_deps4"]) <= 1e-12 * abs(a["dL_deps4"])
    np.testing.assert_allclose(b["dL_dkappa"], a["dL_dkappa"], rtol=0,
                               atol=1e-12 * np.abs(a["dL_dkappa"]).max())

    pf, stf, ut = flow_adjoint_case(q=19, op="mrt", dims=(32, 32, 64), seed=2)
    print("[block order] flow adjoint", flush=True)

    def run_flow():
        d = load_device(pf, stf)
        r = d.flow_adjoint_run(6, ut)
        d.close()
        return r

    a, b = both_orders(monkeypatch, run_flow)
    assert abs(a["loss"] - b["loss"]) <= 1e-12 * a["loss"]
    ga = np.concatenate([a["dL_dforce"], list(a["dL_drates"].values())])
    gb = np.concatenate([b["dL_dforce"], list(b["dL_drates"].values())])
    assert np.abs(ga - gb).max() <= 1e-11 * np.abs(ga).max(), (ga, gb)
