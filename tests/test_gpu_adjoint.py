"""Discrete adjoint of the phase-field + thermal step against central finite differences of
the FP64 oracle (the FP64 reference gradient) and of the device forward.

dL/deps4 and directional / per-cell dL/dkappa, L = 1/2 sum (phi_N - target)^2, with the target
taken from a different forward run (eps4 = 0.03, log-normal kappa).  FP64 adjoint within 1e-6
of the FP64 central differences (h = 1e-4); FP32 adjoint within 1e-3.  Checkpoint spacing does
not change the gradient.
"""
import numpy as np
import pytest
import torch

from materkit_anisotropiclbm_trt_rbc_facet_amd import autograd, mkg
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DIMS = (16, 14, 12)
NSTEPS = 30
H = 1e-4


def params(prec=mkg.PREC_FP64, eps4=0.05, bcs="periodic", rotate=False):
    p = mkg.default_params(nx=DIMS[0], ny=DIMS[1], nz=DIMS[2], phase_q=19, thermal=1,
                           precision=prec, has_kappa=1, eps4=eps4, pf_dt=0.008, pf_wdx2=6.25,
                           alpha_l=0.1, alpha_s=0.06, latent_T=0.5, pf_coupling=1, T_ref=-0.55,
                           ref_h=-0.55, **{"lambda": 3.19})
    if bcs == "walls":
        p.set(bc_g=[1, 1, 0, 0, 1, 1], bc_h=[2, 2, 0, 0, 1, 1], T_face=[-0.5, -0.6, 0, 0, 0, 0])
    if rotate:
        p.ngrains = 1
        R = mkg.quat_to_rot([0.9, 0.2, -0.3, 0.25])
        p.set(rot=R.reshape(-1))
    return p


SEED_C = np.array([[7.3, 6.6, 5.8]])
SEED_R = np.array([4.0])


def kappa_field(seed):
    rng = np.random.default_rng(seed)
    return np.exp(0.2 * rng.standard_normal(np.prod(DIMS)))


def oracle_phi(p, kappa, nsteps=NSTEPS):
    s = O.State.empty(p)
    s.init_seeds(SEED_C, SEED_R)
    if p.ngrains:
        s.grain[:] = np.where(s.grain > 0, 1, 0)
    s.init_equilibrium(T=np.full(s.N, -0.55))
    s.kappa[:] = kappa
    s.step(nsteps, 6, O.OP_BGK)
    return s.g.reshape(19, s.N).sum(0)


def loss(phi, target):
    return 0.5 * float(((phi - target) ** 2).sum())


def make_device(p, kappa):
    d = mkg.Solver(p)
    d.init_seeds(SEED_C, SEED_R)
    d.init_equilibrium(T=np.full(d.N, -0.55))
    d.set(mkg.F_KAPPA, kappa)
    return d


@pytest.fixture(scope="module")
def target():
    return oracle_phi(params(eps4=0.03), kappa_field(11))


@pytest.mark.parametrize("bcs,rotate", [("periodic", False), ("walls", False),
                                        ("periodic", True)])
def test_fp64_adjoint_matches_oracle_finite_differences(target, bcs, rotate):
    p = params(bcs=bcs, rotate=rotate)
    kap = kappa_field(3)
    print(f"[adjoint] fp64 bcs={bcs} rotate={rotate}", flush=True)
    d = make_device(p, kap)
    out = d.adjoint_run(NSTEPS, 7, target)
    # the forward state after the call equals a plain forward run
    phi_dev = d.phi_now()
    d.close()
    phi_orc = oracle_phi(p, kap)
    assert np.abs(phi_dev - phi_orc).max() < 1e-12
    assert abs(out["loss"] - loss(phi_orc, target)) < 1e-10 * max(1.0, out["loss"])
    # dL/deps4
    pp, pm = params(eps4=0.05 + H, bcs=bcs, rotate=rotate), params(eps4=0.05 - H, bcs=bcs,
                                                                     rotate=rotate)
    fd = (loss(oracle_phi(pp, kap), target) - loss(oracle_phi(pm, kap), target)) / (2 * H)
    assert abs(out["dL_deps4"] - fd) <= 1e-6 * abs(fd) + 1e-12, (out["dL_deps4"], fd)
    # directional derivative along a random kappa perturbation
    rng = np.random.default_rng(5)
    v = rng.standard_normal(kap.size)
    fdv = (loss(oracle_phi(p, kap + H * v), target) - loss(oracle_phi(p, kap - H * v), target)) / (
        2 * H)
    adv = float(out["dL_dkappa"] @ v)
    assert abs(adv - fdv) <= 1e-6 * abs(fdv) + 1e-12, (adv, fdv)
    # per-cell derivatives at the three most sensitive cells
    for c in np.argsort(-np.abs(out["dL_dkappa"]))[:3]:
        e = np.zeros(kap.size)
        e[c] = 1.0
        fdc = (loss(oracle_phi(p, kap + H * e), target) -
               loss(oracle_phi(p, kap - H * e), target)) / (2 * H)
        assert abs(out["dL_dkappa"][c] - fdc) <= 1e-5 * abs(fdc) + 1e-12, (c, out["dL_dkappa"][c], fdc)


def test_checkpoint_spacing_does_not_change_the_gradient(target):
    """Checkpoints every 1, 4 or 30 steps, or the tape (ckpt 0, no recompute), give the same
    gradient; the tape leaves the handle at the final forward state like the others."""
    p = params()
    kap = kappa_field(3)
    outs = []
    finals = []
    for ck in (1, 4, 30, 0):
        d = make_device(p, kap)
        outs.append(d.adjoint_run(NSTEPS, ck, target))
        finals.append(d.get(mkg.F_POP_PHASE))
        d.close()
    for fin in finals[1:]:
        np.testing.assert_array_equal(fin, finals[0])
    for o in outs[1:]:
        assert abs(o["dL_deps4"] - outs[0]["dL_deps4"]) <= 1e-12 * abs(outs[0]["dL_deps4"])
        assert np.abs(o["dL_dkappa"] - outs[0]["dL_dkappa"]).max() <= 1e-12 * np.abs(
            outs[0]["dL_dkappa"]).max()


def test_fp32_adjoint_close_to_fp64(target):
    kap = kappa_field(3)
    o64 = make_device(params(), kap)
    r64 = o64.adjoint_run(NSTEPS, 10, target)
    o64.close()
    o32 = make_device(params(prec=mkg.PREC_FP32), kap)
    r32 = o32.adjoint_run(NSTEPS, 10, target)
    o32.close()
    assert abs(r32["dL_deps4"] - r64["dL_deps4"]) <= 1e-3 * abs(r64["dL_deps4"])
    num = np.linalg.norm(r32["dL_dkappa"] - r64["dL_dkappa"])
    assert num <= 1e-3 * np.linalg.norm(r64["dL_dkappa"])


def test_torch_autograd_wrapper(target):
    kap = torch.tensor(kappa_field(3), dtype=torch.float64, requires_grad=True)
    eps = torch.tensor(0.05, dtype=torch.float64, requires_grad=True)

    def make(e):
        return make_device(params(eps4=e), np.ones(np.prod(DIMS)))

    L = autograd.PhaseFieldMismatch.apply(eps, kap, make, NSTEPS, 10, torch.from_numpy(target))
    L.backward()
    d = make_device(params(), kap.detach().numpy())
    ref = d.adjoint_run(NSTEPS, 10, target)
    d.close()
    assert abs(L.item() - ref["loss"]) < 1e-12 * max(1, ref["loss"])
    assert abs(eps.grad.item() - ref["dL_deps4"]) < 1e-12 * abs(ref["dL_deps4"])
    assert torch.allclose(kap.grad, torch.from_numpy(ref["dL_dkappa"]), rtol=1e-12, atol=0)
