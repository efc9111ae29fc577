"""The inverse problem of configuration C5 at its full size: 128^3, 2000 steps, FP32 forward
and adjoint with checkpoints every 100 steps and recompute, FP64 loss reduction.

The target phase field comes from a forward run with eps4 = 0.03 and the log-normal
conductivity field (sigma 0.2, correlation length 8); the gradient of
L = 1/2 sum (phi_N - phi*)^2 is taken at eps4 = 0.05 and uniform conductivity.  dL/deps4 and
the derivative along a random direction of the per-cell conductivity must agree with central
differences (h = 1e-4) of FP64 forward runs on the device to 1e-4 relative.
"""
import numpy as np
import pytest

from materkit_anisotropiclbm_trt_rbc_facet_amd import cases, mkg

pytestmark = pytest.mark.gpu

STEPS = 2000
CKPT = 100
H = 1e-4


def _forward_phi(prec, eps4, kappa, steps=STEPS):
    case = cases.load("C5", precision=prec, eps4=eps4)
    d = mkg.Solver(case.params)
    cases.setup(d, case)
    d.set(mkg.F_KAPPA, kappa)
    d.step(steps, mkg.ALL)
    phi = d.phi_now()
    d.close()
    return phi


def _loss(phi, target):
    return 0.5 * float(((phi - target) ** 2).sum())


def test_c5_fp32_adjoint_matches_fp64_central_differences():
    case = cases.load("C5")
    p = case.params
    dims = (p.nx, p.ny, p.nz)
    assert dims == (128, 128, 128)
    k_true = cases.lognormal_field(dims, *case.conductivity).reshape(-1)
    ones = np.ones(p.ncell)
    print("[adjoint C5] target run", flush=True)
    target = _forward_phi(mkg.PREC_FP64, 0.03, k_true)
    print("[adjoint C5] FP32 adjoint, checkpoints every 100 steps", flush=True)
    case32 = cases.load("C5", precision=mkg.PREC_FP32, eps4=0.05)
    d = mkg.Solver(case32.params)
    cases.setup(d, case32)
    d.set(mkg.F_KAPPA, ones)
    out = d.adjoint_run(STEPS, CKPT, target)
    d.close()
    print(f"[adjoint C5] loss {out['loss']!r} dL/deps4 {out['dL_deps4']!r} times "
          f"{out['forward_ms']:.1f} / {out['recompute_ms']:.1f} / {out['reverse_ms']:.1f} ms",
          flush=True)
    fd_e = (_loss(_forward_phi(mkg.PREC_FP64, 0.05 + H, ones), target) -
            _loss(_forward_phi(mkg.PREC_FP64, 0.05 - H, ones), target)) / (2 * H)
    v = np.random.default_rng(7).standard_normal(p.ncell)
    fd_k = (_loss(_forward_phi(mkg.PREC_FP64, 0.05, ones + H * v), target) -
            _loss(_forward_phi(mkg.PREC_FP64, 0.05, ones - H * v), target)) / (2 * H)
    ad_k = float(out["dL_dkappa"] @ v)
    err_e = abs(out["dL_deps4"] - fd_e) / abs(fd_e)
    err_k = abs(ad_k - fd_k) / abs(fd_k)
    print(f"[adjoint C5] dL/deps4 adjoint {out['dL_deps4']!r} fd {fd_e!r} rel {err_e:.2e}; "
          f"dL/dkappa.v adjoint {ad_k!r} fd {fd_k!r} rel {err_k:.2e}", flush=True)
    assert err_e < 1e-4
    assert err_k < 1e-4
