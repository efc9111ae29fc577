"""Discrete flow adjoint on the device (mkg_flow_adjoint_run) against the oracle's
(orc_flow_adjoint, pinned by finite differences in test_oracle_flow_adjoint.py): loss and the
gradient with respect to the body force and the four collision rates, for D3Q19 and D3Q27,
BGK, TRT and MRT rates, bounce (moving wall), outflow and periodic faces and wall cells; FP64
to 1e-10, FP32 to 1e-3 of the gradient's magnitude.  A finite difference of the device's own
FP64 forward checks it independently of the oracle, and an nz = 64 grid runs the forward's
z-staged O steps."""
import numpy as np
import pytest

from helpers import flow_adjoint_case, load_device, perturbed
from materkit_anisotropiclbm_trt_rbc_facet_amd import mkg
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def device_grad(p, st, nsteps, ut):
    dev = load_device(p, st)
    r = dev.flow_adjoint_run(nsteps, ut)
    f_after = dev.get(mkg.F_POP_FLOW)
    dev.close()
    g = np.concatenate([r["dL_dforce"], [r["dL_drates"][k] for k in ("s_nu", "s_bulk", "s_odd", "s_even")]])
    return r["loss"], g, f_after


def device_loss(p, st, nsteps, ut):
    dev = load_device(p, st)
    dev.step(nsteps, mkg.FLOW)
    u = dev.get(mkg.F_VEL).reshape(-1, 3)
    dev.close()
    m = st["flag"] != mkg.IS_WALLS
    return 0.5 * ((u - ut.reshape(-1, 3))[m] ** 2).sum()


@pytest.mark.parametrize("prec,tol", [(mkg.PREC_FP64, 1e-10), (mkg.PREC_FP32, 1e-3)],
                         ids=["fp64", "fp32"])
@pytest.mark.parametrize("op", ["bgk", "trt", "mrt"])
@pytest.mark.parametrize("q", [19, 27])
def test_device_flow_adjoint_matches_oracle(q, op, prec, tol):
    nsteps = 8
    p, st, ut = flow_adjoint_case(q=q, op=op, prec=prec, dims=(9, 7, 6), seed=2)
    p64, _, _ = flow_adjoint_case(q=q, op=op, prec=mkg.PREC_FP64, dims=(9, 7, 6), seed=2)
    print(f"flow adjoint q {q} op {op} prec {prec}", flush=True)
    lo, go = O.flow_adjoint(p64, st["flag"], st["f"], nsteps, ut)
    ld, gd, f_after = device_grad(p, st, nsteps, ut)
    assert abs(ld - lo) <= tol * lo
    assert np.abs(gd - go).max() <= tol * np.abs(go).max(), (gd, go)
    # the handle is left at step nsteps
    s = O.State.empty(p64)
    s.f[:] = st["f"]
    s.flag[:] = st["flag"]
    s.step(nsteps, 1, O.OP_MRT)
    m = np.repeat(st["flag"] != mkg.IS_WALLS, 1)
    fa = f_after.reshape(q, -1)[:, m]
    fo = s.f.reshape(q, -1)[:, m]
    assert np.abs(fa - fo).max() <= max(tol, 1e-12) * 10 * np.abs(fo).max()


def test_device_flow_adjoint_matches_device_finite_differences():
    """Independent of the oracle: central differences of the device's FP64 forward loss."""
    nsteps = 20
    p, st, ut = flow_adjoint_case(q=19, op="mrt", dims=(16, 12, 10), seed=4)
    ld, gd, _ = device_grad(p, st, nsteps, ut)
    assert abs(ld - device_loss(p, st, nsteps, ut)) <= 1e-12 * ld
    h = 1e-6
    for k in (0, 2, 3, 6):
        fd = (device_loss(perturbed(p, k, h), st, nsteps, ut) -
              device_loss(perturbed(p, k, -h), st, nsteps, ut)) / (2 * h)
        # central differences at h = 1e-6 over 20 steps resolve the gradient to ~1e-6
        assert abs(gd[k] - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(gd).max()), (k, gd[k], fd)


def test_device_flow_adjoint_with_z_staged_o_steps():
    """nz = 64: the forward's O steps move the z populations through LDS."""
    nsteps = 7
    p, st, ut = flow_adjoint_case(q=19, op="trt", dims=(4, 3, 64), seed=5)
    lo, go = O.flow_adjoint(p, st["flag"], st["f"], nsteps, ut)
    ld, gd, _ = device_grad(p, st, nsteps, ut)
    assert abs(ld - lo) <= 1e-10 * lo
    assert np.abs(gd - go).max() <= 1e-10 * np.abs(go).max()


def test_device_flow_adjoint_rejects_unsupported_cases():
    p, st, ut = flow_adjoint_case()
    p.bc_f[1] = mkg.BC_PRESSURE
    dev = load_device(p, st)
    with pytest.raises(mkg.MkgError, match="periodic, bounce or outflow"):
        dev.flow_adjoint_run(2, ut)
    dev.close()
    p, st, ut = flow_adjoint_case()
    dev = load_device(p, st)
    with pytest.raises(mkg.MkgError, match="3 N values"):
        dev.flow_adjoint_run(2, ut[:-1])
    dev.close()
