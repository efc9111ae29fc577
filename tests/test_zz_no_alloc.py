"""mkg_reduce_noalloc (the reductions with no device allocation) against mkg_reduce through the
C ABI: every reduction a case supports gives the same values bit for bit, on two calls in a
row whose second grid is smaller (the shared device-resident array keeps the first call's
partials), and grids above the array's 2^25-cell maximum are refused."""
import numpy as np
import pytest

from helpers import load_device, random_state
from materkit_anisotropiclbm_trt_rbc_facet_amd import mkg
from test_gpu_parity import make_params

pytestmark = pytest.mark.gpu


def _applicable(p):
    out = [mkg.R_NONFINITE]
    if p.flow_q:
        out += [mkg.R_MASS, mkg.R_MOMENTUM, mkg.R_UMAX]
    if p.thermal:
        out += [mkg.R_ENERGY]
    if p.flow_q and p.thermal:
        out += [mkg.R_NUSSELT]
    if p.solute:
        out += [mkg.R_SOLUTE]
    if p.phase_q:
        out += [mkg.R_SOLID, mkg.R_PHI, mkg.R_TIP]
    return out


def _compare(p, seed, walls, nan_cell=None):
    st = random_state(p, seed=seed, with_walls=walls)
    if nan_cell is not None:
        st["f"][nan_cell] = np.nan
    d = load_device(p, st)
    if nan_cell is None:
        d.step(5, mkg.ALL)
    checked = []
    for which in _applicable(p):
        a = d.reduce(which)
        b = d.reduce_noalloc(which)
        print(f"[no-alloc] {p.nx}x{p.ny}x{p.nz} which={which} reduce={a} noalloc={b}",
              flush=True)
        np.testing.assert_array_equal(a, b)
        checked.append(which)
    d.close()
    return checked


CASES = [
    dict(flow_q=19, phase_q=19, thermal=1, op="trt"),
    dict(flow_q=19, phase_q=19, solute=1, op="mrt"),
    dict(flow_q=27, thermal=1, op="bgk"),
]


@pytest.mark.parametrize("prec", [mkg.PREC_FP32, mkg.PREC_FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_reductions_equal_reduce_on_two_calls_in_a_row(case, prec):
    kw = dict(CASES[case])
    op = kw.pop("op")
    big = make_params(op=op, bcs="mixed", prec=prec, dims=(20, 18, 64), **kw)
    small = make_params(op=op, bcs="mixed", prec=prec, dims=(6, 5, 8), **kw)
    n_big = _compare(big, seed=21, walls=bool(kw.get("flow_q")))
    n_small = _compare(small, seed=22, walls=False)
    assert n_big == n_small and len(n_big) >= 5


def test_noalloc_counts_non_finite_populations():
    p = make_params(flow_q=19, thermal=1, op="bgk", prec=mkg.PREC_FP32, dims=(8, 6, 4))
    _compare(p, seed=3, walls=False, nan_cell=17)


def test_noalloc_refuses_grids_above_its_maximum():
    # 8193 x 64 x 64 = 33,558,528 cells > 2^25 (thermal lattice only, about 1 GB in FP32)
    p = mkg.default_params(nx=8193, ny=64, nz=64, thermal=1, precision=mkg.PREC_FP32)
    d = mkg.Solver(p)
    with pytest.raises(mkg.MkgError, match="maximum"):
        d.reduce_noalloc(mkg.R_ENERGY)
    d.close()
