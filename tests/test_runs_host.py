"""Host-side pieces of the configuration runner (no GPU): tip tracking and the log-normal
conductivity field."""
import math

import numpy as np

from materkit_anisotropiclbm_trt_rbc_facet_amd import cases, runs


def test_tip_positions_interpolate_the_zero_crossing():
    dims = (40, 36, 30)
    c = (19.5, 17.0, 14.25)
    X, Y, Z = np.meshgrid(*[np.arange(n) for n in dims], indexing="ij")
    # ellipsoid with semi-axes 9.3, 7.1, 5.6 -> crossings at those distances
    r = np.sqrt(((X - c[0]) / 9.3) ** 2 + ((Y - c[1]) / 7.1) ** 2 + ((Z - c[2]) / 5.6) ** 2)
    phi = -np.tanh((r - 1.0) * 4.0)
    tips = runs.tip_positions(phi.reshape(-1), dims, c)
    # linear interpolation of tanh between cells: within a few hundredths of a cell
    np.testing.assert_allclose(tips, [9.3, 7.1, 5.6], atol=0.05)


def test_lognormal_conductivity_statistics():
    k = cases.lognormal_field((32, 32, 32), 0.2, 4.0, 7)
    g = np.log(k)
    assert abs(g.mean()) < 1e-12 and abs(g.std() - 0.2) < 1e-12
    # correlated: neighbouring cells are close, far cells independent
    c1 = np.corrcoef(g[:-1].ravel(), g[1:].ravel())[0, 1]
    c16 = np.corrcoef(g[:-16].ravel(), g[16:].ravel())[0, 1]
    assert c1 > 0.9 and abs(c16) < 0.3
    # deterministic in the seed
    np.testing.assert_array_equal(k, cases.lognormal_field((32, 32, 32), 0.2, 4.0, 7))
