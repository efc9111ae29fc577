"""Shared helpers of the test-suite: random initial states, oracle/device pairs."""
from __future__ import annotations

import math

import numpy as np

from materkit_anisotropiclbm_trt_rbc_facet_amd import mkg
from oracle import oracle as O


def lat(q):
    e, w, inv, cs2 = O.lattice(q)
    return e, w, inv, cs2


def random_state(params, seed=0, with_walls=False, phi_mode="seeds"):
    """Random pre-collision populations near equilibrium for every present lattice.
    Returns a dict of full-value FP64 arrays in canonical layout."""
    rng = np.random.default_rng(seed)
    p = params
    nx, ny, nz = int(p.nx), int(p.ny), int(p.nz)
    N = nx * ny * nz
    st = {}
    flag = np.zeros(N, dtype=np.int8)
    if with_walls:
        f3 = flag.reshape(nx, ny, nz)
        f3[nx // 3: nx // 3 + 2, ny // 3: ny // 3 + 3, :nz // 2] = mkg.IS_WALLS
    st["flag"] = flag
    if p.flow_q:
        e, w, inv, cs2 = lat(p.flow_q)
        rho = 1.0 + 0.01 * rng.standard_normal(N)
        u = 0.03 * rng.standard_normal((N, 3))
        cu = u @ e.T
        uu = (u * u).sum(1)
        feq = w[None, :] * rho[:, None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * uu[:, None])
        f = feq + 1e-3 * w[None, :] * rng.standard_normal((N, p.flow_q))
        st["f"] = np.ascontiguousarray(f.T).reshape(-1)
    if p.phase_q:
        X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        r = np.sqrt((X - nx / 2.0) ** 2 + (Y - ny / 2.2) ** 2 + (Z - nz / 1.9) ** 2)
        phi = -np.tanh((r - min(nx, ny, nz) / 4.0) / 2.0).reshape(-1)
        # noise carried by the interface, as in the fields the solver produces (|grad phi|
        # grows with 1 - phi^2): noise in a bulk phase would put local extrema of phi there,
        # where the anisotropy and anti-trapping directions grad phi / |grad phi| are
        # undefined while 1 - phi^2 is not zero
        phi = np.clip(phi + 0.02 * (1 - phi ** 2) * rng.standard_normal(N), -1, 1)
        e, w, inv, cs2 = lat(p.phase_q)
        # non-equilibrium part carried by the interface: a bulk phase at rest sits at its
        # equilibrium w_i phi
        g = w[None, :] * phi[:, None] + 1e-3 * (1 - phi ** 2)[:, None] * w[None, :] * \
            rng.standard_normal((N, p.phase_q))
        st["g"] = np.ascontiguousarray(g.T).reshape(-1)
        # dphi/dt is carried by the interface ((1 - phi^2) weight), as at a moving front: the
        # bulk of each phase changes only by rounding
        st["phi_prev"] = phi + 1e-3 * (1 - phi ** 2) * rng.standard_normal(N)
        st["phi"] = phi
    if p.thermal:
        e, w, inv, cs2 = lat(7)
        T = p.T_ref + 0.1 * rng.standard_normal(N)
        h = w[None, :] * T[:, None] * (1 + 0.01 * rng.standard_normal((N, 7)))
        st["h"] = np.ascontiguousarray(h.T).reshape(-1)
    if p.solute:
        e, w, inv, cs2 = lat(7)
        Cv = -0.5 + 0.1 * rng.standard_normal(N)
        c = w[None, :] * Cv[:, None] * (1 + 0.01 * rng.standard_normal((N, 7)))
        st["c"] = np.ascontiguousarray(c.T).reshape(-1)
    if p.has_kappa:
        st["kappa"] = np.exp(0.2 * rng.standard_normal(N))
    if p.ngrains:
        st["grain"] = rng.integers(0, p.ngrains + 1, N).astype(np.int16)
    return st


def load_oracle(params, st) -> O.State:
    s = O.State.empty(params)
    for k in ("f", "g", "h", "c", "phi", "phi_prev", "kappa"):
        if k in st and getattr(s, k) is not None:
            getattr(s, k)[:] = st[k]
    s.flag[:] = st["flag"]
    if "grain" in st:
        s.grain[:] = st["grain"]
    return s


def load_device(params, st) -> mkg.Solver:
    d = mkg.Solver(params)
    d.set(mkg.F_FLAG, st["flag"])
    if "grain" in st:
        d.set(mkg.F_GRAIN, st["grain"])
    for k, fid in (("f", mkg.F_POP_FLOW), ("g", mkg.F_POP_PHASE), ("h", mkg.F_POP_THERMAL),
                   ("c", mkg.F_POP_SOLUTE), ("kappa", mkg.F_KAPPA)):
        if k in st:
            d.set(fid, st[k])
    if "phi" in st:
        d.set(mkg.F_PHI, st["phi"])
        d.set(mkg.F_PHI_PREV, st["phi_prev"])
    return d


def _pop_deviation(p, key, pops):
    """Populations minus the rest part w_i * ref that FP32 storage subtracts."""
    q = {"f": int(p.flow_q), "g": int(p.phase_q), "h": 7, "c": 7}[key]
    ref = {"f": p.ref_f, "g": p.ref_g, "h": p.ref_h, "c": p.ref_c}[key]
    e, w, inv, cs2 = lat(q)
    return pops.reshape(q, -1) - (w * ref)[:, None]


def compare_pops(dev: mkg.Solver, orc: O.State, cells=None):
    """Max error per present lattice, relative to the largest deviation |p_i - w_i ref| of the
    oracle's populations (the part FP32 storage keeps; the rest part w_i ref is exact), and
    the max error of phi relative to its range.  cells (optional, N bools): the cells
    compared."""
    out = {}
    for k, fid in (("f", mkg.F_POP_FLOW), ("g", mkg.F_POP_PHASE), ("h", mkg.F_POP_THERMAL),
                   ("c", mkg.F_POP_SOLUTE)):
        ref = getattr(orc, k)
        if ref is None:
            continue
        q = ref.size // orc.N
        got = dev.get(fid).reshape(q, -1)
        ref = ref.reshape(q, -1)
        scale = max(np.abs(_pop_deviation(orc.params, k, ref)).max(), 1e-300)
        d = np.abs(got - ref)
        if cells is not None:
            d = d[:, cells]
        out[k] = float(d.max() / scale)
    if orc.phi is not None:
        got = dev.get(mkg.F_PHI)
        rng = max(float(orc.phi.max() - orc.phi.min()), 1e-300)
        d = np.abs(got - orc.phi)
        if cells is not None:
            d = d[cells]
        out["phi"] = float(d.max() / rng)
    return out


def step_pair(dev: mkg.Solver, orc: O.State, nsteps: int, op: int, mask: int = 15):
    """Advance device and oracle nsteps steps, one at a time, and return the cells outside the
    light cone of every flag that differs between them.  In FP32 a phase value within rounding
    of the classification threshold can flip a flag (interface / solid), which changes the
    solid fraction the flow sees; its effect spreads by one cell per step.  Returns
    (trusted cells, number of flips)."""
    p = orc.params
    dims = (int(p.nx), int(p.ny), int(p.nz))
    untrusted = np.zeros(dims, dtype=bool)
    flips = 0
    for k in range(1, nsteps + 1):
        dev.step(1, mask)
        orc.step(1, mask, op)
        diff = (dev.get(mkg.F_FLAG) != orc.flag).reshape(dims)
        flips += int(diff.sum())
        if diff.any() and p.flow_q:  # flags act on the flow only (solid fraction)
            r = nsteps - k + 1
            for x, y, z in zip(*np.nonzero(diff)):
                ix = np.arange(x - r, x + r + 1) % dims[0]
                iy = np.arange(y - r, y + r + 1) % dims[1]
                iz = np.arange(z - r, z + r + 1) % dims[2]
                untrusted[np.ix_(ix, iy, iz)] = True
    return ~untrusted.reshape(-1), flips


def field_errors(dev: mkg.Solver, orc: O.State, mask=None):
    """Max and RMS error of every present macroscopic field (rho, u, T, c, phi) of the device
    against the oracle, each relative to the oracle field's range (max - min; for u the
    largest speed; for rho the reference density rho0, since the density of a nearly
    incompressible flow varies by far less than its rounding).  mask (optional, N bools)
    restricts the cells compared."""
    p = orc.params
    m = orc.moments()
    out = {}
    pairs = []
    if p.flow_q:
        pairs.append(("rho", dev.get(mkg.F_RHO), m["rho"]))
        pairs.append(("u", dev.get(mkg.F_VEL).reshape(-1, 3), m["vel"].reshape(-1, 3)))
    if p.thermal:
        pairs.append(("T", dev.get(mkg.F_TEMP), m["T"]))
    if p.solute:
        pairs.append(("c", dev.get(mkg.F_CONC), m["C"]))
    if p.phase_q:
        pairs.append(("phi", dev.get(mkg.F_PHI), orc.phi))
    for name, got, ref in pairs:
        if mask is not None:
            got, ref = got[mask], ref[mask]
        d = got - ref
        if name == "u":
            scale = float(np.sqrt((ref ** 2).sum(1)).max())
            d = np.sqrt((d ** 2).sum(1))
        elif name == "rho":
            scale = float(p.rho0)
        else:
            scale = float(ref.max() - ref.min())
        scale = max(scale, 1e-300)
        out[name] = {"max": float(np.abs(d).max() / scale),
                     "rms": float(np.sqrt((d ** 2).mean()) / scale), "scale": scale}
    return out


def _contour(ph, ymin=2):
    """phi = 1/2 crossings along columns and rows (rows from ymin up, off the wall layer)."""
    nx, ny = ph.shape
    pts = []
    for x in range(nx):
        c = ph[x]
        for j in range(ymin, ny - 1):
            if (c[j] - 0.5) * (c[j + 1] - 0.5) < 0:
                pts.append((x, j + (c[j] - 0.5) / (c[j] - c[j + 1])))
    for y in range(ymin, ny):
        r = ph[:, y]
        for i in range(nx - 1):
            if (r[i] - 0.5) * (r[i + 1] - 0.5) < 0:
                pts.append((i + (r[i] - 0.5) / (r[i] - r[i + 1]), y))
    return np.array(pts)


def contact_angle(ph):
    """Contact angle (degrees, in the phi = 1 fluid) of a 2-D drop on the wall at y = -1/2:
    algebraic circle fit of the contour, cos(theta) = -(height of the centre above the wall)/R."""
    p = _contour(ph)
    x, y = p[:, 0], p[:, 1]
    (a, c, d), *_ = np.linalg.lstsq(np.c_[2 * x, 2 * y, np.ones_like(x)], x * x + y * y,
                                    rcond=None)
    R = math.sqrt(d + a * a + c * c)
    return math.degrees(math.acos(-(c + 0.5) / R)), a


def wetting_drop(theta, prec=None, nx=100, ny=40, R=12.0, W=4.0):
    """A half-disc drop of the dense fluid on the lower y wall (wetting, contact angle theta),
    neutral upper wall, periodic x."""
    m = mkg.mp_params(nx=nx, ny=ny, nz=1, rho_l=1.0, rho_g=1.0, nu_l=0.1, nu_g=0.1,
                      sigma=0.05, width=W, mobility=0.1, tau_h=1.0)
    if prec is not None:
        m.precision = prec
    m.bc[2] = m.bc[3] = mkg.BC_BOUNCE
    m.wet_faces = 1 << 2
    m.contact_angle = theta
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    r = np.sqrt((X - (nx - 1) / 2) ** 2 + (Y + 0.5) ** 2)
    return m, (0.5 - 0.5 * np.tanh(2 * (r - R) / W)).reshape(-1)


def flow_adjoint_case(q=19, op="mrt", prec=mkg.PREC_FP64, dims=(7, 6, 5), seed=0, walls=True):
    """A flow-only case for the adjoint: bounce x- face, outflow x+, periodic y, bounce z faces
    (the z- wall moving at 0.01 along x), random near-equilibrium populations, wall cells, and
    a random velocity target."""
    p = mkg.default_params(nx=dims[0], ny=dims[1], nz=dims[2], flow_q=q, precision=prec)
    om = 1.0 / 0.8
    p.s_nu = om
    if op == "bgk":
        p.s_bulk = p.s_odd = p.s_even = om
    elif op == "trt":
        p.s_bulk = p.s_even = om
        p.s_odd = 1.0 / ((3.0 / 16.0) / (1.0 / om - 0.5) + 0.5)
    else:
        p.s_bulk, p.s_odd, p.s_even = 1.19, 1.2, 1.4
    p.force[0], p.force[1], p.force[2] = 1e-5, -2e-6, 3e-6
    for k, v in enumerate([mkg.BC_BOUNCE, mkg.BC_OUTFLOW, mkg.BC_PERIODIC, mkg.BC_PERIODIC,
                           mkg.BC_BOUNCE, mkg.BC_BOUNCE]):
        p.bc_f[k] = v
    p.u_face[4][0] = 0.01
    st = random_state(p, seed=seed, with_walls=walls)
    N = int(p.nx) * int(p.ny) * int(p.nz)
    ut = 0.01 * np.random.default_rng(seed + 100).standard_normal(3 * N)
    return p, st, ut


def oracle_flow_loss(p, st, nsteps, ut):
    """1/2 sum over non-wall cells |u_N - u_target|^2 after nsteps oracle flow steps (the
    generalised moment collision)."""
    s = O.State.empty(p)
    s.f[:] = st["f"]
    s.flag[:] = st["flag"]
    s.step(nsteps, 1, O.OP_MRT)
    N = s.N
    f = s.f.reshape(int(p.flow_q), N)
    e, w, inv, cs2 = O.lattice(int(p.flow_q))
    rho = f.sum(0)
    u = ((e.T.astype(float) @ f).T + 0.5 * np.array(list(p.force))) / rho[:, None]
    m = st["flag"] != mkg.IS_WALLS
    return 0.5 * ((u - ut.reshape(N, 3))[m] ** 2).sum()


FLOW_ADJ_PARAMS = ("Fx", "Fy", "Fz", "s_nu", "s_bulk", "s_odd", "s_even")


def perturbed(p, k, h):
    """A copy of p with the k-th flow-adjoint parameter moved by h."""
    pp = p.copy()
    if k < 3:
        pp.force[k] = p.force[k] + h
    else:
        setattr(pp, FLOW_ADJ_PARAMS[k], getattr(p, FLOW_ADJ_PARAMS[k]) + h)
    return pp


# ---- Karma-Rappel grains of the multi-phase field

def quat_rot(q) -> np.ndarray:
    """Row-major rotation matrix of the unit quaternion q = (w, x, y, z)."""
    w, x, y, z = np.asarray(q, dtype=float) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


GRAIN_FORMS = {  # aniso_kind, eps4, the two parameters of the other forms
    "cubic": (mkg.ANISO_CUBIC, 0.05, (0.0, 0.0)),
    "cbc": (mkg.ANISO_CBC, 0.0, (0.04, -0.01)),
    "hcp": (mkg.ANISO_HCP, 0.0, (0.03, -0.02)),
    "facet": (mkg.ANISO_FACET, 0.0, (0.05, 0.2)),
}


def grain_params(dims, n, forms, rots, coupling=1, prec=mkg.PREC_FP64, kappa=1.0, **kw):
    """Multi-phase parameters of n Karma-Rappel grains (forms: names of GRAIN_FORMS, rots: 3x3
    rotations), periodic in x and y, bounce-back phase faces and Dirichlet T / U faces in z;
    coupling 1: thermal lattice (latent heat), 2: solute lattice (anti-trapping)."""
    m = mkg.default_mpf_params(nx=dims[0], ny=dims[1], nz=dims[2], n=n, precision=prec,
                               model=mkg.MPF_KARMA_RAPPEL, coupling=coupling, dt=0.008, wdx=2.5,
                               lambda_=3.19, alpha=0.1, latent_T=0.5, T_ref=-0.55)
    m.set(bc=[0, 0, 0, 0, 1, 1])
    for j in range(n):
        m.u[j][0] = 0.0
        m.kappa[j] = kappa
        kind, e4, ap = GRAIN_FORMS[forms[j]]
        m.aniso_kind[j] = kind
        m.eps4[j] = e4
        m.aniso[j][0], m.aniso[j][1] = ap
        R = np.asarray(rots[j], dtype=float).reshape(-1)
        for k in range(9):
            m.rot[j][k] = R[k]
    if coupling == 1:
        m.thermal = 1
        m.set(bc_h=[0, 0, 0, 0, 2, 2], T_face=[0, 0, 0, 0, -0.5, -0.6])
    if coupling == 2:
        m.solute = 1
        m.set(bc_c=[0, 0, 0, 0, 2, 2], C_face=[0, 0, 0, 0, -0.3, -0.35])
        m.D_l, m.D_s, m.k_part, m.q_min = 1.0, 1e-4, 0.15, 1e-3
        m.latent_C, m.at_coef = 0.4, 1.0 / (2.0 * math.sqrt(2.0))
    m.set(**kw)
    return m


def grain_fields(m, centers, radii, seed=4):
    """Seeds of the grains, a noisy T (coupling 1) or c = A(phi) U (coupling 2) around the
    face values."""
    from materkit_anisotropiclbm_trt_rbc_facet_amd import cases
    dims = (int(m.nx), int(m.ny), int(m.nz))
    N = int(np.prod(dims))
    phi = cases.mpf_seed_fields(dims, centers, radii, float(m.wdx))
    rng = np.random.default_rng(seed)
    T = C = None
    if m.thermal:
        T = -0.55 + 0.01 * rng.standard_normal(N)
    if m.solute:
        comb = np.clip((m.n - 1) + phi.sum(0), -1.0, 1.0)
        A = 0.5 * ((1 + m.k_part) - (1 - m.k_part) * comb)
        C = A * (-0.3 + 0.01 * rng.standard_normal(N))
    return phi, T, C


def rot_z(deg):
    t = np.radians(deg)
    return np.array([[np.cos(t), -np.sin(t), 0.0], [np.sin(t), np.cos(t), 0.0], [0.0, 0.0, 1.0]])


def six_fold_orientation(ph, center):
    """Orientation (degrees, in [0, 60)) of the six-fold growth of a 2-D grain: the phase of
    sum over its solid cells of exp(6 i theta) about the seed, divided by 6."""
    nx, ny = ph.shape
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    th = np.arctan2(Y - center[1], X - center[0])
    M = (np.exp(6j * th) * (ph > 0)).sum()
    return float(np.degrees(np.angle(M)) / 6.0) % 60.0


def angle_gap(a, b):
    d = abs(a - b) % 60.0
    return min(d, 60.0 - d)
