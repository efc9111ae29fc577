"""Full configuration runs with the observables each case reports.

    python -m materkit_anisotropiclbm_trt_rbc_facet_amd.runs C3 [--steps N] [--frames F]

C1  duct start-up: centre-line velocity against the transient rectangular-duct series
C2  Rayleigh-Benard: Nusselt number per frame
C3  free dendrite: tip position along +x, +y, +z from the seed centre (the furthest phi = 0
    crossing along each axis, interpolated between cells, reduced on the GPU), tip velocity
    (physical units W0 / tau0) and solid fraction per frame
C4  multi-dendrite growth: solid fraction, solute range and maximum velocity per frame
C5  forward target run of the inverse problem

Each frame prints one JSON line; the last line is a summary with the wall time and MLUPS.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

from . import cases, mkg


def tip_positions(phi: np.ndarray, dims, center) -> np.ndarray:
    """phi = 0 crossing along +x, +y, +z on the axes through the seed centre (cells, from the
    centre), from a host copy of phi: the check of the on-GPU tip reduction."""
    nx, ny, nz = dims
    f = phi.reshape(nx, ny, nz)
    cx, cy, cz = (int(round(c)) for c in center)
    lines = (f[cx:, cy, cz], f[cx, cy:, cz], f[cx, cy, cz:])
    base = (cx, cy, cz)
    out = []
    for a, line in enumerate(lines):
        idx = np.where((line[:-1] >= 0) & (line[1:] < 0))[0]
        if idx.size == 0:
            out.append(float("nan"))
            continue
        i = idx[-1]
        frac = line[i] / (line[i] - line[i + 1])
        out.append(base[a] + i + frac - center[a])
    return np.array(out)


def run(name: str, steps: int | None = None, frames: int | None = None, precision=None,
        emit=print) -> dict:
    over = {} if precision is None else {"precision": precision}
    case = cases.load(name, **over)
    p = case.params
    steps = steps if steps is not None else case.info["numLoops"]
    frames = max(1, frames if frames is not None else min(case.info["numFrames"], 20))
    every = max(1, steps // frames)
    d = mkg.Solver(p)
    cases.setup(d, case)
    if case.conductivity is not None:
        d.set(mkg.F_KAPPA, cases.lognormal_field((p.nx, p.ny, p.nz), *case.conductivity))
    dims = (p.nx, p.ny, p.nz)
    dx_phys = 1.0 / math.sqrt(p.pf_wdx2) if p.pf_wdx2 > 0 else 1.0  # cell size in W0
    dt_phys = p.pf_dt if p.pf_dt > 0 else 1.0
    hist = []
    done = 0
    t_dev = 0.0
    last_tip = None
    while True:
        rec = {"case": name, "step": done}
        if p.phase_q:
            rec["solid_fraction"] = float(d.reduce(mkg.R_SOLID)[0])
            if case.seeds is not None and len(case.seeds[1]) == 1:
                # on-GPU tip reduction (from the domain centre) moved to the seed centre
                centre = 0.5 * (np.array(dims, dtype=float) - 1.0)
                tip = d.reduce(mkg.R_TIP) + centre - np.asarray(case.seeds[0][0], dtype=float)
                rec["tip_cells"] = [round(float(t), 4) for t in tip]
                if last_tip is not None and done > last_tip[0]:
                    v = (tip - last_tip[1]) * dx_phys / ((done - last_tip[0]) * dt_phys)
                    rec["tip_velocity"] = [round(float(x), 6) for x in v]
                last_tip = (done, tip)
        if p.thermal and p.flow_q and p.T_face[4] != p.T_face[5]:
            rec["nusselt"] = float(d.reduce(mkg.R_NUSSELT)[0])
        if p.flow_q:
            rec["umax"] = float(d.reduce(mkg.R_UMAX)[0])
        if p.solute:
            c = d.get(mkg.F_CONC)
            rec["solute_range"] = [float(c.min()), float(c.max())]
        rec["nonfinite"] = float(d.reduce(mkg.R_NONFINITE)[0])
        hist.append(rec)
        emit(json.dumps(rec))
        if done >= steps:
            break
        n = min(every, steps - done)
        d.sync()
        t0 = time.perf_counter()
        d.step(n, mkg.ALL)
        d.sync()
        t_dev += time.perf_counter() - t0
        done += n
    summary = {"case": name, "steps": steps, "cells": p.ncell, "precision": int(p.precision),
               "step_seconds": round(t_dev, 3), "mlups": round(p.ncell * steps / t_dev / 1e6, 1)}
    if name == "C1":
        u = d.get(mkg.F_VEL).reshape(p.nx, p.ny, p.nz, 3)[p.nx // 2, :, :, 0]
        nu = (1.0 / p.s_nu - 0.5) / 3.0
        ref = _duct_series(p.ny, p.nz, nu, p.force[0], steps)
        summary["duct_series_rel_l2"] = float(np.linalg.norm(u - ref) / np.linalg.norm(ref))
        summary["centre_velocity"] = float(u[p.ny // 2, p.nz // 2])
    d.close()
    emit(json.dumps(summary))
    return {"history": hist, "summary": summary}


def _duct_series(ny, nz, nu, g, t, terms=99):
    a, b = float(ny), float(nz)
    Y, Z = np.meshgrid(np.arange(ny) + 0.5, np.arange(nz) + 0.5, indexing="ij")
    u = np.zeros_like(Y)
    for m in range(1, terms + 1, 2):
        for n in range(1, terms + 1, 2):
            lam = nu * math.pi ** 2 * (m * m / a ** 2 + n * n / b ** 2)
            u += 16 * g / (math.pi ** 2 * m * n) / lam * (1 - math.exp(-lam * t)) * np.sin(
                m * math.pi * Y / a) * np.sin(n * math.pi * Z / b)
    return u


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("case")
    ap.add_argument("--steps", type=int)
    ap.add_argument("--frames", type=int)
    ap.add_argument("--fp64", action="store_true")
    a = ap.parse_args()
    run(a.case, a.steps, a.frames, mkg.PREC_FP64 if a.fp64 else None,
        emit=lambda s: print(s, flush=True))


if __name__ == "__main__":
    main()
