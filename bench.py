"""Benchmark of the fused MaterKit time step on one MI355X.

Workload: D3Q19 TRT flow with Boussinesq buoyancy + D3Q7 thermal + D3Q19 anisotropic phase
field on a 256^3 grid, one solid seed of radius 8 at the centre (``cases/
fused_flow_thermal_dendrite_256.xml``); FP32 deviation storage.  A step is one pass of the
phase moment kernel and one pass of the fused collide-stream kernel over all cells.  The FP64
path is timed on the same workload for the FP32/FP64 ratio, and the FP64 CPU oracle (same
step, OpenMP) on a bounded slab of the same workload as the CPU baseline.

Prints one JSON line (rank 0).  Only one GPU is supported (--gpus > 1 exits).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

# The CPU baselines run the OpenMP oracle with its threads pinned to neighbouring cores; the
# OpenMP runtime reads these once, when the first library that uses it loads.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
# CPUs this process may use, read before the GPU runtime pins the main thread
HOST_CPUS = len(os.sched_getaffinity(0))

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell update of each kernel (each population / scalar read once and
# written once; flags read once; phase value read for the gradient stencil once per cell).
#   fused collide-stream: f 19 r+w, h 7 r+w, g 19 r+w, phi^n r, phi^{n-1} r, flag r
#   phase moment pass:    g 19 r, phi^n w, flag r
def alg_bytes(esz: int, fq=19, pq=19, th=7) -> tuple[int, int]:
    fused = 2 * esz * (fq + th + pq) + 2 * esz + 1
    moments = esz * pq + esz + 1
    return fused, moments


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--fp64-steps", type=int, default=30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--adjoint-ckpt", type=int, default=100)
    ap.add_argument("--flow-adjoint-steps", type=int, default=100)
    ap.add_argument("--no-mpf", action="store_true")
    ap.add_argument("--no-fvm", action="store_true")
    ap.add_argument("--no-mp", action="store_true")
    ap.add_argument("--mpf-steps", type=int, default=40)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one plain 1 GiB device copy (k_copy4, the reference kernel for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters of a --pmc run)")
    return ap.parse_args()


def copy_calibration():
    """One plain 1 GiB device copy by the library's copy kernel (k_copy4), the reference that
    scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    from materkit_anisotropiclbm_trt_rbc_facet_amd import mkg

    mkg.stream_copy_rate(1 << 30, 1)


def run_device(prec, size, steps, warmup, faces=False):
    import torch

    from materkit_anisotropiclbm_trt_rbc_facet_amd import cases, mkg

    case = cases.load("fused256", precision=prec, nx=size, ny=size, nz=size)
    if faces:
        # the same step with walls in z: bounce-back flow and phase faces, Dirichlet T faces
        case.params.set(bc_f=[0, 0, 0, 0, mkg.BC_BOUNCE, mkg.BC_BOUNCE],
                        bc_g=[0, 0, 0, 0, mkg.BC_BOUNCE, mkg.BC_BOUNCE],
                        bc_h=[0, 0, 0, 0, mkg.BC_DIRICHLET, mkg.BC_DIRICHLET],
                        T_face=[0, 0, 0, 0, -0.5, -0.6])
    dev = mkg.Solver(case.params)
    cases.setup(dev, case)
    dev.step(warmup, mkg.ALL)
    dev.sync()
    torch.cuda.synchronize()
    dev.enable_timing(True)
    t0 = time.perf_counter()
    dev.step(steps, mkg.ALL)
    dev.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tim = dev.last_timing()
    nonfinite = dev.reduce(mkg.R_NONFINITE)[0]
    solid = dev.reduce(mkg.R_SOLID)[0]
    # one output frame: the phase field copied to a host array (PCIe included)
    readback = []
    for _ in range(3):
        t1 = time.perf_counter()
        dev.get(mkg.F_PHI)
        readback.append(time.perf_counter() - t1)
    tim = dict(tim, readback_phi_s=min(readback))
    dev.close()
    return case.params, wall, tim, nonfinite, solid


def adjoint_cost(steps: int, ckpt: int):
    """Gradient cost of the inverse problem of configuration C5 (phase field D3Q19 + thermal
    D3Q7, 128^3, FP32, per-cell kappa): `steps` steps with checkpoints every `ckpt` steps and
    per-segment recompute (the C5 prescription is 2000 and 100), against a plain forward of the
    same steps.  Three ratios to the plain forward time F:
      beyond_forward   (checkpointed gradient run - F) / F: recompute + reverse + saves
      reverse_only     (recompute + reverse) / F
      whole_gradient   checkpointed gradient run / F
    The tape run (ckpt 0: the forward keeps every step's tape, no recompute) is given beside."""
    import torch

    from materkit_anisotropiclbm_trt_rbc_facet_amd import cases, mkg

    case = cases.load("C5", precision=mkg.PREC_FP32, eps4=0.05)
    kap = cases.lognormal_field((128, 128, 128), *case.conductivity).reshape(-1)

    def fresh():
        dev = mkg.Solver(case.params)
        cases.setup(dev, case)
        dev.set(mkg.F_KAPPA, kap)
        return dev

    dev = fresh()
    dev.step(2, mkg.ALL)
    dev.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.step(steps, mkg.ALL)
    dev.sync()
    fwd = time.perf_counter() - t0
    target = dev.phi_now()
    dev.close()
    runs = {}
    for name, ck in (("checkpointed", ckpt), ("tape", 0)):
        dev = fresh()
        dev.step(2, mkg.ALL)
        out = dev.adjoint_run(steps, ck, target + 0.01)
        dev.close()
        f, r, v = (out["forward_ms"] * 1e-3, out["recompute_ms"] * 1e-3,
                   out["reverse_ms"] * 1e-3)
        runs[name] = {"forward_s": round(f, 4), "recompute_s": round(r, 4),
                      "reverse_s": round(v, 4), "beyond_forward": round((f + r + v - fwd) / fwd, 3),
                      "reverse_only": round((r + v) / fwd, 3),
                      "whole_gradient": round((f + r + v) / fwd, 3)}
    return {"forward_s": round(fwd, 4), "ratio": runs["checkpointed"]["beyond_forward"],
            "checkpointed": runs["checkpointed"], "tape": runs["tape"], "steps": steps,
            "checkpoint_every": ckpt, "cells": 128 ** 3,
            "workload": "C5 inverse problem: phase field (D3Q19) + thermal (D3Q7), 128^3, FP32, "
                        "per-cell kappa"}


def mpf_cost(size: int, steps: int) -> dict:
    """Multi-phase field (two order parameters on D3Q19 + D3Q7 thermal, FP32) on a size^3 grid:
    two grains, one carried at 0.02 in x and repelled by the other (the run_multiPhaseField
    set-up in 3D).  Step time from the wall clock around `steps` steps on a synchronised
    stream; algorithmic bytes: moment pass 19 n + (n + 1) values + flag, collide-stream
    2 x 19 n + 2 x 7 + n + 2 values + flag."""
    import torch

    from materkit_anisotropiclbm_trt_rbc_facet_amd import cases, mkg

    n = 2
    m = mkg.default_mpf_params(nx=size, ny=size, nz=size, n=n, precision=mkg.PREC_FP32,
                               thermal=1, alpha=0.1, latent_T=0.5)
    c = size / 2.0
    phi = cases.mpf_seed_fields((size, size, size), [[c, c, c], [c, c - 16, c]], [8, 8], m.wdx)
    dev = mkg.MpfSolver(m)
    dev.init(phi)
    del phi
    dev.step(4)
    dev.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.step(steps)
    dev.sync()
    dt = (time.perf_counter() - t0) / steps
    dev.close()
    N = size ** 3
    e = 4
    bytes_cell = (19 * n + n + 1) * e + 1 + (2 * 19 * n + 2 * 7 + n + 2) * e + 1
    gbs = bytes_cell * N / dt / 1e9
    return {"workload": f"multi-phase field, {n} order parameters (D3Q19) + thermal (D3Q7), "
                        f"{size}^3, FP32", "mlups": round(N / dt / 1e6, 1),
            "ms_per_step": round(dt * 1e3, 4), "alg_bytes_per_cell": bytes_cell,
            "step_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}


def fvm_cost(size: int, steps: int) -> dict:
    """Explicit finite-volume phase field + solute with anti-trapping current (FP32) on a size^3
    grid, one seed at the centre growing into U = -0.5.  Algorithmic bytes per cell: phase
    kernel phi, c read and phi' written; solute kernel phi, phi', c read and c' written."""
    import torch

    from materkit_anisotropiclbm_trt_rbc_facet_amd import mkg

    p = mkg.default_params(nx=size, ny=size, nz=size, phase_q=19, solute=1, pf_coupling=2,
                           eps4=0.05, pf_dt=0.008, pf_wdx2=6.25, D_l=0.1, D_s=0.0, k_part=0.15,
                           q_min=1e-3, at_coef=0.5 / math.sqrt(2.0), precision=mkg.PREC_FP32,
                           **{"lambda": 3.19})
    x = np.arange(size) - size / 2.0
    r = np.sqrt(x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2)
    phi = -np.tanh((r - 8.0) / (math.sqrt(2.0) * 2.5)).reshape(-1)
    dev = mkg.FvmSolver(p)
    dev.set(mkg.FVM_PHI, phi)
    dev.set(mkg.FVM_CONC, -0.5 * 0.5 * ((1 + 0.15) - (1 - 0.15) * phi))
    del phi, r
    dev.step(4)
    dev.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.step(steps)
    dev.sync()
    dt = (time.perf_counter() - t0) / steps
    dev.close()
    N = size ** 3
    bytes_cell = (3 + 4) * 4
    gbs = bytes_cell * N / dt / 1e9
    return {"workload": f"finite-volume phase field + solute (anti-trapping), {size}^3, FP32",
            "mlups": round(N / dt / 1e6, 1), "ms_per_step": round(dt * 1e3, 4),
            "alg_bytes_per_cell": bytes_cell, "step_gbs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def flow_adjoint_cost(steps: int) -> dict:
    """Flow adjoint (mkg_flow_adjoint_run) on the duct of the Poiseuille case at 128 x 64 x 64
    (D3Q19 BGK, FP32, body force, bounce walls): ratio = (adjoint run - plain forward) / plain
    forward.  The adjoint run tapes every step's populations in FP64 and reverses in FP64."""
    import torch

    from materkit_anisotropiclbm_trt_rbc_facet_amd import cases, mkg

    case = cases.load("C1", precision=mkg.PREC_FP32)
    p = case.params
    p.nx, p.ny, p.nz = 128, 64, 64

    def fresh():
        dev = mkg.Solver(p)
        dev.init_equilibrium()
        dev.step(2, mkg.FLOW)
        return dev

    dev = fresh()
    dev.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.step(steps, mkg.FLOW)
    dev.sync()
    fwd = time.perf_counter() - t0
    target = dev.get(mkg.F_VEL) * 1.01
    dev.close()
    dev = fresh()
    dev.sync()
    t0 = time.perf_counter()
    r = dev.flow_adjoint_run(steps, target)
    adj = time.perf_counter() - t0
    dev.close()
    return {"workload": "flow adjoint, duct 128x64x64, D3Q19 BGK, FP32 forward and reverse, tape written by the forward step",
            "steps": steps, "forward_s": round(fwd, 4), "adjoint_run_s": round(adj, 4),
            "ratio": round((adj - fwd) / fwd, 3), "loss": r["loss"]}


def mp_cost(size: int, steps: int) -> dict:
    """Phase-field two-phase flow (Cahn-Hilliard D3Q19 + velocity-pressure D3Q19, FP32) on a
    size^3 grid: a drop of radius size/4, density ratio 3, under gravity.  Algorithmic bytes per
    cell: phi pass 19 r + 1 w; mu pass 1 r + 1 w; collide-stream 2 x 19 r+w, phi and mu r,
    phi u r+w (3 values)."""
    import torch

    from materkit_anisotropiclbm_trt_rbc_facet_amd import mkg

    m = mkg.mp_params(nx=size, ny=size, nz=size, precision=mkg.PREC_FP32, rho_l=3.0, rho_g=1.0,
                      nu_l=0.1, nu_g=0.05, sigma=0.005, width=4.0, mobility=0.05, tau_h=1.0)
    m.gravity[2] = -1e-6
    x = np.arange(size) - size / 2.0
    r = np.sqrt(x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2)
    phi = (0.5 + 0.5 * np.tanh(2 * (size / 4.0 - r) / 4.0)).reshape(-1)
    dev = mkg.MpSolver(m)
    dev.init(phi)
    del phi, r
    dev.step(4)
    dev.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.step(steps)
    dev.sync()
    dt = (time.perf_counter() - t0) / steps
    dev.close()
    N = size ** 3
    bytes_cell = (19 + 1 + 1 + 1 + 2 * 2 * 19 + 2 + 2 * 3) * 4
    gbs = bytes_cell * N / dt / 1e9
    return {"workload": f"two-phase flow (Cahn-Hilliard D3Q19 + velocity-pressure D3Q19), "
                        f"{size}^3, FP32", "mlups": round(N / dt / 1e6, 1),
            "ms_per_step": round(dt * 1e3, 4), "alg_bytes_per_cell": bytes_cell,
            "step_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}


def host_cpu() -> dict:
    """The host the CPU baselines ran on: CPU model, CPUs this process may use, binding."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "cpus_available": HOST_CPUS,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}


def cpu_baseline(seconds: float):
    """The reference-structured FP64 CPU path (oracle/mk_aos.c: AoS node records of the
    reference's 3640 B layout, one OpenMP sweep per operator) on a 256 x 256 x 16 slab of the
    same workload; the SoA oracle on the same slab for context."""
    from materkit_anisotropiclbm_trt_rbc_facet_amd import cases, mkg
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", str(min(16, os.cpu_count() or 1))))
    case = cases.load("fused256", precision=mkg.PREC_FP64, nz=16)
    p = case.params
    st = O.State.empty(p)
    cases.setup(st, case)  # same initial fields as the device (seed sphere cut by the slab)
    N = p.ncell
    aos = st.copy()
    t1 = aos.aos_step(1, O.OP_TRT)  # warm-up and per-step estimate
    n_aos = max(1, int(0.6 * seconds / max(t1, 1e-6)))
    dt_aos = aos.aos_step(n_aos, O.OP_TRT)
    del aos
    st.step(1, 15, O.OP_TRT)  # warm-up
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.4 * seconds:
        st.step(1, 15, O.OP_TRT)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": round(N * n_aos / dt_aos / 1e6, 3), "unit": "MLUPS", "cores": threads,
            "kind": "port", "host": host_cpu(),
            "sample": f"reference-structured FP64 path (AoS 3640 B nodes, a sweep per "
                      f"operator), fused flow+thermal+phase step, {p.nx}x{p.ny}x{p.nz} slab of "
                      f"the 256^3 workload, {n_aos} steps in {dt_aos:.1f} s",
            "soa_oracle_mlups": round(N * n / dt / 1e6, 3)}


def cpu_other_solvers(seconds: float) -> dict:
    """The FP64 CPU oracle of the other solvers on bounded samples (64^3 grids, about
    seconds/3 each): multi-phase field (two order parameters + thermal), finite-volume phase
    field + solute, two-phase flow.  MLUPS on the same host cores as the fused baseline."""
    from materkit_anisotropiclbm_trt_rbc_facet_amd import cases, mkg
    from oracle import oracle as O

    n = 64
    N = n ** 3
    out = {}

    def timed(fn):
        fn(1)  # warm-up and estimate
        t0 = time.perf_counter()
        fn(1)
        t1 = max(time.perf_counter() - t0, 1e-6)
        k = max(1, int(seconds / 3 / t1))
        t0 = time.perf_counter()
        fn(k)
        return round(N * k / (time.perf_counter() - t0) / 1e6, 3), k

    m = mkg.default_mpf_params(nx=n, ny=n, nz=n, n=2, thermal=1, alpha=0.1, latent_T=0.5)
    c = n / 2.0
    ms = O.MpfState.start(m, cases.mpf_seed_fields((n, n, n), [[c, c, c], [c, c - 12, c]],
                                                   [6, 6], m.wdx))
    v, k = timed(ms.step)
    out["multi_phase_field"] = {"mlups": v, "steps": k, "grid": f"{n}^3"}
    p = mkg.default_params(nx=n, ny=n, nz=n, phase_q=19, solute=1, pf_coupling=2, eps4=0.05,
                           pf_dt=0.008, pf_wdx2=6.25, D_l=0.1, k_part=0.15, q_min=1e-3,
                           at_coef=0.5 / math.sqrt(2.0), **{"lambda": 3.19})
    x = np.arange(n) - n / 2.0
    r = np.sqrt(x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2)
    phi = np.ascontiguousarray(-np.tanh((r - 8.0) / (math.sqrt(2.0) * 2.5)).reshape(-1))
    conc = -0.5 * 0.5 * ((1 + 0.15) - (1 - 0.15) * phi)
    v, k = timed(lambda s: O.fvm_step(p, phi, conc, s))
    out["finite_volume"] = {"mlups": v, "steps": k, "grid": f"{n}^3"}
    mp = mkg.mp_params(nx=n, ny=n, nz=n, rho_l=3.0, rho_g=1.0, nu_l=0.1, nu_g=0.05,
                       sigma=0.005, width=4.0, mobility=0.05, tau_h=1.0)
    st = O.MpState.start(mp, 0.5 + 0.5 * np.tanh(2 * (n / 4.0 - r.reshape(-1)) / 4.0))
    v, k = timed(st.step)
    out["two_phase"] = {"mlups": v, "steps": k, "grid": f"{n}^3"}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              f"(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    from materkit_anisotropiclbm_trt_rbc_facet_amd import mkg

    size = a.size
    N = size ** 3
    if a.copy_calibration:
        copy_calibration()
    p32, wall32, tim32, nf32, solid32 = run_device(mkg.PREC_FP32, size, a.steps, a.warmup)
    if a.no_fp64:
        p64, wall64, tim64, nf64 = None, None, None, 0.0
    else:
        p64, wall64, tim64, nf64, _ = run_device(mkg.PREC_FP64, size, a.fp64_steps, 3)

    b32, m32 = alg_bytes(4)
    b64, m64 = alg_bytes(8)
    mlups32 = N * a.steps / wall32 / 1e6
    k32 = tim32["fused_ms"] / tim32["steps"] * 1e-3  # s per fused launch
    km32 = tim32["phase_moments_ms"] / tim32["steps"] * 1e-3
    ach32 = b32 * N / k32 / 1e9
    traffic = None
    tf = ROOT / "profiles" / "traffic_fused256.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("fused_bytes_per_launch")
        except Exception:
            traffic = None
    out = {
        "metric": "MLUPS (fused D3Q19-TRT + D3Q7 + D3Q19 phase-field step, 256^3)",
        "value": round(mlups32, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall32 / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "fused flow(D3Q19 TRT)+thermal(D3Q7)+phase field(D3Q19), "
                               f"{size}^3, one seed r=8, FP32 deviation storage",
                   "cells": N},
        "roofline": {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach32 / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": "profiles/traffic_fused256.json: rocprofv3 FETCH_SIZE + "
                                       "WRITE_SIZE of this kernel at this size, scaled by the "
                                       "same counters of a plain copy (not measured in this run)",
                     "kernel": "k_step<float,19,19,thermal> (fused collide-stream)",
                     "alg_bytes_per_cell": b32},
        "kernel_ms": {"fused": round(k32 * 1e3, 4), "phase_moments": round(km32 * 1e3, 4)},
        "fp64": None,
        "fp32_over_fp64": None,
        "adjoint_over_forward": None,
        "nonfinite": nf32 + nf64,
        "solid_fraction": solid32,
    }
    # host hand-over: the state stays in HBM; a run hands phi to the host once per output
    # frame (every 100 steps in the reference's dendrite cases: numLoops / numFrames)
    rb = tim32["readback_phi_s"]
    out["pcie_inclusive"] = {
        "readback_phi_ms": round(rb * 1e3, 3),
        "mlups_with_phi_frame_every_100_steps": round(N * 100 / (100 * wall32 / a.steps + rb) / 1e6, 1)}
    if tim64 is not None:
        mlups64 = N * a.fp64_steps / wall64 / 1e6
        k64 = tim64["fused_ms"] / tim64["steps"] * 1e-3
        ach64 = b64 * N / k64 / 1e9
        out["fp64"] = {"value": round(mlups64, 1), "unit": "MLUPS",
                       "fused_kernel_gbs": round(ach64, 1), "frac": round(ach64 / HBM_PEAK_GBS, 4),
                       "alg_bytes_per_cell": b64}
        out["fp32_over_fp64"] = round(mlups32 / mlups64, 3)
    # the face-rule body: walls in z (every z-row block holds a face cell)
    _, wallf, timf, nff, _ = run_device(mkg.PREC_FP32, size, max(10, a.steps // 5), 3,
                                        faces=True)
    kf = timf["fused_ms"] / timf["steps"] * 1e-3
    out["faces_variant"] = {
        "config": "the same step with bounce-back flow/phase and Dirichlet T faces in z",
        "value": round(N * max(10, a.steps // 5) / wallf / 1e6, 1), "unit": "MLUPS",
        "fused_kernel_frac": round(b32 * N / kf / 1e9 / HBM_PEAK_GBS, 4),
        "nonfinite": nff}
    if not a.no_mpf:
        out["mpf"] = mpf_cost(size, a.mpf_steps)
    if not a.no_fvm:
        out["fvm"] = fvm_cost(size, a.mpf_steps)
    if not a.no_mp:
        out["two_phase"] = mp_cost(size, a.mpf_steps)
    if not a.no_adjoint:
        adj = adjoint_cost(a.adjoint_steps, a.adjoint_ckpt)
        out["adjoint_over_forward"] = adj["ratio"]
        out["adjoint"] = adj
        out["flow_adjoint"] = flow_adjoint_cost(a.flow_adjoint_steps)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
        out["cpu_baseline"]["other_solvers"] = cpu_other_solvers(a.cpu_seconds)
    # the whole step on the benchmark's own definition: b32 algorithmic bytes per cell over the
    # step time (k_step plus the phase-moment pass, which re-reads the phase populations)
    copy_gbs = mkg.stream_copy_rate(1 << 30, 10)
    step_s = wall32 / a.steps
    out["whole_step"] = {
        "frac": round(b32 * N / step_s / 1e9 / HBM_PEAK_GBS, 4),
        "alg_bytes_per_cell": b32,
        "moved_bytes_per_cell": b32 + m32,
        "copy_gbs": round(copy_gbs, 1),
        "two_pass_ceiling_frac": round(b32 / (b32 + m32) * copy_gbs / HBM_PEAK_GBS, 4),
        "note": "frac = alg bytes x cells / step time / 8 TB/s; the step moves moved_bytes "
                "(k_step plus the phase-moment pass), so at the measured copy rate it cannot "
                "exceed two_pass_ceiling_frac"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
