"""torch.autograd wrapper of the discrete adjoint (inverse problem of the phase-field +
thermal model).

``PhaseFieldMismatch.apply(eps4, kappa, setup, nsteps, ckpt, target)`` returns the FP64 loss
L = 1/2 sum (phi_N - target)^2 after ``nsteps`` fused steps started from the state that
``setup(solver)`` builds; its backward returns dL/deps4 and dL/dkappa computed by the
hand-written adjoint kernels (forward pass with checkpoints every ``ckpt`` steps, reverse
sweep with recompute).  ``setup`` receives a fresh ``mkg.Solver`` whose parameters already
carry eps4; kappa is uploaded before the run.
"""
from __future__ import annotations

import numpy as np
import torch

from . import mkg


class PhaseFieldMismatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps4, kappa, make_solver, nsteps, ckpt, target):
        e = float(eps4.detach().cpu().item())
        solver = make_solver(e)
        try:
            if kappa is not None:
                solver.set(mkg.F_KAPPA, kappa.detach().double().cpu().numpy())
            tgt = target.detach().double().cpu().numpy() if torch.is_tensor(target) else target
            out = solver.adjoint_run(nsteps, ckpt, tgt, want_kappa=kappa is not None)
        finally:
            solver.close()
        ctx.kappa_dtype = None if kappa is None else kappa.dtype
        ctx.kappa_device = None if kappa is None else kappa.device
        g_e = torch.tensor(out["dL_deps4"], dtype=torch.float64)
        g_k = None if kappa is None else torch.from_numpy(out["dL_dkappa"])
        ctx.save_for_backward(g_e, *( [g_k] if g_k is not None else []))
        ctx.has_k = g_k is not None
        ctx.timing = {k: out[k] for k in ("forward_ms", "recompute_ms", "reverse_ms")}
        return torch.tensor(out["loss"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gL):
        saved = ctx.saved_tensors
        g_e = saved[0] * gL.double().cpu()
        g_k = None
        if ctx.has_k:
            g_k = (saved[1] * gL.double().cpu()).to(ctx.kappa_dtype).to(ctx.kappa_device)
        return g_e, g_k, None, None, None, None
