"""Trajectory gradients of the 3D step through torch.autograd.

`trajectory(tissue, Kv, Ka, Kb, Ks, X0, nsteps, dt, every)` advances a
`capi.Tissues3D` handle from X0 for nsteps Euler steps with per-cell
parameters (Kv, Ka, Kb, Ks) and returns the final positions as a tensor;
its backward is the hand-written discrete adjoint of the step
(dpm3d_adjoint: reverse sweep over segments recomputed from checkpoints
every `every` steps), so any torch loss of the final positions
differentiates with respect to the per-cell parameters and the initial
positions.  The computation runs on the handle's device through the C ABI;
torch carries the tensors and the autograd graph.
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Kv, Ka, Kb, Ks, X0, tissue, nsteps, dt, every):
        np_dtype = np.float64 if tissue.precision == 64 else np.float32
        host = lambda t: t.detach().to("cpu", torch.float64).numpy()
        tissue.set_cell_params(Kv=host(Kv), Ka=host(Ka), Kb=host(Kb), Ks=host(Ks))
        tissue.set_positions(host(X0).astype(np_dtype))
        tissue.trajectory(int(nsteps), float(dt), int(every))
        ctx.tissue = tissue
        ctx.x_dtype = X0.dtype
        XN = tissue.positions(np.float64)
        return torch.as_tensor(XN, dtype=X0.dtype, device=X0.device)

    @staticmethod
    def backward(ctx, gX):
        lam = gX.detach().to("cpu", torch.float64).numpy()
        dth, l0 = ctx.tissue.adjoint(lam)
        dev = gX.device
        g = torch.as_tensor(dth, device=dev)
        dX0 = torch.as_tensor(np.asarray(l0, np.float64), dtype=ctx.x_dtype, device=dev)
        return g[:, 0], g[:, 1], g[:, 2], g[:, 3], dX0, None, None, None, None


def trajectory(tissue: "capi.Tissues3D", Kv, Ka, Kb, Ks, X0, nsteps: int, dt: float, every: int = 20):
    """Final positions [nverts, 3] after nsteps steps; differentiable in
    Kv, Ka, Kb, Ks (per cell, [C]) and X0 ([nverts, 3])."""
    return _Trajectory.apply(Kv, Ka, Kb, Ks, X0, tissue, nsteps, dt, every)


def com_displacement_loss(XN, X0, ncells: int):
    """Mean over cells of |COM_N - COM_0| (torch, differentiable)."""
    d = XN.reshape(ncells, -1, 3).mean(1) - X0.reshape(ncells, -1, 3).mean(1)
    return torch.linalg.norm(d, dim=1).mean()
