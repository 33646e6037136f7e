"""Forward/backward interface of the torque adjoint, in the shape of a
torch.autograd.Function: J = mean shaft torque of a stirred tank over the
last `window` of `steps` runner steps, as a function of the impeller
rotation rate and a per-cell viscosity field.  The forward and backward
passes run on the device through lbm_adjoint_torque (hand-written adjoint
kernels, binomial checkpointing); torch only carries the tensors."""
from __future__ import annotations

import numpy as np
import torch


class TankTorque(torch.autograd.Function):
    """TankTorque.apply(omega, nu, make) -> J.

    omega: 0-d tensor, impeller rotation rate (lattice units).
    nu: tensor of the grid's cell count, kinematic viscosity per cell.
    make: callable(omega: float) -> (Solver, steps, window, snapshots) that
    builds the tank at that rotation rate from rest (same geometry each
    call); the viscosity field is applied here as the solver's omega field.
    """

    @staticmethod
    def forward(ctx, omega, nu, make):
        solver, steps, window, snaps = make(float(omega))
        nu_np = nu.detach().cpu().double().numpy().reshape(-1)
        solver.set_omega_field((1.0 / (3.0 * nu_np + 0.5)).astype(np.float32))
        need = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        out = solver.adjoint_torque(steps, window, snaps, gradients=need)
        if need:
            ctx.save_for_backward(torch.tensor(out["dJ_domega"], dtype=torch.float64),
                                  torch.from_numpy(out["dJ_dnu"]).reshape(nu.shape))
        solver.close()
        return torch.tensor(out["J"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        dw, dnu = ctx.saved_tensors
        return gJ * dw, gJ * dnu, None
