"""torch.autograd wrapper of the fluid adjoint, for verification and fitting.

``FluidRollout.apply(omega0, forcing, ctx)`` runs ``len(ctx.dts)`` steps of fluid_step
on the GPU from the initial vorticity ``omega0`` with forcing strength ``forcing`` and
returns the final vorticity; its backward calls the library's checkpointed adjoint
(fluid_adjoint_backward).  Tensors are CPU float64: the C ABI takes host arrays.

The library's adjoint keeps the checkpoints of the last forward pass only, so a context
serves one rollout at a time: a backward whose forward was followed by another forward
on the same context recomputes its own forward first (from the omega0 and forcing it
saved), so it never returns the gradient of the other rollout.
"""
import numpy as np
import torch

import ggb


class RolloutContext:
    """A fluid, its adjoint and a fixed dt schedule (dt is not differentiated)."""

    def __init__(self, fluid, dts, checkpoint_every=10):
        self.fluid = fluid
        self.dts = np.ascontiguousarray(dts, np.float64)
        self.adj = ggb.FluidAdjoint(fluid, checkpoint_every)
        self.generation = 0  # forward passes run on this context


class FluidRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, omega0, forcing, rc):
        rc.fluid.set_forcing(float(forcing))
        w0 = np.ascontiguousarray(omega0.detach().cpu().numpy(), np.float64)
        out = rc.adj.forward(w0, rc.dts)
        rc.generation += 1
        ctx.rc = rc
        ctx.forcing = float(forcing)
        ctx.omega0 = w0
        ctx.generation = rc.generation
        return torch.from_numpy(out)

    @staticmethod
    def backward(ctx, g_out):
        rc = ctx.rc
        rc.fluid.set_forcing(ctx.forcing)
        if rc.generation != ctx.generation:
            # another rollout ran on this context since: restore this one's checkpoints
            rc.adj.forward(ctx.omega0, rc.dts)
            rc.generation += 1
            ctx.generation = rc.generation
        g0, gk = rc.adj.backward(g_out.detach().cpu().numpy())
        return torch.from_numpy(g0), torch.tensor(gk, dtype=torch.float64), None


def area_loss(omega_T, target, area):
    """0.5 * sum(area * (omega_T - target)^2): the fitting loss of the tests."""
    return 0.5 * torch.sum(area * (omega_T - target) ** 2)
