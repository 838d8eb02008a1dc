"""Drag sensitivities of the FENE-P / IBM step behind torch.autograd.

``DragObjective.apply(wi, beta, L2, force)`` runs ``nsteps`` steps of a case
(FP64, or the mixed-precision step when the case asks for it) on the GPU from
its initial state and returns the time-averaged drag
J = sum_s w_s D_s with D_s = -sum_markers Fxc dX^3 of step s (the forces of
ibm_Helmholtz, fenep_mpi.f90:1845-1853).  Its backward runs the hand-written
discrete adjoint of every kernel of the step (fp_adj_backward) with
checkpointed recomputation and returns dJ/dWi, dJ/dbeta, dJ/d(L^2) and dJ/dF
for the body-force field F (3, NZ, NY, NX).  For a mixed-precision forward the
reverse sweep is the FP64 adjoint along FP64 recomputations of each segment
from the mixed step's checkpoints.

Import torch before this module's library is first used in a process (torch
ships its own HIP runtime).
"""
import math

import numpy as np
import torch

from . import api


class DragRun:
    """A case bound to a device context, with the objective's step weights."""

    def __init__(self, case, nsteps, every, avg_last):
        self.case = case
        self.nsteps = int(nsteps)
        self.every = int(every)
        self.weights = np.zeros(self.nsteps)
        self.weights[self.nsteps - int(avg_last):] = 1.0 / int(avg_last)
        self.g = api.FenepGPU(case.params)
        if case.markers is not None:
            self.g.set_markers(case.markers)
        self.replay = False

    def forward(self, wi, beta, L2, force):
        """J at (wi, beta, L2, force); force is a (3, nz, ny, nx) array."""
        g = self.g
        g.set_physics(wi, beta, math.sqrt(L2))
        g.set_state(self.case.state)
        f = np.ascontiguousarray(force, dtype=np.float64)
        g.set_body_force([f[0], f[1], f[2]])
        drag = g.adj_record(self.nsteps, self.every, replay=self.replay)
        return float(np.dot(self.weights, drag)), drag

    def close(self):
        self.g.close()


class DragObjective(torch.autograd.Function):
    """J(wi, beta, L2, force) for a DragRun; gradients by the discrete adjoint."""

    @staticmethod
    def forward(ctx, wi, beta, L2, force, run):
        J, _ = run.forward(float(wi), float(beta), float(L2), force.detach().cpu().numpy())
        ctx.run = run
        ctx.fshape = tuple(force.shape)
        ctx.fdevice = force.device
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        run = ctx.run
        fb = torch.zeros(ctx.fshape, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        w = run.weights * float(gJ)
        dp = run.g.adj_backward(weights=w, bar_force=fb.data_ptr())
        mk = lambda v: torch.tensor(v, dtype=torch.float64)
        return mk(dp[0]), mk(dp[1]), mk(dp[2]), fb.to(ctx.fdevice), None


def drag_objective(run, wi, beta, L2, force):
    return DragObjective.apply(wi, beta, L2, force, run)
