"""torch.autograd.Function over the acoustic solver: receiver traces as a
differentiable function of the material coefficients (kappa on MM, 1/rho on
NM and MN).  Forward = checkpointed forward run (tw_adjoint_forward),
backward = the hand-written discrete adjoint (tw_adjoint_run) with the
incoming trace gradient as the receiver sources.

The solver keeps its own configuration, source and receivers; each forward
starts from the state captured when the function was built.
"""
import numpy as np
import torch

from . import C_BINV_VX, C_BINV_VY, C_KAPPA


class _Traces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kappa, binv_vx, binv_vy, op):
        op.solver.restore(op.state0)
        for k, t in ((C_KAPPA, kappa), (C_BINV_VX, binv_vx), (C_BINV_VY, binv_vy)):
            op.solver.set_coef(k, t.detach().cpu().double().numpy())
        op.solver.adjoint_forward(op.nsteps, op.interval)
        tr = op.solver.receivers()[:, -op.nsteps:]
        ctx.op = op
        return torch.from_numpy(np.ascontiguousarray(tr)).to(kappa.dtype)

    @staticmethod
    def backward(ctx, grad):
        op = ctx.op
        gk, gx, gy = op.solver.adjoint_run(grad.detach().cpu().double().numpy())
        return torch.from_numpy(gk), torch.from_numpy(gx), torch.from_numpy(gy), None


class AcousticTraces:
    """traces = AcousticTraces(solver, nsteps)(kappa, binv_vx, binv_vy)"""

    def __init__(self, solver, nsteps, interval=0):
        self.solver = solver
        self.nsteps = int(nsteps)
        self.interval = int(interval)
        self.state0 = solver.checkpoint()

    def __call__(self, kappa, binv_vx, binv_vy):
        return _Traces.apply(kappa, binv_vx, binv_vy, self)
