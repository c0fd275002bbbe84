"""torch.autograd.Function over the acoustic solver: receiver traces as a
differentiable function of the material coefficients (kappa on MM, 1/rho on
NM and MN, or kappa and rho itself).  Forward = checkpointed forward run
(tw_adjoint_forward), backward = the hand-written discrete adjoint
(tw_adjoint_run) with the incoming trace gradient as the receiver sources.

The solver keeps its own configuration, source and receivers; each forward
starts from the state captured when the operator was built.  The C-ABI takes
host arrays, so coefficients and gradients cross to the host (FP64) at each
call; outputs come back on the device and dtype of the input.

Several forwards may be taken before their backwards (two models, a line
search, a sum of losses): each backward checks that the solver's checkpoints
belong to its own forward and, if not, reruns that forward first.
"""
import itertools

import numpy as np
import torch

from . import C_BINV_VX, C_BINV_VY, C_KAPPA

_ids = itertools.count(1)


def _run_forward(op, coefs):
    op.solver.restore(op.state0)
    for k, v in zip((C_KAPPA, C_BINV_VX, C_BINV_VY), coefs):
        op.solver.set_coef(k, v)
    op.solver.adjoint_forward(op.nsteps, op.interval)


class _Traces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kappa, binv_vx, binv_vy, op):
        coefs = [t.detach().to("cpu", torch.float64).numpy().copy() for t in (kappa, binv_vx, binv_vy)]
        _run_forward(op, coefs)
        fid = next(_ids)
        op.last_forward = fid
        tr = op.solver.receivers()[:, -op.nsteps:]
        ctx.op, ctx.fid, ctx.coefs = op, fid, coefs
        ctx.like = kappa
        return torch.from_numpy(np.ascontiguousarray(tr)).to(device=kappa.device, dtype=kappa.dtype)

    @staticmethod
    def backward(ctx, grad):
        op = ctx.op
        if op.last_forward != ctx.fid:
            # the solver's checkpoints are those of a later forward: recompute ours
            _run_forward(op, ctx.coefs)
            op.last_forward = ctx.fid
        g = op.solver.adjoint_run(grad.detach().to("cpu", torch.float64).numpy())
        like = ctx.like
        return tuple(torch.from_numpy(x).to(device=like.device, dtype=like.dtype) for x in g) + (None,)


class AcousticTraces:
    """traces = AcousticTraces(solver, nsteps)(kappa, binv_vx, binv_vy)"""

    def __init__(self, solver, nsteps, interval=0):
        self.solver = solver
        self.nsteps = int(nsteps)
        self.interval = int(interval)
        self.state0 = solver.checkpoint()
        self.last_forward = 0

    def __call__(self, kappa, binv_vx, binv_vy):
        return _Traces.apply(kappa, binv_vx, binv_vy, self)


class AcousticTracesKappaRho(AcousticTraces):
    """traces = AcousticTracesKappaRho(solver, nsteps)(kappa, rho_vx, rho_vy):
    the same operator parametrized by the density on the two velocity grids;
    the chain rule d/d rho = -(1/rho^2) d/d(1/rho) is applied by torch."""

    def __call__(self, kappa, rho_vx, rho_vy):
        return _Traces.apply(kappa, 1.0 / rho_vx, 1.0 / rho_vy, self)
