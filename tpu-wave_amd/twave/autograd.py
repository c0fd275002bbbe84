"""torch.autograd.Function over the solver: receiver traces as a
differentiable function of the material coefficients.  Acoustic: kappa on MM,
1/rho on NM and MN (or kappa and rho itself).  Elastic: lambda and 2 mu on
MM, mu on NN, 1/rho on NM and MN (or lambda, mu and rho).  Either physics:
the model of set_model, rho, Vp and Vs on the model nodes (ModelTraces).
Forward =
checkpointed forward run (tw_adjoint_forward), backward = the hand-written
discrete adjoint (tw_adjoint_run_coefs) with the incoming trace gradient as
the receiver sources.

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

from . import C_BINV_VX, C_BINV_VY, C_KAPPA, C_LAMBDA, C_MU, C_TWOMU, ELASTIC, model_gradient

_ids = itertools.count(1)

ACOUSTIC_COEFS = (C_KAPPA, C_BINV_VX, C_BINV_VY)
ELASTIC_COEFS = (C_LAMBDA, C_TWOMU, C_MU, C_BINV_VX, C_BINV_VY)


def _run_forward(op, coefs):
    op.solver.restore(op.state0)
    for k, v in zip(op.coefs, coefs):
        op.solver.set_coef(k, v)
    op.solver.adjoint_forward(op.nsteps, op.interval)


class _Traces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, op, *tensors):
        coefs = [t.detach().to("cpu", torch.float64).numpy().copy() for t in tensors]
        _run_forward(op, coefs)
        fid = next(_ids)
        op.last_forward = fid
        tr = op.solver.receivers()[:, -op.nsteps:]
        ctx.op, ctx.fid, ctx.coefs = op, fid, coefs
        like = tensors[0]
        ctx.like = like
        return torch.from_numpy(np.ascontiguousarray(tr)).to(device=like.device, dtype=like.dtype)

    @staticmethod
    def backward(ctx, grad):
        op = ctx.op
        if op.last_forward != ctx.fid:
            # the solver's checkpoints are those of a later forward: recompute ours
            _run_forward(op, ctx.coefs)
            op.last_forward = ctx.fid
        g = op.solver.adjoint_run_coefs(grad.detach().to("cpu", torch.float64).numpy())
        like = ctx.like
        return (None,) + tuple(torch.from_numpy(g[k]).to(device=like.device, dtype=like.dtype) for k in op.coefs)


class _Op:
    coefs = ()

    def __init__(self, solver, nsteps, interval=0):
        self.solver = solver
        self.nsteps = int(nsteps)
        self.interval = int(interval)
        self.state0 = solver.checkpoint()
        self.last_forward = 0


class AcousticTraces(_Op):
    """traces = AcousticTraces(solver, nsteps)(kappa, binv_vx, binv_vy)"""
    coefs = ACOUSTIC_COEFS

    def __call__(self, kappa, binv_vx, binv_vy):
        return _Traces.apply(self, kappa, binv_vx, binv_vy)


class AcousticTracesKappaRho(AcousticTraces):
    """traces = AcousticTracesKappaRho(solver, nsteps)(kappa, rho_vx, rho_vy):
    the same operator parametrized by the density on the two velocity grids;
    the chain rule d/d rho = -(1/rho^2) d/d(1/rho) is applied by torch."""

    def __call__(self, kappa, rho_vx, rho_vy):
        return _Traces.apply(self, kappa, 1.0 / rho_vx, 1.0 / rho_vy)


class ElasticTraces(_Op):
    """traces = ElasticTraces(solver, nsteps)(lam, twomu, mu, binv_vx, binv_vy):
    lambda and 2 mu on MM, mu on NN, 1/rho on NM and MN, as the solver's
    coefficient arrays."""
    coefs = ELASTIC_COEFS

    def __call__(self, lam, twomu, mu, binv_vx, binv_vy):
        return _Traces.apply(self, lam, twomu, mu, binv_vx, binv_vy)


class ElasticTracesLameRho(ElasticTraces):
    """traces = ElasticTracesLameRho(solver, nsteps)(lam, mu_mm, mu_nn, rho_vx,
    rho_vy): density and the Lame moduli, mu on the cells (MM, where the
    solver keeps 2 mu) and on the nodes (NN); torch applies the chain rule."""

    def __call__(self, lam, mu_mm, mu_nn, rho_vx, rho_vy):
        return _Traces.apply(self, lam, 2.0 * mu_mm, mu_nn, 1.0 / rho_vx, 1.0 / rho_vy)


class _ModelTraces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, op, rho, vp, vs):
        model = [t.detach().to("cpu", torch.float64).numpy().copy() for t in (rho, vp, vs)]
        op.solver.set_model(*model, op.num_dxm, op.den_dxm)
        coefs = [op.solver.get_coef(k) for k in op.coefs]
        _run_forward(op, coefs)
        fid = next(_ids)
        op.last_forward = fid
        tr = op.solver.receivers()[:, -op.nsteps:]
        ctx.op, ctx.fid, ctx.coefs, ctx.model, ctx.like = op, fid, coefs, model, rho
        return torch.from_numpy(np.ascontiguousarray(tr)).to(device=rho.device, dtype=rho.dtype)

    @staticmethod
    def backward(ctx, grad):
        op = ctx.op
        if op.last_forward != ctx.fid:
            _run_forward(op, ctx.coefs)
            op.last_forward = ctx.fid
        g = op.solver.adjoint_run_coefs(grad.detach().to("cpu", torch.float64).numpy())
        rho, vp, vs = ctx.model
        out = model_gradient(op.solver.cfg, g, rho, vp, vs if op.solver.cfg.physics == ELASTIC else None,
                             op.num_dxm, op.den_dxm)
        gs = out[2] if out[2] is not None else np.zeros_like(rho)
        like = ctx.like
        return (None,) + tuple(torch.from_numpy(np.ascontiguousarray(a)).to(device=like.device, dtype=like.dtype)
                               for a in (out[0], out[1], gs))


class ModelTraces(_Op):
    """traces = ModelTraces(solver, nsteps, num_dxm, den_dxm)(rho, vp, vs): the
    model of set_model on its (nxm+1) x (nym+1) nodes with spacing
    num_dxm/den_dxm (the reference's inversion parameters rho, Vp, Vs,
    cuda/src/cpp/inverse.hpp:10); each forward sets the model, backward maps
    the coefficient gradients to the nodes (twave.model_gradient).  Acoustic
    solvers ignore vs (its gradient is zero)."""

    def __init__(self, solver, nsteps, num_dxm, den_dxm, interval=0):
        super().__init__(solver, nsteps, interval)
        self.num_dxm, self.den_dxm = int(num_dxm), int(den_dxm)
        self.coefs = ELASTIC_COEFS if solver.cfg.physics == ELASTIC else ACOUSTIC_COEFS

    def __call__(self, rho, vp, vs):
        return _ModelTraces.apply(self, rho, vp, vs)
