"""Inversion driver on top of the adjoint: misfit bookkeeping per shot and per
receiver, aggregated over shots as the reference's Class_Inverse_Specs does
(aggregate_data_misfit, cuda/src/cpp/inverse.hpp:34-43; per-receiver misfit
Class_Forward_Specs::calculate_data_misfit, cuda/src/cpp/forward.cpp:414-445),
and a model update loop the reference leaves unwritten: steepest descent with
a backtracking line search on the bulk modulus (acoustic) or on the elastic
moduli.

The problem object is a twave.workloads.AdjointShots or ElasticShots (or
anything with their `run(coefs, shot, residual_of)`, `misfit`, `shots`,
`data`, `m0`, `s` interface).
"""
import numpy as np


class ShotMisfit:
    """Misfit of one shot: per receiver record and total (tw_data_misfit)."""

    def __init__(self, per_record, total):
        self.per_record = per_record
        self.total = total


def shot_misfits(problem, coefs):
    """Forward run of every shot with coefs; the reference's per-receiver
    misfits of each shot and the aggregate over shots."""
    out, agg = [], 0.0
    for i in range(len(problem.shots)):
        problem.run(coefs, i)
        per, tot = problem.s.data_misfit(problem.data[i])
        out.append(ShotMisfit(per, tot))
        agg += tot
    return out, agg


class SteepestDescent:
    """Update the coefficients `update` (indices into the problem's
    coefficients: kappa, index 0, by default; for ElasticShots e.g. (0, 1, 2),
    lambda, 2 mu and mu) along the negative misfit gradient, one step length
    for all of them, scaled so that the first trial step changes none by more
    than `max_rel` of its value; backtracking halves the step until the misfit
    decreases.  The other coefficients (1/rho) are held fixed.  `history`
    records the misfit of each accepted model."""

    def __init__(self, problem, max_rel=0.02, max_halvings=6, bounds=None, update=(0,)):
        self.p = problem
        self.coefs = [np.array(m, copy=True) for m in problem.m0]
        self.max_rel = max_rel
        self.max_halvings = max_halvings
        self.bounds = bounds
        self.update = tuple(update)
        self.J, g = problem.misfit(self.coefs, grad=True)
        self.g = g
        self.history = [self.J]

    def iterate(self):
        """One accepted update; returns the new misfit (unchanged when no
        trial step decreased it)."""
        gmax = max(float(np.abs(self.g[i] / self.coefs[i]).max()) for i in self.update)
        if gmax == 0:
            return self.J
        alpha = self.max_rel / gmax
        for _ in range(self.max_halvings + 1):
            cand = list(self.coefs)
            for i in self.update:
                trial = self.coefs[i] - alpha * self.g[i]
                if self.bounds is not None:
                    trial = np.clip(trial, *self.bounds)
                cand[i] = trial
            Jt, _ = self.p.misfit(cand)
            if Jt < self.J:
                self.coefs = cand
                self.J, self.g = self.p.misfit(cand, grad=True)
                self.history.append(self.J)
                return self.J
            alpha *= 0.5
        return self.J
