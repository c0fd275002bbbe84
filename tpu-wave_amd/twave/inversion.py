"""Inversion driver on top of the adjoint: misfit bookkeeping per shot and per
receiver, aggregated over shots as the reference's Class_Inverse_Specs does
(aggregate_data_misfit, cuda/src/cpp/inverse.hpp:34-43; per-receiver misfit
Class_Forward_Specs::calculate_data_misfit, cuda/src/cpp/forward.cpp:414-445),
and a model update loop the reference leaves unwritten: steepest descent on
the bulk modulus with a backtracking line search.

The problem object is a twave.workloads.AdjointShots (or anything with its
`run(coefs, shot, residual_of)`, `shots`, `data`, `m0`, `s` interface).
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
    """Update kappa (index 0 of the coefficients) along the negative misfit
    gradient, scaled so that the first trial step changes kappa by at most
    `max_rel` of its value; backtracking halves the step until the misfit
    decreases.  1/rho is held fixed.  `history` records the misfit of each
    accepted model."""

    def __init__(self, problem, max_rel=0.02, max_halvings=6, bounds=None):
        self.p = problem
        self.coefs = [np.array(m, copy=True) for m in problem.m0]
        self.max_rel = max_rel
        self.max_halvings = max_halvings
        self.bounds = bounds
        self.J, g = problem.misfit(self.coefs, grad=True)
        self.g = g
        self.history = [self.J]

    def iterate(self):
        """One accepted update; returns the new misfit (unchanged when no
        trial step decreased it)."""
        k = self.coefs[0]
        gk = self.g[0]
        gmax = np.abs(gk / k).max()
        if gmax == 0:
            return self.J
        alpha = self.max_rel / gmax
        for _ in range(self.max_halvings + 1):
            trial = k - alpha * gk
            if self.bounds is not None:
                trial = np.clip(trial, *self.bounds)
            cand = [trial] + self.coefs[1:]
            Jt, _ = self.p.misfit(cand)
            if Jt < self.J:
                self.coefs = cand
                self.J, self.g = self.p.misfit(cand, grad=True)
                self.history.append(self.J)
                return self.J
            alpha *= 0.5
        return self.J
