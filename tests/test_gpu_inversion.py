"""Inversion driver (GPU): the reference's per-receiver misfit bookkeeping
aggregated over shots (forward.cpp:414-445, inverse.hpp:34-43), and a
steepest-descent loop on kappa driven by the adjoint gradient whose misfit
decreases at every accepted step."""
import numpy as np
import pytest

import twave as T
import twave.inversion as INV
import twave.workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    P = W.AdjointShots(n=96, nt=400, nshots=4, nrcv=48, dx_km=0.02, f0=6.0, precision=T.FP32, seed=3)
    yield P
    P.close()


def test_misfit_bookkeeping(problem):
    print("case inversion bookkeeping", flush=True)
    P = problem
    shots, agg = INV.shot_misfits(P, P.m0)
    J, _ = P.misfit(P.m0)
    assert J > 0
    np.testing.assert_allclose(agg, J, rtol=1e-12)
    for i, sm in enumerate(shots):
        tr, _ = P.run(P.m0, i)
        ref = np.sum((tr - P.data[i]) ** 2, axis=1)
        np.testing.assert_allclose(sm.per_record, ref, rtol=1e-12, atol=0)
        np.testing.assert_allclose(sm.total, ref.sum(), rtol=1e-12)
    # the true model fits its own data
    P.s.set_model(P.rho, P.vp_true, P.vp_true, 1, int(round(1 / P.s.cfg.dx)))
    truth = [P.s.get_coef(k) for k in P.coefs]
    _, agg0 = INV.shot_misfits(P, truth)
    assert agg0 == 0.0


def test_steepest_descent_decreases_misfit(problem):
    print("case inversion descent", flush=True)
    opt = INV.SteepestDescent(problem, max_rel=0.02)
    for _ in range(3):
        opt.iterate()
    h = opt.history
    assert len(h) >= 3, h
    assert all(b < a for a, b in zip(h, h[1:])), h
    assert h[-1] < 0.9 * h[0], h
