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


def test_acoustic_model_gradient_and_autograd(problem):
    """Acoustic gradient with respect to the model of set_model (rho, Vp on the
    nodes; kappa = rho Vp^2 interpolated, 1/rho after the interpolation)
    through twave.model_gradient against Richardson-extrapolated differences
    of the FP32 shot misfit's FP64 counterpart, and twave.autograd.ModelTraces
    giving the same gradient through torch."""
    print("case acoustic model gradient", flush=True)
    import torch
    import twave.autograd as TA
    P = W.AdjointShots(n=64, nt=300, nshots=1, nrcv=32, dx_km=0.02, f0=6.0, precision=T.FP64, seed=3)
    den = int(round(1.0 / P.s.cfg.dx))
    vp0 = np.full_like(P.rho, 2.0)  # the background (vbg) without the anomaly
    model0 = (P.rho, vp0)

    def coefs_of(model):
        P.s.set_model(model[0], model[1], model[1], 1, den)
        return [P.s.get_coef(k) for k in P.coefs]

    P.s.restore(P.state0)
    for k, v in zip(P.coefs, coefs_of(model0)):
        P.s.set_coef(k, v)
    P.s.set_source_ricker(T.GRID_MM, 4, P.shots[0], P.f0, P.t0)
    P.s.adjoint_forward(P.nt)
    tr = P.s.receivers()[:, -P.nt:]
    g = P.s.adjoint_run_coefs(2.0 * (tr - P.data[0]))
    g_rho, g_vp, g_vs = T.model_gradient(P.s.cfg, g, P.rho, vp0, None, 1, den)
    assert g_vs is None
    rng = np.random.default_rng(2)
    d = [m * W.smooth_direction(rng, m.shape, 6) for m in model0]

    def J(h):
        return P.misfit(coefs_of([m + h * e for m, e in zip(model0, d)]))[0]

    h = 1e-5
    cd1, cd2 = (J(h) - J(-h)) / (2 * h), (J(h / 2) - J(-h / 2)) / h
    fd = (4 * cd2 - cd1) / 3
    ad = float(np.sum(g_rho * d[0]) + np.sum(g_vp * d[1]))
    assert abs(fd - ad) <= 1e-7 * abs(ad), (fd, ad)
    # the same through torch (vs is ignored by the acoustic solver: zero gradient)
    P.s.restore(P.state0)
    P.s.set_source_ricker(T.GRID_MM, 4, P.shots[0], P.f0, P.t0)
    op = TA.ModelTraces(P.s, P.nt, 1, den)
    m = [torch.tensor(a, requires_grad=True) for a in (P.rho, vp0, vp0)]
    loss = ((op(*m) - torch.from_numpy(P.data[0])) ** 2).sum()
    loss.backward()
    np.testing.assert_allclose(m[0].grad.numpy(), g_rho, rtol=1e-12, atol=1e-12 * np.abs(g_rho).max())
    np.testing.assert_allclose(m[1].grad.numpy(), g_vp, rtol=1e-12, atol=1e-12 * np.abs(g_vp).max())
    assert not m[2].grad.numpy().any()
    P.close()
