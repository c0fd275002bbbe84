"""Inversion driver on the elastic solver (GPU): the gradient of the shot
misfit with respect to lambda, 2 mu, mu and 1/rho (ElasticShots, through
tw_adjoint_run_coefs) against Richardson-extrapolated central differences of
the FP64 misfit, the per-receiver bookkeeping, and steepest descent on the
three moduli lowering the misfit at every accepted step."""
import numpy as np
import pytest

import twave as T
import twave.inversion as INV
import twave.workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    P = W.ElasticShots(n=96, nt=600, nshots=2, nrcv=32, precision=T.FP64, seed=7)
    yield P
    P.close()


def test_elastic_misfit_gradient_matches_fd(problem):
    print("case elastic inversion gradient", flush=True)
    P = problem
    J0, g = P.misfit(P.m0, grad=True)
    assert J0 > 1e-6 * sum(float(np.sum(d ** 2)) for d in P.data)  # the anomaly is seen
    rng = np.random.default_rng(4)
    d = [m * W.smooth_direction(rng, m.shape, 8) for m in P.m0]  # smooth relative perturbation

    def cd(h):
        jp = P.misfit([m + h * e for m, e in zip(P.m0, d)])[0]
        jm = P.misfit([m - h * e for m, e in zip(P.m0, d)])[0]
        return (jp - jm) / (2 * h)

    h = 1e-5
    fd = (4 * cd(h / 2) - cd(h)) / 3
    ad = sum(float(np.sum(a * b)) for a, b in zip(g, d))
    assert abs(fd - ad) <= 1e-8 * abs(ad), (fd, ad)  # measured 1.3e-11


def test_elastic_misfit_bookkeeping(problem):
    print("case elastic inversion bookkeeping", flush=True)
    P = problem
    shots, agg = INV.shot_misfits(P, P.m0)
    J, _ = P.misfit(P.m0)
    np.testing.assert_allclose(agg, J, rtol=1e-12)
    for i, sm in enumerate(shots):
        tr, _ = P.run(P.m0, i)
        np.testing.assert_allclose(sm.per_record, np.sum((tr - P.data[i]) ** 2, axis=1), rtol=1e-12, atol=0)


def test_elastic_steepest_descent_decreases_misfit(problem):
    print("case elastic inversion descent", flush=True)
    opt = INV.SteepestDescent(problem, max_rel=0.02, update=(0, 1, 2))
    for _ in range(2):
        opt.iterate()
    h = opt.history
    assert len(h) == 3, h
    assert all(b < a for a, b in zip(h, h[1:])), h
    assert h[-1] < 0.8 * h[0], h  # measured 0.56
    # only the moduli moved
    for i in (3, 4):
        np.testing.assert_array_equal(opt.coefs[i], problem.m0[i])


def test_model_gradient_matches_fd(problem):
    """Gradient with respect to the model of set_model (rho, Vp, Vs on the
    model nodes, the reference's inversion parameters, inverse.hpp:10) by the
    chain rule through the transposed interpolation (twave.model_gradient),
    against Richardson-extrapolated central differences of the misfit of the
    models themselves."""
    print("case elastic model gradient", flush=True)
    P = problem
    den = int(round(1.0 / P.s.cfg.dx))
    model0 = (P.rho, P.vp0, P.vs0)

    def coefs_of(model):
        P.s.set_model(*model, 1, den)
        return [P.s.get_coef(k) for k in P.coefs]

    _, g = P.misfit(coefs_of(model0), grad=True)
    gm = T.model_gradient(P.s.cfg, dict(zip(P.coefs, g)), *model0, 1, den)
    rng = np.random.default_rng(9)
    d = [m * W.smooth_direction(rng, m.shape, 8) for m in model0]

    def cd(h):
        jp = P.misfit(coefs_of([m + h * e for m, e in zip(model0, d)]))[0]
        jm = P.misfit(coefs_of([m - h * e for m, e in zip(model0, d)]))[0]
        return (jp - jm) / (2 * h)

    h = 1e-5
    fd = (4 * cd(h / 2) - cd(h)) / 3
    ad = sum(float(np.sum(a * b)) for a, b in zip(gm, d))
    assert abs(fd - ad) <= 1e-7 * abs(ad), (fd, ad)
    # each parameter on its own
    for i in range(3):
        di = [np.zeros_like(m) for m in model0]
        di[i] = d[i]
        jp = P.misfit(coefs_of([m + h * e for m, e in zip(model0, di)]))[0]
        jm = P.misfit(coefs_of([m - h * e for m, e in zip(model0, di)]))[0]
        jp2 = P.misfit(coefs_of([m + 0.5 * h * e for m, e in zip(model0, di)]))[0]
        jm2 = P.misfit(coefs_of([m - 0.5 * h * e for m, e in zip(model0, di)]))[0]
        fdi = (4 * (jp2 - jm2) / h - (jp - jm) / (2 * h)) / 3
        adi = float(np.sum(gm[i] * d[i]))
        assert abs(fdi - adi) <= 1e-6 * abs(adi), (i, fdi, adi)


def test_model_traces_autograd(problem):
    """twave.autograd.ModelTraces: the traces of one shot as a function of the
    model tensors (rho, Vp, Vs); backward equals model_gradient of the C-ABI
    gradient of the same misfit, and a second forward before the first
    backward does not change it."""
    print("case elastic model autograd", flush=True)
    import torch
    import twave.autograd as TA
    P = problem
    den = int(round(1.0 / P.s.cfg.dx))
    model0 = (P.rho, P.vp0, P.vs0)
    P.s.set_model(*model0, 1, den)
    coefs = [P.s.get_coef(k) for k in P.coefs]
    _, g = P.run(coefs, 0, P.data[0])
    want = T.model_gradient(P.s.cfg, dict(zip(P.coefs, g)), *model0, 1, den)
    P.s.restore(P.state0)
    P.s.set_source_ricker(T.GRID_NM, P.src_depth, P.shots[0], P.f0, P.t0)
    op = TA.ModelTraces(P.s, P.nt, 1, den)
    m = [torch.tensor(a, requires_grad=True) for a in model0]
    data = torch.from_numpy(P.data[0])
    loss = ((op(*m) - data) ** 2).sum()
    other = ((op(*[t.detach() * 1.01 for t in m]) - data) ** 2).sum()  # a later forward
    loss.backward()
    assert float(loss.detach()) > 0 and float(other) != float(loss.detach())
    for t, w in zip(m, want):
        np.testing.assert_allclose(t.grad.numpy(), w, rtol=1e-12, atol=1e-12 * np.abs(w).max())
