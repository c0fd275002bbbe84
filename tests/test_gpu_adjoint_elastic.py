"""Discrete adjoint of the elastic step (GPU): gradients of receiver misfits
with respect to lambda, 2 mu (MM), mu (NN) and 1/rho (NM, MN), checked against
central finite differences of the forward solver across boundary types
(periodic, free surface, characteristic: the characteristic sides also carry
the impedance sensitivities of the normal and tangential velocities),
checkpoint intervals and precisions, and through the torch.autograd wrappers.

The forward solver is itself pinned to the oracle bitwise (test_gpu_parity),
so finite differences of it are the reference for the transposed step.
"""
import numpy as np
import pytest

import common as C
import twave as T

pytestmark = pytest.mark.gpu

COEFS = (T.C_LAMBDA, T.C_TWOMU, T.C_MU, T.C_BINV_VX, T.C_BINV_VY)


def elastic_solver(bc, precision="fp64", mx=20, my=24, nt=40, seed=1, receivers_per_field=4):
    kw = C.cfg_kwargs(physics="elastic", mx=mx, my=my, bc=bc, precision=precision, cpst="0", nt=4 * nt)
    s = T.Solver(T.config(**kw))
    rng = np.random.default_rng(seed)
    # vp = sqrt((lambda + 2 mu) / rho) <= sqrt(3.6 * 1.2) ~ 2.08: CFL about 0.52 at the default dt
    for k in COEFS:
        lo, hi = {T.C_TWOMU: (1.6, 2.4)}.get(k, (0.8, 1.2))
        s.set_coef(k, rng.uniform(lo, hi, s.get_coef(k).shape))
    for f in s.fields:
        s.set_field(f, 1e-3 * rng.standard_normal(s.shape(f)))
    s.set_source_ricker(T.GRID_MM, mx // 2, my // 3, 25.0, 0.01)
    fs, xs, ys = [], [], []
    for f in s.fields:
        r, c = s.shape(f)
        for _ in range(receivers_per_field):
            fs.append(f)
            xs.append(int(rng.integers(0, r)))
            ys.append(int(rng.integers(0, c)))
    s.set_receivers(fs, xs, ys)
    return s


def traces(s, blob, coefs, nt):
    s.restore(blob)
    for k, v in zip(COEFS, coefs):
        s.set_coef(k, v)
    s.step(nt)
    return s.receivers()[:, -nt:]


def central(s, blob, m0, d, w, nt, eps):
    lp = np.sum(w * traces(s, blob, [m + eps * dd for m, dd in zip(m0, d)], nt))
    lm = np.sum(w * traces(s, blob, [m - eps * dd for m, dd in zip(m0, d)], nt))
    return (lp - lm) / (2 * eps)


def richardson(s, blob, m0, d, w, nt, eps):
    """Central differences at eps and eps/2 combined to cancel the eps^2 term
    (the traces are polynomial in the coefficients, so plain differences at
    1e-4 carry a truncation error of up to ~1e-5 relative)."""
    return (4 * central(s, blob, m0, d, w, nt, eps / 2) - central(s, blob, m0, d, w, nt, eps)) / 3


def fd_check(s, nt, ndir=3, eps=1e-4, seed=5, smooth=False, extrapolate=True):
    """L = sum(w * traces): adjoint gradient vs (Richardson-extrapolated)
    central differences on random directions; returns the worst relative
    error and the gradients."""
    rng = np.random.default_rng(seed)
    blob = s.checkpoint()
    m0 = [s.get_coef(k) for k in COEFS]
    w = rng.standard_normal((s.nrcv, nt))
    s.adjoint_forward(nt)
    g = s.adjoint_run_coefs(w)
    worst = 0.0
    for _ in range(ndir):
        if smooth:  # low-wavenumber relative perturbations
            d = []
            for m in m0:
                r, c = m.shape
                kx, ky = rng.uniform(0.5, 2.0, 2)
                ph = rng.uniform(0, 2 * np.pi)
                x = np.arange(r)[:, None] / r
                y = np.arange(c)[None, :] / c
                d.append(m * np.sin(2 * np.pi * (kx * x + ky * y) + ph))
        else:
            d = [rng.standard_normal(m.shape) for m in m0]
        fd = (richardson if extrapolate else central)(s, blob, m0, d, w, nt, eps)
        ad = sum(np.sum(g[k] * dd) for k, dd in zip(COEFS, d))
        worst = max(worst, abs(fd - ad) / max(abs(fd), 1e-300))
    return worst, g


@pytest.mark.parametrize("bc", ["PPPP", "FFFF", "CCCC", "FFPP", "FCPP", "PPFC", "FCCC", "CFFC"])
def test_elastic_adjoint_matches_finite_differences_fp64(bc):
    print("case elastic adjoint fd", bc, flush=True)
    s = elastic_solver(bc)
    worst, _ = fd_check(s, nt=40)
    assert worst < 1e-9, worst


def test_elastic_adjoint_each_coefficient():
    """Extrapolated central differences along one coefficient array at a
    time, so a gradient that is wrong for one array alone cannot hide behind
    the others."""
    s = elastic_solver("FCCC", nt=30, seed=3)
    rng = np.random.default_rng(11)
    blob = s.checkpoint()
    m0 = [s.get_coef(k) for k in COEFS]
    w = rng.standard_normal((s.nrcv, 30))
    s.adjoint_forward(30)
    g = s.adjoint_run_coefs(w)
    eps = 1e-4
    for i, k in enumerate(COEFS):
        d = rng.standard_normal(m0[i].shape)
        dirs = [d if j == i else np.zeros_like(m) for j, m in enumerate(m0)]
        fd = richardson(s, blob, m0, dirs, w, 30, eps)
        ad = np.sum(g[k] * d)
        assert abs(fd - ad) <= 1e-9 * abs(fd), (k, fd, ad)


def test_elastic_adjoint_checkpoint_interval_invariant():
    s = elastic_solver("FCCC", nt=30)
    blob = s.checkpoint()
    w = np.random.default_rng(2).standard_normal((s.nrcv, 30))
    out = []
    for interval in (1, 4, 7, 30):
        s.restore(blob)
        s.adjoint_forward(30, interval)
        out.append(s.adjoint_run_coefs(w))
    for g in out[1:]:
        for k in COEFS:
            np.testing.assert_array_equal(g[k], out[0][k])


def test_elastic_adjoint_fp32_close_to_fp64():
    g = {}
    for prec in ("fp64", "fp32"):
        s = elastic_solver("FCCC", precision=prec, nt=40)
        w = np.random.default_rng(3).standard_normal((s.nrcv, 40))
        s.adjoint_forward(40)
        g[prec] = s.adjoint_run_coefs(w)
        s.close()
    for k in COEFS:
        assert C.rel_l2(g["fp32"][k], g["fp64"][k]) < 1e-4, k


def test_elastic_adjoint_smooth_directions_eps_1e3():
    """A scaled inversion-like setting (free-surface top, characteristic
    sides and bottom, a shot near the top, receivers on every field): central
    differences at eps = 1e-3 along smooth relative perturbations agree with
    the adjoint within 1e-3."""
    s = elastic_solver("FCCC", mx=32, my=40, nt=60, seed=9)
    worst, _ = fd_check(s, nt=60, ndir=4, eps=1e-3, smooth=True, extrapolate=False)
    assert worst < 1e-3, worst


def test_elastic_adjoint_rejects_unsupported():
    kw = C.cfg_kwargs(physics="elastic", mx=16, my=16, bc="FCCC", precision="bf16", cpst="K8", nt=8)
    s = T.Solver(T.config(**kw))
    with pytest.raises(T.TwError):
        s.adjoint_forward(2)
    s2 = elastic_solver("FCCC", nt=4)
    with pytest.raises(T.TwError):
        s2.adjoint_run_coefs(np.zeros((s2.nrcv, 2)))  # no checkpointed run yet
    s2.adjoint_forward(4)
    with pytest.raises(T.TwError):
        s2.adjoint_run(np.zeros((s2.nrcv, 4)))  # the acoustic-only entry point


def test_elastic_autograd_wrappers():
    torch = pytest.importorskip("torch")
    import twave.autograd as TA
    s = elastic_solver("FCCC", nt=25)
    op = TA.ElasticTraces(s, 25)
    m = [torch.tensor(s.get_coef(k), requires_grad=True) for k in COEFS]
    w = torch.tensor(np.random.default_rng(4).standard_normal((s.nrcv, 25)))
    tr = op(*m)
    (tr * w).sum().backward()
    s.restore(op.state0)
    s.adjoint_forward(25)
    ref = s.adjoint_run_coefs(w.numpy())
    for t, k in zip(m, COEFS):
        np.testing.assert_array_equal(t.grad.numpy(), ref[k])
    np.testing.assert_array_equal(tr.detach().numpy(),
                                  traces(s, op.state0, [t.detach().numpy() for t in m], 25))
    # density and Lame moduli: torch's chain rule, checked by central differences
    op2 = TA.ElasticTracesLameRho(s, 25)
    lam = torch.tensor(s.get_coef(T.C_LAMBDA), requires_grad=True)
    mu_mm = torch.tensor(s.get_coef(T.C_TWOMU) / 2, requires_grad=True)
    mu_nn = torch.tensor(s.get_coef(T.C_MU), requires_grad=True)
    rho = [torch.tensor(1.0 / s.get_coef(k), requires_grad=True) for k in (T.C_BINV_VX, T.C_BINV_VY)]
    args = [lam, mu_mm, mu_nn] + rho
    (op2(*args) * w).sum().backward()
    rng = np.random.default_rng(8)
    d = [rng.standard_normal(t.shape) for t in args]
    eps = 1e-6

    def loss(sign):
        a = [torch.tensor(t.detach().numpy() + sign * eps * dd) for t, dd in zip(args, d)]
        with torch.no_grad():
            return float((op2(*a) * w).sum())

    fd = (loss(1) - loss(-1)) / (2 * eps)
    ad = sum(float((t.grad * torch.tensor(dd)).sum()) for t, dd in zip(args, d))
    assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)


def test_elastic_adjoint_benchmark_model_512():
    """The benchmark's elastic workload at 512^2 (layered model, 15 Hz Ricker
    vertical force at 100 m depth, free-surface top, characteristic sides,
    256 surface receivers on vx), 1000 steps: the misfit
    sum (trace - data)^2 against data from a model with 2% stiffer moduli in a
    disc around the source, differentiated by the adjoint (FP64) and by extrapolated central
    differences along smooth relative perturbations of all five coefficient
    arrays; FP32 storage's gradient against FP64's."""
    import twave.workloads as W
    m, nt = 512, 1000
    grads = {}
    for prec in (T.FP64, T.FP32):
        s = T.Solver(W.elastic_config(m=m, precision=prec, nt=nt + 8, dx_km=0.005))
        W.setup_elastic(s, m, 0.005)
        blob = s.checkpoint()
        m0 = [s.get_coef(k) for k in COEFS]
        x = np.arange(s.get_coef(T.C_LAMBDA).shape[0])[:, None]
        y = np.arange(s.get_coef(T.C_LAMBDA).shape[1])[None, :]
        disc = ((x - 60) ** 2 + (y - m // 2) ** 2 < 30 ** 2) * 0.02 + 1.0  # 300 m deep, radius 150 m
        true = [mm * (disc if k == T.C_LAMBDA or k == T.C_TWOMU else 1.0) for k, mm in zip(COEFS, m0)]
        data = traces(s, blob, true, nt)
        s.restore(blob)
        for k, v in zip(COEFS, m0):
            s.set_coef(k, v)
        s.adjoint_forward(nt)
        tr = s.receivers()[:, -nt:]
        g = s.adjoint_run_coefs(2 * (tr - data))
        grads[prec] = g
        misfit = np.sum((tr - data) ** 2)
        print("precision", prec, "misfit %.6e, data energy %.6e" % (misfit, np.sum(data ** 2)), flush=True)
        assert misfit > 1e-6 * np.sum(data ** 2)  # the perturbation reaches the receivers
        if prec == T.FP64:
            rng = np.random.default_rng(21)
            worst = 0.0
            for _ in range(3):
                d = []
                for mm in m0:
                    r, c = mm.shape
                    kx, ky = rng.uniform(0.5, 3.0, 2)
                    ph = rng.uniform(0, 2 * np.pi)
                    d.append(mm * np.sin(2 * np.pi * (kx * np.arange(r)[:, None] / r + ky * np.arange(c)[None, :] / c) + ph))

                def J(eps):
                    return np.sum((traces(s, blob, [mm + eps * dd for mm, dd in zip(m0, d)], nt) - data) ** 2)

                e = 1e-4
                c1 = (J(e) - J(-e)) / (2 * e)
                c2 = (J(e / 2) - J(-e / 2)) / e
                fd = (4 * c2 - c1) / 3
                ad = sum(np.sum(g[k] * dd) for k, dd in zip(COEFS, d))
                worst = max(worst, abs(fd - ad) / abs(fd))
            print("worst rel err adjoint vs extrapolated differences %.3e" % worst, flush=True)
            assert worst < 1e-6, worst
        s.close()
    for k in COEFS:
        err = C.rel_l2(grads[T.FP32][k], grads[T.FP64][k])
        print("fp32 vs fp64 gradient rel-L2", k, "%.3e" % err, flush=True)
        assert err < 1e-3, (k, err)


def test_adjoint_receivers_sharing_a_point():
    """Receivers at one grid point (listed apart, on different fields too):
    their adjoint sources are added point by point in list order, so the
    gradient equals that of one receiver per point carrying the summed
    weights (to rounding), and repeated runs are bitwise equal."""
    s = elastic_solver("FCCC", nt=20, seed=4, receivers_per_field=1)
    blob = s.checkpoint()
    f, x, y = s.get_receivers()
    f, x, y = list(f), list(x), list(y)
    # duplicate the first three receivers, interleaved with the others
    f2, x2, y2 = f[:3] + f + f[:3], x[:3] + x + x[:3], y[:3] + y + y[:3]
    rng = np.random.default_rng(13)
    w2 = rng.standard_normal((len(f2), 20))
    s.set_receivers(f2, x2, y2)
    s.restore(blob)
    s.adjoint_forward(20)
    g2 = s.adjoint_run_coefs(w2)
    g2b = s.adjoint_run_coefs(w2)
    w1 = w2[3:3 + len(f)].copy()
    w1[:3] += w2[:3] + w2[3 + len(f):]
    s.set_receivers(f, x, y)
    s.restore(blob)
    s.adjoint_forward(20)
    g1 = s.adjoint_run_coefs(w1)
    for k in COEFS:
        np.testing.assert_array_equal(g2[k], g2b[k])
        assert C.rel_l2(g2[k], g1[k]) < 1e-12, k
