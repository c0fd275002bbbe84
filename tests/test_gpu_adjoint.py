"""Discrete adjoint of the acoustic step (GPU): gradients of receiver misfits
with respect to kappa and 1/rho, checked against central finite differences
of the forward solver, across boundary types, checkpoint intervals and
precisions, and through the torch.autograd wrapper.

The forward solver is itself pinned to the oracle bitwise (test_gpu_parity),
so finite differences of it are the reference for the transposed step.
"""
import os

import numpy as np
import pytest

import common as C
import twave as T

pytestmark = pytest.mark.gpu

COEFS = (T.C_KAPPA, T.C_BINV_VX, T.C_BINV_VY)


def acoustic_solver(bc, precision="fp64", mx=20, my=24, nt=40, seed=1, dx=1.0 / 125):
    kw = C.cfg_kwargs(physics="acoustic", mx=mx, my=my, bc=bc, precision=precision, cpst="0", nt=4 * nt,
                      dx=dx, stencil_form=1)
    s = T.Solver(T.config(**kw))
    rng = np.random.default_rng(seed)
    for k in COEFS:
        s.set_coef(k, rng.uniform(0.8, 1.2, s.get_coef(k).shape))
    for f in s.fields:
        s.set_field(f, 1e-3 * rng.standard_normal(s.shape(f)))
    s.set_source_ricker(T.GRID_MM, mx // 2, my // 3, 25.0, 0.01)
    fs, xs, ys = [], [], []
    for f in s.fields:
        r, c = s.shape(f)
        for _ in range(4):
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


def linear_functional_check(s, nt, ndir=3, eps=1e-4, seed=5):
    """L = sum(w * traces): adjoint gradient vs central differences on random
    directions; returns the worst relative error."""
    rng = np.random.default_rng(seed)
    blob = s.checkpoint()
    m0 = [s.get_coef(k) for k in COEFS]
    w = rng.standard_normal((s.nrcv, nt))
    s.adjoint_forward(nt)
    g = s.adjoint_run(w)
    worst = 0.0
    for _ in range(ndir):
        d = [rng.standard_normal(m.shape) for m in m0]
        lp = np.sum(w * traces(s, blob, [m + eps * dd for m, dd in zip(m0, d)], nt))
        lm = np.sum(w * traces(s, blob, [m - eps * dd for m, dd in zip(m0, d)], nt))
        fd = (lp - lm) / (2 * eps)
        ad = sum(np.sum(gg * dd) for gg, dd in zip(g, d))
        worst = max(worst, abs(fd - ad) / max(abs(fd), 1e-300))
    return worst, g


@pytest.mark.parametrize("bc", ["PPPP", "FFFF", "CCCC", "FCPP", "PPCF", "CFFC"])
def test_adjoint_matches_finite_differences_fp64(bc):
    print("case adjoint fd", bc, flush=True)
    s = acoustic_solver(bc)
    worst, _ = linear_functional_check(s, nt=40)
    assert worst < 1e-6, worst


def test_adjoint_checkpoint_interval_invariant():
    s = acoustic_solver("FCCC", nt=30)
    blob = s.checkpoint()
    w = np.random.default_rng(2).standard_normal((s.nrcv, 30))
    out = []
    for interval in (1, 4, 7, 30):
        s.restore(blob)
        s.adjoint_forward(30, interval)
        rec, en = s.receivers().copy(), [e.copy() for e in s.energy()]
        out.append(s.adjoint_run(w))
        # the sweep's recomputed steps record nothing: the forward run's
        # receiver values and energies are what the solver still holds
        np.testing.assert_array_equal(s.receivers(), rec)
        for a, b in zip(s.energy(), en):
            np.testing.assert_array_equal(a, b)
    for g in out[1:]:
        for a, b in zip(g, out[0]):
            np.testing.assert_array_equal(a, b)


def test_adjoint_fp32_close_to_fp64():
    g = {}
    for prec in ("fp64", "fp32"):
        s = acoustic_solver("FCCC", precision=prec, nt=40)
        w = np.random.default_rng(3).standard_normal((s.nrcv, 40))
        s.adjoint_forward(40)
        g[prec] = s.adjoint_run(w)
        s.close()
    for a, b in zip(g["fp32"], g["fp64"]):
        assert C.rel_l2(a, b) < 1e-4


def test_adjoint_rejects_unsupported():
    s, _ = C.make_pair(physics="acoustic", mx=16, my=16, bc="FFPP", precision="bf16", nt=4)
    with pytest.raises(T.TwError):
        s.adjoint_forward(2)
    s1, _ = C.make_pair(physics="acoustic", mx=16, my=16, bc="FFPP", precision="fp32", cpst="6R", nt=4)
    with pytest.raises(T.TwError):
        s1.adjoint_forward(2)
    s2 = acoustic_solver("PPPP")
    with pytest.raises(T.TwError):
        s2.adjoint_run(np.zeros((s2.nrcv, 2)))  # no checkpointed run yet


def test_autograd_wrapper():
    torch = pytest.importorskip("torch")
    import twave.autograd as TA
    s = acoustic_solver("FCCC", nt=25)
    op = TA.AcousticTraces(s, 25)
    m = [torch.tensor(s.get_coef(k), requires_grad=True) for k in COEFS]
    w = torch.tensor(np.random.default_rng(4).standard_normal((s.nrcv, 25)))
    tr = op(*m)
    (tr * w).sum().backward()
    s.restore(op.state0)
    s.adjoint_forward(25)
    ref = s.adjoint_run(w.numpy())
    for t, r in zip(m, ref):
        np.testing.assert_array_equal(t.grad.numpy(), r)
    # forward values equal a plain run from the same state
    np.testing.assert_array_equal(tr.detach().numpy(),
                                  traces(s, op.state0, [t.detach().numpy() for t in m], 25))


def test_autograd_two_forwards_then_backwards():
    """Two forwards (different models) before their backwards: each gradient
    is that of its own trajectory, equal to a direct adjoint run there."""
    torch = pytest.importorskip("torch")
    import twave.autograd as TA
    s = acoustic_solver("FCCC", nt=25)
    op = TA.AcousticTraces(s, 25)
    base = [s.get_coef(k) for k in COEFS]
    ma = [torch.tensor(b, requires_grad=True) for b in base]
    mb = [torch.tensor(b * f, requires_grad=True) for b, f in zip(base, (1.05, 0.97, 1.02))]
    rng = np.random.default_rng(5)
    wa = torch.tensor(rng.standard_normal((s.nrcv, 25)))
    wb = torch.tensor(rng.standard_normal((s.nrcv, 25)))
    la = (op(*ma) * wa).sum()
    lb = (op(*mb) * wb).sum()
    (la + lb).backward()
    for m, w in ((ma, wa), (mb, wb)):
        s.restore(op.state0)
        for k, t in zip(COEFS, m):
            s.set_coef(k, t.detach().numpy())
        s.adjoint_forward(25)
        ref = s.adjoint_run(w.numpy())
        for t, r in zip(m, ref):
            np.testing.assert_array_equal(t.grad.numpy(), r)


def test_autograd_density_parametrization():
    """Gradient with respect to rho on the velocity grids equals the chain
    rule -(1/rho^2) d/d(1/rho) of the direct adjoint, and matches central
    differences of the traces in rho."""
    torch = pytest.importorskip("torch")
    import twave.autograd as TA
    s = acoustic_solver("FCCC", nt=25)
    op = TA.AcousticTracesKappaRho(s, 25)
    kap = torch.tensor(s.get_coef(T.C_KAPPA), requires_grad=True)
    rho = [torch.tensor(1.0 / s.get_coef(k), requires_grad=True) for k in (T.C_BINV_VX, T.C_BINV_VY)]
    w = torch.tensor(np.random.default_rng(6).standard_normal((s.nrcv, 25)))
    (op(kap, *rho) * w).sum().backward()
    s.restore(op.state0)
    for k, t in zip(COEFS, (kap, 1.0 / rho[0], 1.0 / rho[1])):
        s.set_coef(k, t.detach().numpy())
    s.adjoint_forward(25)
    gk, gx, gy = s.adjoint_run(w.numpy())
    np.testing.assert_array_equal(kap.grad.numpy(), gk)
    for t, g in zip(rho, (gx, gy)):
        r = t.detach().numpy()
        np.testing.assert_allclose(t.grad.numpy(), -g / r ** 2, rtol=1e-14, atol=0)
    d = [np.random.default_rng(7 + i).standard_normal(t.shape) for i, t in enumerate(rho)]
    eps = 1e-6

    def loss(sign):
        r = [torch.tensor(t.detach().numpy() + sign * eps * dd) for t, dd in zip(rho, d)]
        with torch.no_grad():
            return float((op(kap.detach(), *r) * w).sum())

    fd = (loss(1) - loss(-1)) / (2 * eps)
    ad = sum(float((t.grad * torch.tensor(dd)).sum()) for t, dd in zip(rho, d))
    assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)


class C5Scaled:
    """Scaled-down adjoint configuration: true model = unit background
    velocity plus a circular 10% anomaly, starting model = a 3% slower
    background without it; free-surface top and characteristic sides, Ricker
    shots along the top, receivers on P along row 4, misfit
    sum (trace - data)^2 (cuda/src/cpp/forward.cpp:437-438) with data from the
    true model."""

    def __init__(self, precision, n=96, nt=400, nshots=4):
        dx = 1.0 / n
        kw = C.cfg_kwargs(physics="acoustic", mx=n, my=n, bc="FCCC", precision=precision, cpst="0", nt=2 * nt,
                          dx=dx, num_dx=1, den_dx=n, dt=0.4 * dx / (np.sqrt(2.0) * 1.1), stencil_form=1)
        self.s = T.Solver(T.config(**kw))
        self.n, self.nt = n, nt
        X, Y = np.meshgrid(np.arange(n + 1) * dx, np.arange(n + 1) * dx, indexing="ij")
        rho = np.ones((n + 1, n + 1))
        vp_true = 1.0 + 0.1 * (((X - 0.35) ** 2 + (Y - 0.5) ** 2) < 0.12 ** 2)
        ry = list(range(4, n - 4, 3))
        self.s.set_receivers([T.P] * len(ry), [4] * len(ry), ry)
        self.shots = [int(v) for v in np.linspace(n // 8, n - n // 8, nshots)]
        self.state0 = self.s.checkpoint()
        self.s.set_model(rho, vp_true, vp_true, 1, n)
        truth = [self.s.get_coef(k) for k in COEFS]
        self.data = [self.run(truth, i)[0] for i in range(nshots)]
        self.s.set_model(rho, 0.97 * np.ones_like(rho), 0.97 * np.ones_like(rho), 1, n)
        self.m0 = [self.s.get_coef(k) for k in COEFS]

    def run(self, coefs, shot, grad_of=None):
        s = self.s
        s.restore(self.state0)
        for k, v in zip(COEFS, coefs):
            s.set_coef(k, v)
        s.set_source_ricker(T.GRID_MM, 6, self.shots[shot], 12.0, 0.1)
        if grad_of is None:
            s.step(self.nt)
            return s.receivers()[:, -self.nt:], None
        s.adjoint_forward(self.nt)
        tr = s.receivers()[:, -self.nt:]
        return tr, s.adjoint_run(2 * (tr - grad_of))

    def misfit(self, coefs, grad=False):
        total, g = 0.0, [np.zeros_like(m) for m in self.m0]
        for i in range(len(self.shots)):
            tr, gi = self.run(coefs, i, self.data[i] if grad else None)
            total += float(np.sum((tr - self.data[i]) ** 2))
            if grad:
                for a, b in zip(g, gi):
                    a += b
        return total, g


def smooth_direction(rng, shape, corr):
    """Unit-peak random model perturbation, Gaussian-correlated over `corr`
    cells (model-space directions; white noise would scatter so strongly that
    eps = 1e-3 leaves the linear regime)."""
    from scipy.ndimage import gaussian_filter
    d = gaussian_filter(rng.standard_normal(shape), corr, mode="nearest")
    return d / np.abs(d).max()


def test_adjoint_c5_scaled_fd_directions():
    """Misfit gradient on 32 random smooth directions against central
    differences (eps = 1e-3 of the coefficients), FP64; the FP32 adjoint
    agrees with the FP64 one to 1e-3."""
    print("case adjoint c5 scaled", flush=True)
    P = C5Scaled("fp64")
    J0, g = P.misfit(P.m0, grad=True)
    assert J0 > 0
    rng = np.random.default_rng(11)
    eps = 1e-3
    fds, ads = [], []
    for _ in range(32):
        d = [smooth_direction(rng, m.shape, P.n / 16) * m for m in P.m0]
        jp, _ = P.misfit([m + eps * dd for m, dd in zip(P.m0, d)])
        jm, _ = P.misfit([m - eps * dd for m, dd in zip(P.m0, d)])
        fds.append((jp - jm) / (2 * eps))
        ads.append(sum(float(np.sum(a * dd)) for a, dd in zip(g, d)))
    fds, ads = np.array(fds), np.array(ads)
    # relative error, with directions of nearly zero slope measured against
    # the typical slope
    scale = np.maximum(np.abs(fds), 0.1 * np.median(np.abs(fds)))
    err = np.abs(fds - ads) / scale
    print("max rel err", err.max(), flush=True)
    assert err.max() < 1e-3, err
    P32 = C5Scaled("fp32")
    _, g32 = P32.misfit(P32.m0, grad=True)
    for a, b in zip(g32, g):
        assert C.rel_l2(a, b) < 1e-3


def test_adjoint_shots_workload_small():
    """The adjoint-gradient workload class at a small size: gradient against
    central differences on smooth directions (FP64)."""
    import twave.workloads as W
    print("case adjoint shots small", flush=True)
    P = W.AdjointShots(n=64, nt=240, nshots=2, nrcv=32, dx_km=0.04, f0=3.0, precision=T.FP64, start_scale=0.98)
    J0, g = P.misfit(P.m0, grad=True)
    assert J0 > 0
    rng = np.random.default_rng(21)
    errs = []
    for _ in range(4):
        d = [W.smooth_direction(rng, m.shape, 4) * m for m in P.m0]
        jp, _ = P.misfit([m + 1e-3 * dd for m, dd in zip(P.m0, d)])
        jm, _ = P.misfit([m - 1e-3 * dd for m, dd in zip(P.m0, d)])
        fd = (jp - jm) / 2e-3
        ad = sum(float(np.sum(a * dd)) for a, dd in zip(g, d))
        errs.append(abs(fd - ad) / abs(fd))
    P.close()
    assert max(errs) < 1e-3, errs


def test_adjoint_c5_full_scale(progress):
    """The adjoint configuration at full size: acoustic 2048^2, 16 Ricker
    12 Hz shots along the top, 512 receivers, 3000 steps, true model = 2 km/s
    background + seeded circular 10% anomaly, starting model = background.

    FP64 adjoint gradient against finite differences of the FP64 forward
    misfit on smooth random directions (4 by default, 32 with TW_RUN_SLOW=1,
    or TW_C5_DIRS; direction seed TW_C5_SEED), and the FP32 gradient (sqrt(N)
    checkpoints) against the FP64 one.  With the receivers beside the sources, the first-order change
    of the misfit is tiny next to its curvature along the direct wave:
    central differences at eps = 1e-3 are dominated by their O(eps^2) term
    (FD / adjoint = 46 at 1e-3, 1.45 at 1e-4, 1.0045 at 1e-5, 1.0006 at 1e-6
    on the first direction), so the differences are Richardson-extrapolated
    from h = 1e-5 and 2h (error O(h^4)).  For the same reason the directional
    derivatives are ~1e-7 of |g| |d|, below FP32 resolution of the gradient,
    so FP32 is measured by its distance to the FP64 gradient."""
    import time
    import twave.workloads as W
    ndir = int(os.environ.get("TW_C5_DIRS", "32" if os.environ.get("TW_RUN_SLOW") else "4"))
    seed = int(os.environ.get("TW_C5_SEED", "11"))
    h = 1e-5
    log = progress
    t = time.perf_counter()
    P64 = W.AdjointShots(precision=T.FP64, log=log)
    log("fp64 data %.1f s" % (time.perf_counter() - t))
    t = time.perf_counter()
    J0, g64 = P64.misfit(P64.m0, grad=True)
    log("fp64 misfit %.6e, gradient %.1f s" % (J0, time.perf_counter() - t))
    P32 = W.AdjointShots(precision=T.FP32, data=P64.data)
    t = time.perf_counter()
    J32, g32 = P32.misfit(P32.m0, grad=True)
    t32 = time.perf_counter() - t
    P32.close()
    rl = [C.rel_l2(a, b) for a, b in zip(g32, g64)]
    log("fp32 misfit %.6e, gradient %.1f s; rel-L2 vs fp64 (kappa, 1/rho NM, 1/rho MN) %s" % (J32, t32, rl))
    rng = np.random.default_rng(seed)

    # the 16 shots as one batch (one launch per half step for all of them,
    # receivers recorded): traces equal to the shot-by-shot runs bit for bit
    t = time.perf_counter()
    Jb = P64.misfit_batched(P64.m0)
    log("fp64 misfit, shots batched: %.6e (%.1f s)" % (Jb, time.perf_counter() - t))
    assert Jb == J0

    def cd(d, e):
        jp = P64.misfit([m + e * dd for m, dd in zip(P64.m0, d)])[0]
        jm = P64.misfit([m - e * dd for m, dd in zip(P64.m0, d)])[0]
        return (jp - jm) / (2 * e)

    fds, a64 = [], []
    for i in range(ndir):
        d = [W.smooth_direction(rng, m.shape, P64.n / 16) * m for m in P64.m0]
        d1, d2 = cd(d, h), cd(d, 2 * h)
        fds.append((4 * d1 - d2) / 3)
        a64.append(sum(float(np.sum(a * dd)) for a, dd in zip(g64, d)))
        log("direction %d: cd(h) %.9e cd(2h) %.9e richardson %.9e adjoint64 %.9e" % (i, d1, d2, fds[-1], a64[-1]))
    P64.close()
    fds, a64 = np.array(fds), np.array(a64)
    scale = np.maximum(np.abs(fds), 0.1 * np.median(np.abs(fds)))
    e64 = np.abs(fds - a64) / scale
    log("max rel err fp64 adjoint vs finite differences %.3e (median %.3e)" % (e64.max(), np.median(e64)))
    assert e64.max() < 1e-3
    assert max(rl) < 1e-3
