"""Synthetic workloads of the benchmark configurations (seeded, regenerable).

Units are km, s, g/cm^3 and GPa, a consistent system for the velocity-stress
equations (1 GPa / (g/cm^3 * km) = 1 km/s^2), so velocities and stresses are
within an impedance factor (~2-10) of each other and 16-bit storage stays in
range.  Physical 'Y' (depth) is data x (rows), physical 'X' is data y.
"""
import math

import numpy as np

from . import (BF16, C_BINV_VX, C_BINV_VY, C_KAPPA, ELASTIC, ACOUSTIC, FP16, FP32, FP64, GRID_MM, GRID_NM, VX,
               config)


def gaussian_noise(shape, corr_cells, rng):
    """Unit-variance Gaussian-correlated noise (Gaussian filter of white noise)."""
    from scipy.ndimage import gaussian_filter
    w = rng.standard_normal(shape).astype(np.float32)
    g = gaussian_filter(w, sigma=corr_cells, mode="wrap").astype(np.float64)
    s = g.std()
    return g / (s if s > 0 else 1.0)


_cache = {}


def layered_elastic_model(m, dx_km, seed=2024, nlayers=20, vp_range=(1.5, 4.0), pert=0.05, corr_km=0.05):
    """20-layer model on the (m+1)^2 model nodes: Vp uniform per layer in
    [1.5, 4.0] km/s times (1 + 5% Gaussian-correlated noise, length 50 m),
    Vs = Vp/1.73, rho = 0.31 Vp^0.25 (g/cm^3; 310 Vp^0.25 in kg/m^3, m/s).
    The last model built is cached (read-only arrays)."""
    key = (m, dx_km, seed, nlayers, vp_range, pert, corr_km)
    if key in _cache:
        return _cache[key]
    _cache.clear()
    _cache[key] = _layered(m, dx_km, seed, nlayers, vp_range, pert, corr_km)
    return _cache[key]


def _layered(m, dx_km, seed, nlayers, vp_range, pert, corr_km):
    rng = np.random.default_rng(seed)
    n = m + 1
    bounds = np.sort(rng.choice(np.arange(1, n), size=nlayers - 1, replace=False))
    layer = np.searchsorted(bounds, np.arange(n), side="right")
    vl = rng.uniform(vp_range[0], vp_range[1], nlayers)
    vp = vl[layer][:, None] * (1.0 + pert * gaussian_noise((n, n), corr_km / dx_km, rng))
    vs = vp / 1.73
    rho = 0.31 * (vp * 1000.0) ** 0.25
    return rho, vp, vs


def smooth_acoustic_model(m, dx_km, seed, vrange=(1.5, 4.5), corr_km=0.1):
    """Smooth random velocity (Gaussian-filtered noise, length 100 m) mapped to
    [1.5, 4.5] km/s, constant density 1 g/cm^3."""
    rng = np.random.default_rng(seed)
    g = gaussian_noise((m + 1, m + 1), corr_km / dx_km, rng)
    lo, hi = g.min(), g.max()
    vp = vrange[0] + (vrange[1] - vrange[0]) * (g - lo) / (hi - lo if hi > lo else 1.0)
    rho = np.ones_like(vp)
    return rho, vp, vp / 1.73


def elastic_config(m=4096, precision=FP32, cpst="0", coef16=0, nt=2000, dx_km=0.005, cfl=0.5, vmax=None,
                   bc="FCCC", source_scaling=1.0):
    """The elastic benchmark case: free-surface top (data x left), characteristic
    elsewhere.  dt from the reference's CFL form cfl = (dt/dx) sqrt(2) vmax
    (cuda/src/cpp/namespace_input.hpp:116)."""
    if vmax is None:
        vmax = 4.0 * 1.25
    dt = cfl * dx_km / (math.sqrt(2.0) * vmax)
    # dx = 5 m = 1/200 km as a rational for the location maps
    num, den = 1, int(round(1.0 / dx_km))
    return config(physics=ELASTIC, mx=m, my=m, dx=dx_km, num_dx=num, den_dx=den, dt=dt, bc=bc,
                  precision=precision, cpst=cpst, coef16=coef16, nt_max=nt, source_scaling=source_scaling)


def setup_elastic(solver, m, dx_km, seed=2024, f0=15.0, depth_km=0.1, nrcv=256):
    """Material, 15 Hz Ricker vertical force at 100 m depth (vx, data (N,M)),
    and nrcv receivers of vx along the free surface."""
    rho, vp, vs = layered_elastic_model(m, dx_km, seed)
    den = int(round(1.0 / dx_km))
    solver.set_model(rho, vp, vs, 1, den)
    ix = int(round(depth_km / dx_km))
    iy = m // 2
    solver.set_source_ricker(GRID_NM, ix, iy, f0, 1.25 / f0)
    cols = np.linspace(m // 16, m - m // 16, nrcv).astype(np.int32)
    solver.set_receivers(np.full(nrcv, VX, np.int32), np.zeros(nrcv, np.int32), cols)
    return dict(vpmax=float(vp.max()), src=(ix, iy))


class AdjointShots:
    """The adjoint-gradient configuration: acoustic n x n cells, a uniform
    background velocity plus a seeded circular 10% anomaly (the true model),
    Ricker shots on p along row 4 under the free surface (top 'F',
    characteristic elsewhere), receivers on p along the same row, misfit
    J = sum over shots and samples of (trace - data)^2, the reference's form
    without 1/2 (cuda/src/cpp/forward.cpp:437-438), data from the true model.

    m0 is the starting model's (kappa, 1/rho on NM, 1/rho on MN): the
    background without the anomaly, times `start_scale`."""

    def __init__(self, n=2048, nt=3000, nshots=16, nrcv=512, dx_km=0.01, vbg=2.0, f0=12.0, precision=FP32,
                 cpst="0", seed=5, start_scale=1.0, log=None, data=None):
        from . import Solver, P
        self.n, self.nt, self.log = n, nt, log
        den = int(round(1.0 / dx_km))
        dt = 0.4 * dx_km / (math.sqrt(2.0) * 1.1 * vbg)
        cfg = config(physics=ACOUSTIC, mx=n, my=n, dx=dx_km, num_dx=1, den_dx=den, dt=dt, bc="FCCC",
                     precision=precision, cpst=cpst, nt_max=nt, stencil_form=1)
        self.s = Solver(cfg)
        rng = np.random.default_rng(seed)
        L = n * dx_km
        # shallow enough that its reflections reach the surface receivers
        # within nt steps: centre 8-15% of the width deep, radius 3-5%
        cx, cy = rng.uniform(0.08, 0.15) * L, rng.uniform(0.35, 0.65) * L
        rad = rng.uniform(0.03, 0.05) * L
        X, Y = np.meshgrid(np.arange(n + 1) * dx_km, np.arange(n + 1) * dx_km, indexing="ij")
        self.rho = np.ones((n + 1, n + 1))
        self.vp_true = vbg * (1.0 + 0.1 * (((X - cx) ** 2 + (Y - cy) ** 2) < rad ** 2))
        self.anomaly = (cx, cy, rad)
        ry = np.linspace(4, n - 5, nrcv).round().astype(np.int32)
        self.s.set_receivers(np.full(nrcv, P, np.int32), np.full(nrcv, 4, np.int32), ry)
        self.shots = [int(v) for v in np.linspace(n // 16, n - n // 16, nshots)]
        self.f0, self.t0 = f0, 1.2 / f0
        self.state0 = self.s.checkpoint()
        self.coefs = (C_KAPPA, C_BINV_VX, C_BINV_VY)
        self.s.set_model(self.rho, self.vp_true, self.vp_true, 1, den)
        truth = [self.s.get_coef(k) for k in self.coefs]
        self.data = data
        if data is None:
            self.data = []
            for i in range(nshots):
                self.data.append(self.run(truth, i)[0])
                self._log("data shot %d/%d" % (i + 1, nshots))
        vb = start_scale * vbg * np.ones_like(self.rho)
        self.s.set_model(self.rho, vb, vb, 1, den)
        self.m0 = [self.s.get_coef(k) for k in self.coefs]

    def _log(self, msg):
        if self.log:
            self.log(msg)

    def run(self, coefs, shot, residual_of=None):
        """Traces of one shot for the coefficients; with residual_of (the
        shot's data), also the adjoint gradient of sum (trace - data)^2."""
        s = self.s
        s.restore(self.state0)
        for k, v in zip(self.coefs, coefs):
            s.set_coef(k, v)
        s.set_source_ricker(GRID_MM, 4, self.shots[shot], self.f0, self.t0)
        if residual_of is None:
            s.step(self.nt)
            return s.receivers()[:, -self.nt:], None
        s.adjoint_forward(self.nt)
        tr = s.receivers()[:, -self.nt:]
        return tr, s.adjoint_run(2.0 * (tr - residual_of))

    def misfit_batched(self, coefs):
        """J for the coefficients with all shots advanced together as one
        batch (one solver per shot, one launch per half step for all of
        them, the step loop in a hipGraph); the same traces as misfit()."""
        from . import Batch, Solver
        if getattr(self, "_shot_solvers", None) is None:
            self._shot_solvers = []
            for i in range(len(self.shots)):
                s = Solver(self.s.cfg)
                f, x, y = self.s.get_receivers()
                s.set_receivers(f, x, y)
                s.set_source_ricker(GRID_MM, 4, self.shots[i], self.f0, self.t0)
                self._shot_solvers.append(s)
            self._shot_state0 = [s.checkpoint() for s in self._shot_solvers]
            self._batch = Batch(self._shot_solvers)
        for s, blob in zip(self._shot_solvers, self._shot_state0):
            s.restore(blob)
            for k, v in zip(self.coefs, coefs):
                s.set_coef(k, v)
        self._batch.step(self.nt)
        total = 0.0
        for s, d in zip(self._shot_solvers, self.data):
            tr = s.receivers()[:, -self.nt:]
            total += float(np.sum((tr - d) ** 2))
        return total

    def misfit(self, coefs, grad=False):
        """J and (with grad) its gradient, summed over the shots."""
        total, g = 0.0, [np.zeros_like(m) for m in self.m0]
        for i in range(len(self.shots)):
            tr, gi = self.run(coefs, i, self.data[i] if grad else None)
            total += float(np.sum((tr - self.data[i]) ** 2))
            if grad:
                for a, b in zip(g, gi):
                    a += b
                self._log("gradient shot %d/%d" % (i + 1, len(self.shots)))
        return total, g

    def close(self):
        if getattr(self, "_batch", None) is not None:
            self._batch.close()
            for s in self._shot_solvers:
                s.close()
            self._batch, self._shot_solvers = None, None
        self.s.close()


class ElasticShots:
    """The elastic counterpart of AdjointShots: n x n cells of the layered
    model (`nlayers` layers, 5% correlated perturbation) with a seeded circular
    anomaly of +10% in Vp and Vs as the true model; vertical-force Ricker shots
    on vx (data (N,M)) `src_depth` cells below the free surface (top 'F',
    characteristic elsewhere), receivers on vx along the surface, misfit
    J = sum over shots and samples of (trace - data)^2 with data from the true
    model.  The coefficients are (lambda, 2 mu, mu, 1/rho on NM, 1/rho on MN);
    m0 is the layered model without the anomaly."""

    def __init__(self, n=96, nt=600, nshots=2, nrcv=32, dx_km=0.005, f0=15.0, precision=FP64, seed=7,
                 nlayers=6, src_depth=4):
        from . import Solver, C_LAMBDA, C_MU, C_TWOMU
        self.n, self.nt = n, nt
        cfg = elastic_config(m=n, precision=precision, nt=nt, dx_km=dx_km)
        self.s = Solver(cfg)
        den = int(round(1.0 / dx_km))
        rho, vp, vs = _layered(n, dx_km, seed, nlayers, (1.5, 4.0), 0.05, 0.05)
        rng = np.random.default_rng(seed + 1)
        cx, cy, rad = rng.uniform(0.15, 0.25) * n, rng.uniform(0.35, 0.65) * n, rng.uniform(0.06, 0.1) * n
        X, Y = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
        bump = 1.0 + 0.1 * (((X - cx) ** 2 + (Y - cy) ** 2) < rad ** 2)
        self.rho, self.vp0, self.vs0 = rho, vp, vs
        self.vp_true, self.vs_true = vp * bump, vs * bump
        self.anomaly = (cx, cy, rad)
        self.src_depth = int(src_depth)
        ry = np.linspace(4, n - 5, nrcv).round().astype(np.int32)
        self.s.set_receivers(np.full(nrcv, VX, np.int32), np.zeros(nrcv, np.int32), ry)
        self.shots = [int(v) for v in np.linspace(n // 4, n - n // 4, nshots)]
        self.f0, self.t0 = f0, 1.25 / f0
        self.state0 = self.s.checkpoint()
        self.coefs = (C_LAMBDA, C_TWOMU, C_MU, C_BINV_VX, C_BINV_VY)
        self.s.set_model(rho, self.vp_true, self.vs_true, 1, den)
        truth = [self.s.get_coef(k) for k in self.coefs]
        self.data = [self.run(truth, i)[0] for i in range(nshots)]
        self.s.set_model(rho, vp, vs, 1, den)
        self.m0 = [self.s.get_coef(k) for k in self.coefs]

    def run(self, coefs, shot, residual_of=None):
        """Traces of one shot for the coefficients; with residual_of (the
        shot's data), also the adjoint gradient of sum (trace - data)^2, one
        array per coefficient in `coefs` order."""
        s = self.s
        s.restore(self.state0)
        for k, v in zip(self.coefs, coefs):
            s.set_coef(k, v)
        s.set_source_ricker(GRID_NM, self.src_depth, self.shots[shot], self.f0, self.t0)
        if residual_of is None:
            s.step(self.nt)
            return s.receivers()[:, -self.nt:], None
        s.adjoint_forward(self.nt)
        tr = s.receivers()[:, -self.nt:]
        g = s.adjoint_run_coefs(2.0 * (tr - residual_of))
        return tr, [g[k] for k in self.coefs]

    def misfit(self, coefs, grad=False):
        """J and (with grad) its gradient, summed over the shots."""
        total, g = 0.0, [np.zeros_like(m) for m in self.m0]
        for i in range(len(self.shots)):
            tr, gi = self.run(coefs, i, self.data[i] if grad else None)
            total += float(np.sum((tr - self.data[i]) ** 2))
            if grad:
                for a, b in zip(g, gi):
                    a += b
        return total, g

    def close(self):
        self.s.close()


def smooth_direction(rng, shape, corr):
    """Unit-peak random model perturbation, Gaussian-correlated over `corr`
    cells."""
    from scipy.ndimage import gaussian_filter
    d = gaussian_filter(rng.standard_normal(shape), corr, mode="nearest")
    return d / np.abs(d).max()
