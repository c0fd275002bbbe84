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
