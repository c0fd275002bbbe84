"""The benchmark configurations at their full sizes (GPU).

The oracle cannot run thousands of steps on 4096^2 and 8192^2 grids in a test,
so full sizes are checked two ways:

* a short window against the oracle: the real workload (layered model, source
  at 100 m depth, 256 surface receivers, free-surface top and characteristic
  sides) from seeded random initial fields, a few steps, fields, residual codes
  and receiver traces bit for bit -- every 256-column strip, 16-row chunk and
  SBP band tile of the full-size launch geometry takes part;
* size-independent properties over the full step count: a repeated run is
  bitwise identical, the discrete energy E = E_k + (E_p0 + E_p1)/2 does not
  grow once the source has stopped (free surface conserves, characteristic
  sides dissipate), FP32 storage stays within 1e-4 relative L2 of FP64, and
  the compensated 16-bit variants keep their measured error and stay below
  the uncompensated one.
The batched-shots configuration is checked at its full grid sizes by
bitwise equality of each problem with the same problem stepped alone.
"""
import math

import numpy as np
import pytest

import common as C
import oracle as O
import twave as T
import twave.workloads as W

pytestmark = pytest.mark.gpu

DX_KM = 0.005


def elastic_pair(m, precision, cpst, coef16, nt, with_oracle=True):
    cfg = W.elastic_config(m=m, precision=precision, cpst=cpst, coef16=coef16, nt=nt)
    s = T.Solver(cfg)
    W.setup_elastic(s, m, DX_KM)
    o = None
    if with_oracle:
        # 8-bit coefficients: the oracle takes the solver's decoded values as FP32
        o = O.Oracle(O.make_config(physics=0, mx=m, my=m, dx=cfg.dx, num_dx=cfg.num_dx, den_dx=cfg.den_dx,
                                   dt=cfg.dt, bc="FCCC", precision=precision, cpst=cpst,
                                   coef16=0 if coef16 >= 2 else coef16, nt_max=nt))
        W.setup_elastic(o, m, DX_KM)
        if coef16 >= 2:
            for k in (T.C_BINV_VX, T.C_BINV_VY, T.C_MU, T.C_LAMBDA, T.C_TWOMU):
                o.set_coef(k, s.get_coef(k))
    return s, o


def random_state(s, o, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    for f in s.fields:
        v = amp * rng.standard_normal(s.shape(f))
        s.set_field(f, v)
        o.set_field(f, v)


@pytest.mark.parametrize("m,precision,cpst,coef16,steps", [
    (4096, T.FP32, "0", 0, 6),     # plain elastic FP32
    (8192, T.BF16, "K8", 1, 3),    # mixed precision: BF16 state, 8-bit residual codes, 16-bit coefficients
    (8192, T.BF16, "K8", 2, 3),    # the same with 8-bit coefficients (per-row affine codes)
    (8192, T.FP16, "K8S", 3, 3),   # FP16, residual-fed stencils, 16-bit coefficient codes
])
def test_full_size_window_bitwise(m, precision, cpst, coef16, steps):
    print("case full-size window", m, precision, cpst, coef16, steps, flush=True)
    O.set_threads(16)
    s, o = elastic_pair(m, precision, cpst, coef16, steps + 2)
    random_state(s, o, 11)
    s.step(1)
    s.step(steps - 1)  # two calls: the carried work crosses a tw_step boundary
    o.step(steps)
    C.compare_state(s, o, exact=True, companions=True)
    np.testing.assert_array_equal(s.receivers(), o.receivers())
    # energies: the same FP64 terms summed in another order (block tree vs the
    # oracle's sequential loop); a sequential sum of n positive terms is within
    # (n - 1) u of exact, here n = every point of every field
    n = sum(r * c for r, c in (s.shape(f) for f in s.fields))
    for a, b in zip(s.energy(), o.energy()):
        np.testing.assert_allclose(a, b, rtol=n * 2.0 ** -53, atol=0)
    s.close()


def run_full(m, precision, cpst, coef16, nt, model_from=None):
    s, _ = elastic_pair(m, precision, cpst, coef16, nt, with_oracle=False)
    if model_from is not None:  # the same (quantized) material in every run
        for k in (T.C_BINV_VX, T.C_BINV_VY, T.C_MU, T.C_LAMBDA, T.C_TWOMU):
            s.set_coef(k, model_from.get_coef(k))
    s.step(nt)
    return s


def state(s):
    """Fields as stored value plus decoded compensation, and receiver traces."""
    return {f: s.get_field_compensated(f) for f in s.fields}, s.receivers()


def total_energy(s):
    ek, ep0, ep1 = s.energy()
    return ek + 0.5 * (ep0 + ep1)


def source_off_step(cfg_dt, f0=15.0):
    """First step after which the Ricker source (t0 = 1.25/f0) is below 1e-6
    of its peak: |t - t0| > sqrt(ln(1e6)) / (pi f0) plus one period."""
    t_end = 1.25 / f0 + math.sqrt(math.log(1e6)) / (math.pi * f0) + 1.0 / f0
    return int(math.ceil(t_end / cfg_dt))


def test_c2_elastic_fp32_full(progress):
    """4096^2, FP32, 2000 steps: determinism, energy after the source stops,
    FP32 vs FP64 within 1e-4 relative L2 (fields and receiver traces)."""
    m, nt = 4096, 2000
    print("case C2 full", m, nt, flush=True)
    a = run_full(m, T.FP32, "0", 0, nt)
    progress("C2 first FP32 run done")
    b = run_full(m, T.FP32, "0", 0, nt)
    progress("C2 second FP32 run done")
    fa, ra = state(a)
    fb, rb = state(b)
    for f in fa:
        np.testing.assert_array_equal(fa[f], fb[f], err_msg=T.FIELD_NAMES[f])
    np.testing.assert_array_equal(ra, rb)
    b.close()
    E = total_energy(a)
    n0 = source_off_step(W.elastic_config(m=m, nt=nt).dt)
    assert n0 < nt - 100
    tail = E[n0:nt]
    assert np.all(np.isfinite(tail)) and tail[0] > 0
    growth = float(np.max(np.diff(tail) / tail[:-1]))
    print("C2 energy after the source (step %d): %.6e -> %.6e, largest one-step relative growth %.2e"
          % (n0, tail[0], tail[-1], growth), flush=True)
    assert tail.max() <= tail[0] * (1 + 1e-5)
    assert tail[-1] <= tail[0]
    s64 = run_full(m, T.FP64, "0", 0, nt, model_from=a)
    f64, r64 = state(s64)
    s64.close()
    err_f = max(C.rel_l2(fa[f], f64[f]) for f in f64)
    err_r = C.rel_l2(ra, r64)
    print("C2 FP32 vs FP64 rel-L2: fields %.3e receivers %.3e" % (err_f, err_r), flush=True)
    assert err_f < 1e-4 and err_r < 1e-4
    a.close()


def test_c3_mixed_precision_full(progress):
    """8192^2, 2000 steps (fields as stored value plus decoded residual).
    Against FP64 on the same quantized material: the K8 variant (BF16 + 8-bit
    residual codes) stays within its measured error and well below
    uncompensated BF16, and a repeated K8 run is bitwise identical; with the
    residual-fed stencils (K8S) FP16 storage is within 1e-4 and BF16 well below
    K8.  Against FP64 on the exact material: FP16 + K8S with 16-bit
    coefficient codes within 1e-4, 8-bit codes not (the coefficient
    quantization is then the larger error)."""
    m, nt = 8192, 2000
    print("case C3 full", m, nt, flush=True)

    def err(run, ref):
        (fa, ra), (fb, rb) = run, ref
        return max(C.rel_l2(fa[f], fb[f]) for f in fb), C.rel_l2(ra, rb)

    def full(precision, cpst, coef16, model_from=None):
        s = run_full(m, precision, cpst, coef16, nt, model_from=model_from)
        out = state(s)
        progress("C3 run done: precision", precision, cpst, "coef16", coef16)
        return s, out

    k8, rk8 = full(T.BF16, "K8", 1)
    k8b, rk8b = full(T.BF16, "K8", 1, model_from=k8)
    k8b.close()
    for f in rk8[0]:
        np.testing.assert_array_equal(rk8[0][f], rk8b[0][f], err_msg=T.FIELD_NAMES[f])
    np.testing.assert_array_equal(rk8[1], rk8b[1])
    del rk8b
    s64, ref16 = full(T.FP64, "0", 0, model_from=k8)  # the 16-bit coefficients' values
    s64.close()
    e = {"bf16-K8": err(rk8, ref16)}
    del rk8
    for name, prec, cp in (("bf16-0", T.BF16, "0"), ("fp16-K8", T.FP16, "K8")):
        x, r = full(prec, cp, 1, model_from=k8)
        x.close()
        e[name] = err(r, ref16)
        del r
    k8.close()
    del ref16
    # 16-bit coefficient codes: FP64 on their decoded values and on the exact material
    c16, r16 = full(T.FP16, "K8S", 3)
    s64, refc = full(T.FP64, "0", 0, model_from=c16)
    s64.close()
    e["fp16-K8S-c16"] = err(r16, refc)
    b16, rb16 = full(T.BF16, "K8S", 3, model_from=c16)
    b16.close()
    e["bf16-K8S-c16"] = err(rb16, refc)
    del rb16, refc
    c16.close()
    s64, refx = full(T.FP64, "0", 0)
    s64.close()
    e["fp16-K8S-c16 exact material"] = err(r16, refx)
    del r16
    c8, r8 = full(T.FP16, "K8S", 2)
    c8.close()
    e["fp16-K8S-c8 exact material"] = err(r8, refx)
    del r8, refx
    print("C3 rel-L2 vs FP64 (fields, receivers):", e, flush=True)
    assert e["bf16-K8"][0] < 5e-2 and e["bf16-K8"][1] < 5e-2
    assert e["bf16-K8"][0] < 0.5 * e["bf16-0"][0]
    assert e["fp16-K8"][0] < 1e-2 and e["fp16-K8"][1] < 1e-2
    assert e["fp16-K8S-c16"][0] < 1e-4 and e["fp16-K8S-c16"][1] < 1e-4
    assert e["bf16-K8S-c16"][0] < 0.05 * e["bf16-K8"][0]
    assert e["fp16-K8S-c16 exact material"][0] < 1e-4 and e["fp16-K8S-c16 exact material"][1] < 1e-4
    assert e["fp16-K8S-c8 exact material"][0] > 1e-4


def test_c4_batch_full_sizes():
    """Batched acoustic shots at the configuration's extreme grid sizes
    (512^2 and 2048^2, plus one in between), smooth random velocity, Ricker at
    a random interior point, BF16 storage: every problem of the batch equals
    the same problem stepped alone, bit for bit."""
    print("case C4 full sizes", flush=True)
    sizes = [512, 1300, 2048]
    dx, vmax = 0.01, 4.5

    def build():
        rng = np.random.default_rng(4)
        out = []
        for i, n in enumerate(sizes):
            rho, vp, _ = W.smooth_acoustic_model(n, dx, seed=1000 + i)
            cfg = T.config(physics=T.ACOUSTIC, mx=n, my=n, dx=dx, num_dx=1, den_dx=100,
                           dt=0.5 * dx / (math.sqrt(2.0) * vmax), bc="CCCC", precision=T.BF16, cpst="0",
                           coef16=1, nt_max=300, record_energy=0, stencil_form=1)
            s = T.Solver(cfg)
            s.set_model(rho, vp, vp, 1, 100)
            f0 = float(rng.uniform(10.0, 25.0))
            ix, iy = (int(v) for v in rng.integers(n // 8, n - n // 8, 2))
            s.set_source_ricker(T.GRID_MM, ix, iy, f0, 1.25 / f0)
            out.append(s)
        return out

    batched, solo = build(), build()
    B = T.Batch(batched)
    B.step(150)
    B.step(50)
    for s in solo:
        s.step(200)
    for x, y in zip(batched, solo):
        for f in x.fields:
            np.testing.assert_array_equal(x.get_field(f), y.get_field(f), err_msg=T.FIELD_NAMES[f])
        assert np.any(x.get_field(T.P) != 0)
    B.close()
    for s in batched + solo:
        s.close()


def c4_problems(sizes, dx=0.01, vmax=4.5, nt=1600, seed=4, only=None):
    """Problems of the batched-shots configuration (the kind bench.py --workload c4 runs):
    n x n cells, smooth random velocity 1.5-4.5 km/s, 10-25 Hz Ricker at a
    random interior point, characteristic sides, BF16 storage, 16-bit
    coefficients; returns the solvers (of the indices in `only`, default all:
    the same problems whatever the selection) and (f0, t0, ix, iy) of each
    source."""
    rng = np.random.default_rng(seed)
    params = []
    for n in sizes:
        f0 = float(rng.uniform(10.0, 25.0))
        ix, iy = (int(v) for v in rng.integers(n // 8, n - n // 8, 2))
        params.append((f0, ix, iy))
    out, src = [], []
    for i, n in enumerate(sizes):
        if only is not None and i not in only:
            continue
        rho, vp, _ = W.smooth_acoustic_model(n, dx, seed=1000 + i)
        cfg = T.config(physics=T.ACOUSTIC, mx=n, my=n, dx=dx, num_dx=1, den_dx=100,
                       dt=0.5 * dx / (math.sqrt(2.0) * vmax), bc="CCCC", precision=T.BF16, cpst="0",
                       coef16=1, nt_max=nt, record_energy=0, stencil_form=1)
        s = T.Solver(cfg)
        s.set_model(rho, vp, vp, 1, 100)
        f0, ix, iy = params[i]
        s.set_source_ricker(T.GRID_MM, ix, iy, f0, 1.25 / f0)
        out.append(s)
        src.append((f0, 1.25 / f0, ix, iy))
    return out, src


def test_c4_batch_as_configured(progress):
    """The batched-shots configuration as configured: 64 seeded problems with
    n uniform in [512, 2048], the tile height chosen by the autotuner, the
    1500-step loop as one graph.  The smallest and the largest problem equal
    the same problems stepped alone, bit for bit; then a 3-step window of the
    batch against the CPU oracle started from the smallest problem's state."""
    print("case C4 as configured", flush=True)
    rng = np.random.default_rng(4)
    sizes = [int(v) for v in rng.integers(512, 2049, 64)]
    K = 1500
    batched, src = c4_problems(sizes)
    lo, hi = int(np.argmin(sizes)), int(np.argmax(sizes))
    solo, _ = c4_problems(sizes, only=(lo, hi))
    B = T.Batch(batched)
    rows = B.autotune(20)
    assert rows in (4, 8, 16, 32, 64)
    assert all(s.current_step() == 0 for s in batched)
    progress("C4 autotuned to %d rows" % rows)
    B.step(K)
    progress("C4 batch of %d problems advanced %d steps" % (len(batched), K))
    for s in solo:
        s.step(K)
    for k, s in zip((lo, hi), solo):
        for f in s.fields:
            np.testing.assert_array_equal(batched[k].get_field(f), s.get_field(f), err_msg="%d %s" % (sizes[k], f))
        assert np.any(s.get_field(T.P) != 0)
    # window: the oracle from the smallest problem's state at step K
    x = batched[lo]
    cfg = x.cfg
    o = O.Oracle(O.make_config(physics=1, mx=cfg.mx, my=cfg.my, dx=cfg.dx, num_dx=cfg.num_dx, den_dx=cfg.den_dx,
                               dt=cfg.dt, bc="CCCC", precision=O.BF16, cpst="0", coef16=1, nt_max=8,
                               record_energy=0, stencil_form=1))
    for k in (T.C_BINV_VX, T.C_BINV_VY, T.C_KAPPA):
        o.set_coef(k, x.get_coef(k))
    for f in x.fields:
        o.set_field(f, x.get_field(f))
    f0, t0, ix, iy = src[lo]
    inc = [O.ricker_increment(o.cfg, f0, t0, K + t, 1) for t in range(3)]
    o.set_source(O.GRID_MM, ix, iy, np.array(inc))
    B.step(3)
    o.step(3)
    for f in x.fields:
        np.testing.assert_array_equal(x.get_field(f), o.get_field(f), err_msg="window %s" % T.FIELD_NAMES[f])
    B.close()
    for s in batched + solo:
        s.close()
