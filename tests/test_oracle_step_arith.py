"""One time step of the oracle restated independently in exact rational
arithmetic, rounding by rounding, on periodic grids (every point interior).

16-bit storage computes each derivative-sum step v = v + S*c and each
coefficient product R = R + v*P as one fused multiply-add (one rounding to
FP32); FP32 storage as a separate product and sum (two roundings), the order of
the reference's host code (cuda/src/cpp/grid_drvt.tpp:23-34, 404-425;
singlecore/src/cpp/namespaces.hpp:566-629).  The new value is
store(float(old) + acc).  Swapping the two arithmetics, reordering the taps or
the terms, or dropping a rounding fails these tests; the restatement uses the
oracle's stored fields, coefficients and stencil table only as inputs.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle as O

M = 32  # cells per axis
# resolution of the stress / pressure data, 2^-Q: the partners of the first
# half step are 1 + k 2^-Q (k small integers) and the velocities start at 0,
# so the stencil sums cancel to a few storage ulps and the products' rounding
# (present in one arithmetic, absent in the other) reaches the stored values
Q = {O.BF16: 7, O.FP16: 10, O.FP32: 20}


def _rne(x, p, emin):
    """x (Fraction) rounded to nearest even with p significant bits and
    minimum normal exponent emin (subnormals below), as a Python float."""
    if x == 0:
        return 0.0
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fr(2) ** e > a:
        e -= 1
    e = max(e, emin)
    ulp = Fr(2) ** (e - (p - 1))
    q = a / ulp
    n = q.numerator // q.denominator
    r = q - n
    if r > Fr(1, 2) or (r == Fr(1, 2) and n % 2 == 1):
        n += 1
    v = float(n * ulp)
    return -v if x < 0 else v


def f32(x):
    return _rne(Fr(x), 24, -126)


STORE = {O.FP32: f32, O.BF16: lambda x: _rne(Fr(x), 8, -126), O.FP16: lambda x: _rne(Fr(x), 11, -14)}


def mac_fused(a, x, c):
    return _rne(Fr(x) * Fr(c) + Fr(a), 24, -126)


def mac_split(a, x, c):
    return f32(Fr(f32(Fr(x) * Fr(c))) + Fr(a))


class Step:
    """The restated step on the fields and coefficients read from the oracle."""

    def __init__(self, prec, mac, st):
        self.store, self.mac, self.st = STORE[prec], mac, st

    def dsum(self, vals):
        v = 0.0
        for x, c in zip(vals, self.st):
            v = self.mac(v, x, c)
        return v

    def dx(self, P, i, j, out_n):
        o = -2 if out_n else -1
        return self.dsum([P[(i + o + k) % M, j] for k in range(4)])

    def dy(self, P, i, j, out_n):
        o = -2 if out_n else -1
        return self.dsum([P[i, (j + o + k) % M] for k in range(4)])

    def upd(self, old, acc):
        return self.store(f32(Fr(old) + Fr(acc)))


def _oracle(physics, prec, seed):
    cfg = O.make_config(physics=physics, mx=M, my=M, dx=0.01, num_dx=1, den_dx=100, dt=1e-3, bc="PPPP",
                        precision=prec, cpst="0", stencil_form=1, nt_max=2)
    o = O.Oracle(cfg)
    rng = np.random.default_rng(seed)
    for f in o.fields:
        if f in (O.VX, O.VY):
            o.set_field(f, np.zeros(o.shape(f)))
        else:
            o.set_field(f, 1.0 + rng.integers(-8, 9, o.shape(f)) * 2.0 ** -Q[prec])
    coefs = (O.C_KAPPA, O.C_BINV_VX, O.C_BINV_VY) if physics == O.ACOUSTIC else \
        (O.C_LAMBDA, O.C_TWOMU, O.C_MU, O.C_BINV_VX, O.C_BINV_VY)
    k = {}
    for c in coefs:
        shape = O.field_shape(cfg, {O.C_KAPPA: O.P, O.C_LAMBDA: O.SXX, O.C_TWOMU: O.SXX, O.C_MU: O.SXY,
                                    O.C_BINV_VX: O.VX, O.C_BINV_VY: O.VY}[c])
        o.set_coef(c, rng.uniform(0.5, 2.0, shape) * (100.0 if c in (O.C_LAMBDA, O.C_TWOMU, O.C_MU, O.C_KAPPA)
                                                         else 1.0))
        k[c] = o.get_coef(c, shape)
    s0 = {f: o.get_field(f) for f in o.fields}
    st = list(o.table(0))
    o.step(1)
    return s0, k, st, {f: o.get_field(f) for f in o.fields}


def acoustic_step(S, s0, k):
    vx, vy, p = s0[O.VX].copy(), s0[O.VY].copy(), s0[O.P].copy()
    # velocity half (cpu_acoustic.cpp:231-322): vx from dp/dx, vy from dp/dy
    nvx = np.zeros_like(vx)
    for i in range(M + 1):
        for j in range(M):
            nvx[i, j] = S.upd(vx[i, j], S.mac(0.0, S.dx(p, i, j, True), k[O.C_BINV_VX][i, j]))
    nvy = np.zeros_like(vy)
    for i in range(M):
        for j in range(M + 1):
            nvy[i, j] = S.upd(vy[i, j], S.mac(0.0, S.dy(p, i, j, True), k[O.C_BINV_VY][i, j]))
    npp = np.zeros_like(p)
    for i in range(M):
        for j in range(M):
            kap = k[O.C_KAPPA][i, j]
            acc = S.mac(0.0, S.dx(nvx, i, j, False), kap)
            acc = S.mac(acc, S.dy(nvy, i, j, False), kap)
            npp[i, j] = S.upd(p[i, j], acc)
    return {O.VX: nvx, O.VY: nvy, O.P: npp}


def elastic_step(S, s0, k):
    vx, vy = s0[O.VX], s0[O.VY]
    sxx, syy, sxy = s0[O.SXX], s0[O.SYY], s0[O.SXY]
    # velocity half: vy (MN) from d sxy/dx and d syy/dy; vx (NM) from d sxx/dx and d sxy/dy
    nvy = np.zeros_like(vy)
    for i in range(M):
        for j in range(M + 1):
            b = k[O.C_BINV_VY][i, j]
            acc = S.mac(0.0, S.dx(sxy, i, j, False), b)
            acc = S.mac(acc, S.dy(syy, i, j, True), b)
            nvy[i, j] = S.upd(vy[i, j], acc)
    nvx = np.zeros_like(vx)
    for i in range(M + 1):
        for j in range(M):
            b = k[O.C_BINV_VX][i, j]
            acc = S.mac(0.0, S.dx(sxx, i, j, True), b)
            acc = S.mac(acc, S.dy(sxy, i, j, False), b)
            nvx[i, j] = S.upd(vx[i, j], acc)
    # stress half: sxy (NN) from d vy/dx and d vx/dy; sxx, syy (MM) share each derivative
    nxy = np.zeros_like(sxy)
    for i in range(M + 1):
        for j in range(M + 1):
            mu = k[O.C_MU][i, j]
            acc = S.mac(0.0, S.dx(nvy, i, j, True), mu)
            acc = S.mac(acc, S.dy(nvx, i, j, True), mu)
            nxy[i, j] = S.upd(sxy[i, j], acc)
    nxx, nyy = np.zeros_like(sxx), np.zeros_like(syy)
    for i in range(M):
        for j in range(M):
            lam, tmu = k[O.C_LAMBDA][i, j], k[O.C_TWOMU][i, j]
            v = S.dx(nvx, i, j, False)
            a0 = S.mac(0.0, v, lam)
            a1 = S.mac(0.0, v, lam)
            a0 = S.mac(a0, v, tmu)
            w = S.dy(nvy, i, j, False)
            a0 = S.mac(a0, w, lam)
            a1 = S.mac(a1, w, lam)
            a1 = S.mac(a1, w, tmu)
            nxx[i, j] = S.upd(sxx[i, j], a0)
            nyy[i, j] = S.upd(syy[i, j], a1)
    return {O.VX: nvx, O.VY: nvy, O.SXX: nxx, O.SYY: nyy, O.SXY: nxy}


@pytest.mark.parametrize("physics", [O.ACOUSTIC, O.ELASTIC])
@pytest.mark.parametrize("prec", [O.BF16, O.FP16, O.FP32])
def test_one_step_rounding_by_rounding(physics, prec):
    s0, k, st, got = _oracle(physics, prec, seed=3 + prec)
    fused = prec in (O.BF16, O.FP16)
    step = acoustic_step if physics == O.ACOUSTIC else elastic_step
    want = step(Step(prec, mac_fused if fused else mac_split, st), s0, k)
    other = step(Step(prec, mac_split if fused else mac_fused, st), s0, k)
    ndiff = 0
    for f in got:
        np.testing.assert_array_equal(got[f], want[f], err_msg="field %d" % f)
        ndiff += int(np.sum(other[f] != want[f]))
    # the other arithmetic gives a different step on these inputs: the check
    # tells the two apart
    assert ndiff > 0
