"""Pins the oracle's compensated update and its residual codes point by point.

The compensation variants of the leapfrog update (cuda/src/cuda/grid_update.cut:
28-151; singlecore/src/cpp/namespaces.hpp:633-668) are error-free
transformations or approximations of them, and the 16-bit residual codes have
a closed form.  Each test checks the oracle's own per-point function (the one
its time step calls) against an independent statement of the mathematics:

* 2Sum (6R, 6C) is exact for every input: S + R == s + r in exact rational
  arithmetic, and S is the rounded sum;
* Fast2Sum (3R, -3R, 3C, -3C) is exact whenever |s| >= |r|, and on inputs
  built with |r| > |s| it is NOT exact where 2Sum is, so 3 and 6 cannot be
  confused;
* the 'C' variants first round Y = r + c, then transform (s, Y), and leave R = 0;
* the K8 code of a 16-bit store is clamp(rint((t - hi) * 2^(bias - max(e, 1))),
  -127, 127) with hi the store rounded by the library's conversion (numpy /
  torch), e its exponent field; its decode is within ulp(hi)/512 of t (ulp/256
  where the code was clamped); the K residual is the 16-bit rounding of t - hi.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O


def _exact(*vals):
    return sum((Fraction(float(v)) for v in vals), Fraction(0))


def _inputs(dtype, n, seed, ordered=None):
    """Random (s, r) pairs over many magnitudes, with exact cancellations and
    ties; ordered=True keeps |s| >= |r|, False builds |r| > |s|."""
    rng = np.random.default_rng(seed)
    s = rng.standard_normal(n) * 2.0 ** rng.integers(-30, 30, n)
    r = rng.standard_normal(n) * 2.0 ** rng.integers(-30, 30, n)
    s, r = s.astype(dtype).astype(np.float64), r.astype(dtype).astype(np.float64)
    k = n // 8
    r[:k] = -s[:k]  # exact cancellation
    r[k:2 * k] = (s[k:2 * k] * 2.0 ** -np.finfo(dtype).nmant * 1.5).astype(dtype)  # ties at the last bit
    if ordered is not None:
        swap = (np.abs(r) > np.abs(s)) if ordered else (np.abs(s) >= np.abs(r))
        s[swap], r[swap] = r[swap], s[swap]
        if not ordered:
            keep = np.abs(r) > np.abs(s)
            s, r = s[keep], r[keep]
    return s, r


@pytest.mark.parametrize("prec,dtype", [(O.FP64, np.float64), (O.FP32, np.float32)])
@pytest.mark.parametrize("cpst", ["6R", "6C"])
def test_two_sum_exact(prec, dtype, cpst):
    s, r = _inputs(dtype, 4000, 1)
    c = np.zeros_like(s)
    if cpst == "6C":
        c = (np.random.default_rng(2).standard_normal(s.size) * np.abs(r) * 1e-3).astype(dtype).astype(np.float64)
    S, R, C = O.update_points(prec, cpst, s, r, c)
    y = (r.astype(dtype) + c.astype(dtype)).astype(np.float64) if cpst == "6C" else r
    np.testing.assert_array_equal(S, (s.astype(dtype) + y.astype(dtype)).astype(np.float64))
    err = C if cpst == "6C" else R
    for i in range(s.size):
        assert _exact(S[i], err[i]) == _exact(s[i], y[i]), (i, s[i], y[i], S[i], err[i])
    if cpst == "6C":
        assert not R.any()


@pytest.mark.parametrize("prec,dtype", [(O.FP64, np.float64), (O.FP32, np.float32)])
@pytest.mark.parametrize("cpst", ["3R", "-3R", "3C", "-3C"])
def test_fast_two_sum_exact_when_ordered(prec, dtype, cpst):
    s, r = _inputs(dtype, 4000, 3, ordered=True)
    c = np.zeros_like(s)
    S, R, C = O.update_points(prec, cpst, s, r, c)
    np.testing.assert_array_equal(S, (s.astype(dtype) + r.astype(dtype)).astype(np.float64))
    err = C if cpst.endswith("C") else R
    for i in range(s.size):
        assert _exact(S[i], err[i]) == _exact(s[i], r[i])
    if cpst.endswith("C"):
        assert not R.any()


@pytest.mark.parametrize("prec,dtype", [(O.FP64, np.float64), (O.FP32, np.float32)])
def test_fast_two_sum_inexact_when_unordered(prec, dtype):
    """With |r| > |s| Fast2Sum loses bits that 2Sum keeps.  Constructed case
    (binary32 / binary64 alike): s = 3 u/2, r = 1 with u = ulp(1): the sum
    rounds to 1 + 2u, Fast2Sum's correction z = (1 + 2u) - 3u/2 rounds to 1,
    so its residual is 0, while the true residual is -u/2."""
    u = float(np.finfo(dtype).eps)
    S3, R3, _ = O.update_points(prec, "3R", [1.5 * u], [1.0])
    S6, R6, _ = O.update_points(prec, "6R", [1.5 * u], [1.0])
    assert S3[0] == S6[0] == 1.0 + 2 * u
    assert R3[0] == 0.0 and R6[0] == -0.5 * u
    s, r = _inputs(dtype, 4000, 5, ordered=False)
    _, R3, _ = O.update_points(prec, "3R", s, r)
    S6, R6, _ = O.update_points(prec, "6R", s, r)
    _, Rm3, _ = O.update_points(prec, "-3R", s, r)
    np.testing.assert_array_equal(R3, Rm3)  # -3R is the same three operations
    inexact = sum(_exact(S6[i], R3[i]) != _exact(s[i], r[i]) for i in range(s.size))
    assert all(_exact(S6[i], R6[i]) == _exact(s[i], r[i]) for i in range(s.size))
    assert inexact > s.size // 10, inexact


@pytest.mark.parametrize("prec,dtype", [(O.FP64, np.float64), (O.FP32, np.float32)])
def test_plain_update(prec, dtype):
    s, r = _inputs(dtype, 1000, 7)
    S, R, C = O.update_points(prec, "0", s, r, np.ones_like(s))
    np.testing.assert_array_equal(S, (s.astype(dtype) + r.astype(dtype)).astype(np.float64))
    assert not R.any()
    np.testing.assert_array_equal(C, 1.0)  # the '0' update leaves any C alone


def _k8_cases(prec, n, seed):
    rng = np.random.default_rng(seed)
    t = rng.standard_normal(n) * 2.0 ** rng.integers(-20, 12, n)
    if prec == O.FP16:
        t = np.concatenate([t[np.abs(t) < 6e4],
                            rng.standard_normal(400) * 2.0 ** -20,  # FP16 subnormal range
                            rng.standard_normal(200) * 2.0 ** -14])
    else:
        t = np.concatenate([t, rng.standard_normal(200) * 2.0 ** -130])  # BF16 subnormal range
    one = np.float64(1.0)
    nm = 7 if prec == O.BF16 else 10
    # residuals exactly at +-ulp/2 and at code ties (k + 1/2) ulp/256
    t = np.concatenate([t, [one + 2.0 ** -(nm + 1), one - 2.0 ** -(nm + 2), -(one + 2.0 ** -(nm + 1))],
                        one + (np.arange(-127, 128) + 0.5) * 2.0 ** -(nm + 8),
                        2.0 ** np.arange(-10, 10) * (one - 2.0 ** -24)])
    return t.astype(np.float32).astype(np.float64)


def _round16(prec, x):
    import torch
    if prec == O.BF16:
        return torch.from_numpy(np.asarray(x, np.float32)).to(torch.bfloat16).double().numpy()
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float64)


def _exp_field(prec, hi):
    import torch
    if prec == O.BF16:
        bits = torch.from_numpy(hi.astype(np.float32)).to(torch.bfloat16).view(torch.int16).numpy().astype(np.int64)
        return (bits >> 7) & 0xFF, 142, 134
    bits = hi.astype(np.float16).view(np.uint16).astype(np.int64)
    return (bits >> 10) & 0x1F, 33, 25


@pytest.mark.parametrize("prec", [O.BF16, O.FP16])
def test_k8_codec_closed_form(prec):
    t = _k8_cases(prec, 20000, 11)
    hi, q = O.store16_points(prec, "K8", t)
    hi_ref = _round16(prec, t)
    np.testing.assert_array_equal(hi, hi_ref)
    e, bias, ulp_bias = _exp_field(prec, hi_ref)
    emax = np.maximum(e, 1)
    res = t - hi_ref  # exact in binary64
    q_ref = np.clip(np.rint(np.ldexp(res, bias - emax)), -127, 127)
    np.testing.assert_array_equal(q, q_ref)
    dec = O.k8_decode_points(prec, hi, q)
    np.testing.assert_array_equal(dec, np.ldexp(q_ref, emax - bias))
    ulp = np.ldexp(1.0, emax - ulp_bias)
    clamped = np.abs(np.rint(np.ldexp(res, bias - emax))) > 127
    bound = np.where(clamped, ulp / 256, ulp / 512)
    assert np.all(np.abs(t - (hi + dec)) <= bound)
    assert clamped.any()  # the +-ulp/2 residuals are clamped from 128 to 127
    # the code carries 8 more bits: the decoded value is 2^8 times closer than hi
    big = np.abs(res) > ulp / 8
    assert np.all(np.abs(t - (hi + dec))[big] <= np.abs(res[big]) / 64)


@pytest.mark.parametrize("prec", [O.BF16, O.FP16])
def test_k_residual_closed_form(prec):
    t = _k8_cases(prec, 5000, 13)
    hi, r = O.store16_points(prec, "K", t)
    np.testing.assert_array_equal(hi, _round16(prec, t))
    np.testing.assert_array_equal(r, _round16(prec, t - hi))
    hi0, r0 = O.store16_points(prec, "0", t)
    np.testing.assert_array_equal(hi0, hi)
    assert not r0.any()


def _state_with_residuals(o, prec, fields, rng):
    """Random 16-bit values with random residual codes on `fields`; returns the
    compensated values hi + decode(code), exact in binary32."""
    vals = {}
    for f in fields:
        shape = o.shape(f)
        hi = _round16(prec, rng.standard_normal(shape))
        o.set_field(f, hi)
        vals[f] = hi
    return vals


@pytest.mark.parametrize("prec", [O.BF16, O.FP16])
@pytest.mark.parametrize("bc", ["FCCC", "CFPP", "PPPP"])
def test_k8s_stencils_read_the_residual(prec, bc):
    """K8S: every stencil, SAT and boundary-value input reads the stored value
    plus its decoded residual.  Start from stress fields with residuals (the
    K8 state after a first step from stress-only initial values, which the
    oracle produces itself) and compare one step of K8S and K8 with an FP32
    run started from the exact compensated values: K8S stays within the
    codes' resolution of FP32 everywhere (closure rows and SAT boundary rows
    included), K8, whose stencils drop the residuals, is far off."""
    m = 24
    kw = dict(physics=O.ELASTIC, mx=m, my=m + 4, bc=bc, nt_max=4, dt=0.3 / 125 / (2 * 2 ** 0.5))
    rng = np.random.default_rng(3)
    runs = {}
    for cp in ("K8S", "K8"):
        o = O.Oracle(O.make_config(precision=prec, cpst=cp, **kw))
        rng = np.random.default_rng(3)
        _state_with_residuals(o, prec, [O.SXX, O.SYY, O.SXY], rng)
        o.step(1)  # creates residual codes on every field
        comp = {f: o.get_field(f) + O.k8_decode_points(prec, o.get_field(f), o.get_companion(f)).reshape(o.shape(f))
                for f in (O.VX, O.VY, O.SXX, O.SYY, O.SXY)}
        f32 = O.Oracle(O.make_config(precision=O.FP32, cpst="0", **kw))
        for f, v in comp.items():
            f32.set_field(f, v)
            np.testing.assert_array_equal(f32.get_field(f), v)  # exact in binary32
        o.step(1)
        f32.step(1)
        runs[cp] = {f: (o.get_field(f) + O.k8_decode_points(prec, o.get_field(f), o.get_companion(f)).reshape(
            o.shape(f))) - f32.get_field(f) for f in comp}
        scale = {f: np.abs(f32.get_field(f)).max() for f in comp}
        assert all(np.any(o.get_companion(f)) for f in comp)
    ulp = 2.0 ** -(7 if prec == O.BF16 else 10)
    for f in runs["K8S"]:
        e8s = np.abs(runs["K8S"][f]).max() / scale[f]
        e8 = np.abs(runs["K8"][f]).max() / scale[f]
        assert e8s < ulp / 64, (f, e8s)
        assert e8 > 8 * e8s, (f, e8, e8s)
