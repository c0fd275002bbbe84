"""The allocation-free entry point tw_check_finite_noalloc (GPU), called
through the C-ABI and checked against tw_check_finite on the same states:
two calls in a row on one device, the second on a smaller solver, so a flag
left set by the first call would show; both orders of finite and non-finite
states; every field of both storage widths."""
import ctypes

import numpy as np
import pytest

import common as C
import twave as T

pytestmark = pytest.mark.gpu


def solver(mx, my, precision, nan_field=None, seed=0):
    s, _ = C.make_pair(physics="elastic", mx=mx, my=my, bc="FCCC", precision=precision,
                       cpst="K8" if precision in ("bf16", "fp16") else "6R", nt=8)
    rng = np.random.default_rng(seed)
    for f in s.fields:
        v = 1e-3 * rng.standard_normal(s.shape(f))
        if f == nan_field:
            v[-1, -1] = np.nan  # last point of the field: the end of the launch grid
        s.set_field(f, v)
    s.step(2)
    return s


def codes(s):
    L = T.lib()
    return L.tw_check_finite(s.h), L.tw_check_finite_noalloc(s.h)


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_noalloc_matches_check_finite(precision):
    print("case noalloc", precision, flush=True)
    for f in (T.VX, T.SXY):
        big_bad = solver(96, 300, precision, nan_field=f, seed=1)
        small_ok = solver(40, 48, precision, seed=2)
        a = codes(big_bad)
        b = codes(small_ok)  # second call, smaller input: the flag must start cleared
        assert a == (5, 5), a  # TW_ENONFINITE from both
        assert b == (0, 0), b
        # the other order, noalloc first each time
        L = T.lib()
        assert L.tw_check_finite_noalloc(small_ok.h) == 0
        assert L.tw_check_finite_noalloc(big_bad.h) == 5
        assert L.tw_check_finite_noalloc(small_ok.h) == 0
        assert small_ok.check_finite_noalloc() and not big_bad.check_finite_noalloc()
        big_bad.close()
        small_ok.close()


def test_noalloc_null_handle():
    assert T.lib().tw_check_finite_noalloc(ctypes.c_void_p()) == 1  # TW_EINVAL
