"""C-ABI checks that need no GPU: the library loads, exports every entry
point declared in include/tw.h, and its host-only helpers validate input."""
import ctypes
import os
import re

import numpy as np
import pytest

import twave as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "tw.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|double|const char\*)\s+(tw_\w+)\s*\(", src, re.M)))


def test_library_exports_header():
    lib = T.lib()
    decl = declared_symbols()
    assert len(decl) >= 30
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(T.exported_symbols()) == decl


def test_config_struct_layout_matches_oracle():
    import oracle as O
    assert ctypes.sizeof(T.Config) == ctypes.sizeof(O.TwConfig)
    for (a, _), (b, _) in zip(T.Config._fields_, O.TwConfig._fields_):
        assert a == b
        assert getattr(T.Config, a).offset == getattr(O.TwConfig, b).offset


def test_config_default_and_check():
    c = T.Config()
    T.lib().tw_config_default(ctypes.byref(c))
    assert c.physics == T.ELASTIC and c.bc == b"FFPP" and c.precision == T.FP32
    T.check_config(c)
    for field, bad in (("mx", 4), ("dt", 0.0), ("precision", 9), ("cpst", 99), ("source_scaling", 0.0)):
        d = T.config()
        setattr(d, field, bad)
        with pytest.raises(T.TwError):
            T.check_config(d)
    with pytest.raises(T.TwError):
        T.check_config(T.config(bc="PFPP"))        # periodic must pair
    for cp in ("K", "K8"):
        with pytest.raises(T.TwError):
            T.check_config(T.config(precision="fp32", cpst=cp))
        T.check_config(T.config(precision="bf16", cpst=cp))
    with pytest.raises(T.TwError):
        T.check_config(T.config(precision="bf16", cpst="6R"))
    # coefficient storage: 8- and 16-bit codes only with 16-bit state, nothing beyond 3
    for prec in ("bf16", "fp16"):
        for codes in (2, 3):
            T.check_config(T.config(precision=prec, cpst="K8", coef16=codes))
    for prec in ("fp32", "fp64"):
        for codes in (2, 3):
            with pytest.raises(T.TwError):
                T.check_config(T.config(precision=prec, coef16=codes))
    with pytest.raises(T.TwError):
        T.check_config(T.config(precision="bf16", coef16=4))
    with pytest.raises(T.TwError):
        T.check_config(T.config(energy_form=1))    # acoustic periodic only
    with pytest.raises(T.TwError) as e:
        T.check_config(T.config(mx=140000, my=140000))   # over the 2^34-cell guard
    assert e.value.code == 3                    # TW_ENOMEM, before any allocation


def test_field_shapes():
    c = T.config(mx=10, my=12)
    assert T.field_shape(c, T.VX) == (11, 12)
    assert T.field_shape(c, T.VY) == (10, 13)
    assert T.field_shape(c, T.SXX) == (10, 12)
    assert T.field_shape(c, T.SXY) == (11, 13)
    with pytest.raises(T.TwError):
        T.field_shape(c, T.P)


def test_locate_matches_oracle_and_kat():
    """Shipped case (cuda/data/homogeneous/SrcRcv.txt): Src_1_MN_401_250_20_250
    on 600x600, dx = 1/125 lands on data (N,M) at (10, 200), receiver
    Rcv_1_1_MN_801_250_20_250 at (10, 400)."""
    import oracle as O
    c = T.config(mx=600, my=600, num_dx=1, den_dx=125)
    # physical X -> data y, physical Y -> data x; "MN" physical = data (N, M) = GRID_NM
    assert T.locate(c, T.GRID_NM, 20, 250, 401, 250) == (10, 200)
    assert T.locate(c, T.GRID_NM, 20, 250, 801, 250) == (10, 400)
    oc = O.make_config(mx=600, my=600, num_dx=1, den_dx=125)
    rng = np.random.default_rng(0)
    for _ in range(200):
        g = int(rng.integers(0, 4))
        nx, ny = int(rng.integers(0, 1300)), int(rng.integers(0, 1300))
        den = int(rng.choice([250, 500, 125]))
        e, ix, iy = O.locate(oc, g, nx, den, ny, den)
        if e:
            with pytest.raises(T.TwError):
                T.locate(c, g, nx, den, ny, den)
        else:
            assert T.locate(c, g, nx, den, ny, den) == (ix, iy)


def test_interpolation_matches_oracle():
    import oracle as O
    c = T.config(mx=40, my=30, num_dx=1, den_dx=100)
    oc = O.make_config(mx=40, my=30, num_dx=1, den_dx=100)
    rng = np.random.default_rng(1)
    model = rng.uniform(1, 2, (21, 16))
    for g in range(4):
        np.testing.assert_array_equal(T.interpolate_model(c, g, model, 1, 50), O.interpolate(oc, g, model, 1, 50))


def test_ricker_matches_oracle():
    import oracle as O
    c = T.config(dt=1e-4, dx=1.0 / 125)
    oc = O.make_config(dt=1e-4, dx=1.0 / 125)
    for it in (0, 1, 10, 2499, 2500, 10000):
        for half in (0, 1):
            assert T.ricker_increment(c, 5.0, 0.25, it, half) == O.ricker_increment(oc, 5.0, 0.25, it, half)


@pytest.mark.parametrize("spacing", [(1, 100, 41, 31), (1, 50, 21, 16), (3, 200, 28, 21)])
def test_interpolation_adjoint_is_the_transpose(spacing):
    """<interp(x), y> = <x, interp^T(y)> on every grid, model spacings equal
    to, twice and 1.5 times the grid's (nodes beyond the model included), and
    interp^T of one grid point is that point's four weights."""
    num, den, nx, ny = spacing
    c = T.config(mx=40, my=30, num_dx=1, den_dx=100)
    rng = np.random.default_rng(7)
    for g in range(4):
        x = rng.standard_normal((nx, ny))
        y = rng.standard_normal(T.interpolate_model(c, g, x, num, den).shape)
        lhs = float(np.sum(T.interpolate_model(c, g, x, num, den) * y))
        rhs = float(np.sum(x * T.interpolate_model_adjoint(c, g, y, nx - 1, ny - 1, num, den)))
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1.0), (g, lhs, rhs)
        e = np.zeros_like(y)
        e[3, 5] = 1.0
        col = T.interpolate_model_adjoint(c, g, e, nx - 1, ny - 1, num, den)
        assert abs(col.sum() - 1.0) < 1e-15 and np.count_nonzero(col) <= 4
        np.testing.assert_allclose(np.sum(col * x), T.interpolate_model(c, g, x, num, den)[3, 5], rtol=1e-14)
    with pytest.raises(T.TwError):  # a model too small to cover the grid
        T.interpolate_model_adjoint(c, 0, np.zeros((41, 31)), 5, 5, 1, 100)


@pytest.mark.parametrize("physics", [T.ACOUSTIC, T.ELASTIC])
def test_model_gradient_chain_rule(physics):
    """twave.model_gradient against central differences of a linear functional
    of the coefficients formed as tw_set_model forms them (moduli on the model
    nodes, then interpolated; 1/rho after the interpolation), host-only."""
    c = T.config(physics=physics, mx=24, my=20, num_dx=1, den_dx=100)
    rng = np.random.default_rng(5)
    shp = (25, 21)
    rho, vp = rng.uniform(1.0, 2.0, shp), rng.uniform(2.0, 4.0, shp)
    vs = vp / 1.73 if physics == T.ELASTIC else None
    ids = (T.C_LAMBDA, T.C_TWOMU, T.C_MU, T.C_BINV_VX, T.C_BINV_VY) if physics == T.ELASTIC else \
        (T.C_KAPPA, T.C_BINV_VX, T.C_BINV_VY)

    def coefs(r, p, s):
        node = {T.C_KAPPA: r * p ** 2, T.C_MU: r * s ** 2 if s is not None else None}
        if s is not None:
            node[T.C_TWOMU] = 2 * r * s ** 2
            node[T.C_LAMBDA] = r * p ** 2 - 2 * r * s ** 2
        out = {}
        for k in ids:
            g = T.COEF_GRID[k]
            out[k] = 1.0 / T.interpolate_model(c, g, r, 1, 100) if k in (T.C_BINV_VX, T.C_BINV_VY) else \
                T.interpolate_model(c, g, node[k], 1, 100)
        return out

    c0 = coefs(rho, vp, vs)
    W = {k: rng.standard_normal(v.shape) for k, v in c0.items()}
    gr, gp, gs = T.model_gradient(c, W, rho, vp, vs, 1, 100)
    assert (gs is None) == (physics == T.ACOUSTIC)
    d = [rng.standard_normal(shp) for _ in range(3)]

    def F(h):
        r, p = rho + h * d[0], vp + h * d[1]
        s = None if vs is None else vs + h * d[2]
        return sum(float(np.sum(W[k] * v)) for k, v in coefs(r, p, s).items())

    h = 1e-6
    fd = (F(h) - F(-h)) / (2 * h)
    ad = float(np.sum(gr * d[0]) + np.sum(gp * d[1]) + (0.0 if gs is None else np.sum(gs * d[2])))
    assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)
