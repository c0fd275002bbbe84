"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where results are checked against it.  The product library
(libtwave.so) never loads it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ELASTIC, ACOUSTIC = 0, 1
FP64, FP32, BF16, FP16 = 0, 1, 2, 3
CPST = {"0": 0, "3R": 1, "-3R": 2, "6R": 3, "3C": 4, "-3C": 5, "6C": 6, "K": 7, "K8": 8, "K8S": 9}
VX, VY, SXX, SYY, SXY, P = 0, 1, 2, 3, 4, 5
GRID_NN, GRID_MN, GRID_NM, GRID_MM = 0, 1, 2, 3
C_BINV_VX, C_BINV_VY, C_MU, C_LAMBDA, C_TWOMU, C_KAPPA = 0, 1, 2, 3, 4, 5


class TwConfig(ctypes.Structure):
    """Mirror of tw_config in include/tw.h."""
    _fields_ = [
        ("physics", ctypes.c_int32),
        ("mx", ctypes.c_int32),
        ("my", ctypes.c_int32),
        ("dx", ctypes.c_double),
        ("num_dx", ctypes.c_int64),
        ("den_dx", ctypes.c_int64),
        ("dt", ctypes.c_double),
        ("bc", ctypes.c_char * 4),
        ("precision", ctypes.c_int32),
        ("coef16", ctypes.c_int32),
        ("cpst", ctypes.c_int32),
        ("stencil_form", ctypes.c_int32),
        ("energy_form", ctypes.c_int32),
        ("record_energy", ctypes.c_int32),
        ("nt_max", ctypes.c_int32),
        ("source_scaling", ctypes.c_double),
        ("device", ctypes.c_int32),
    ]


def make_config(physics=ELASTIC, mx=64, my=64, dx=1.0 / 125, num_dx=1, den_dx=125,
                dt=1e-4, bc="FFPP", precision=FP64, coef16=0, cpst="0",
                stencil_form=0, energy_form=0, record_energy=1, nt_max=100,
                source_scaling=1.0, device=0):
    c = TwConfig()
    c.physics = physics
    c.mx, c.my = mx, my
    c.dx = dx
    c.num_dx, c.den_dx = num_dx, den_dx
    c.dt = dt
    c.bc = bc.encode()
    c.precision = precision
    c.coef16 = coef16
    c.cpst = CPST[cpst] if isinstance(cpst, str) else cpst
    c.stencil_form = stencil_form
    c.energy_form = energy_form
    c.record_energy = record_energy
    c.nt_max = nt_max
    c.source_scaling = source_scaling
    c.device = device
    return c


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        cfgp = ctypes.POINTER(TwConfig)
        L.or_create.restype = vp
        L.or_create.argtypes = [cfgp]
        L.or_destroy.argtypes = [vp]
        L.or_error.restype = ctypes.c_char_p
        L.or_set_model.argtypes = [vp, dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_long]
        L.or_set_coef.argtypes = [vp, ctypes.c_int, dp]
        L.or_get_coef.argtypes = [vp, ctypes.c_int, dp]
        L.or_set_source.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int]
        L.or_set_source_ricker.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.or_set_receivers.argtypes = [vp, ctypes.c_int, ip, ip, ip]
        L.or_step.argtypes = [vp, ctypes.c_int]
        L.or_read_receivers.argtypes = [vp, dp, ip]
        L.or_read_energy.argtypes = [vp, dp, dp, dp, ip]
        L.or_get_field.argtypes = [vp, ctypes.c_int, dp]
        L.or_set_field.argtypes = [vp, ctypes.c_int, dp]
        L.or_get_companion.argtypes = [vp, ctypes.c_int, dp]
        L.or_field_shape.argtypes = [cfgp, ctypes.c_int, ip, ip]
        L.or_locate.argtypes = [cfgp, ctypes.c_int, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_long, ip, ip]
        L.or_interpolate.argtypes = [cfgp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_long, dp]
        L.or_ricker_increment.restype = ctypes.c_double
        L.or_ricker_increment.argtypes = [cfgp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.or_table.argtypes = [vp, ctypes.c_int, dp]
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_update_points.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, dp, dp, dp]
        L.or_store16_points.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, dp, dp, dp]
        L.or_k8_decode_points.argtypes = [ctypes.c_int, ctypes.c_long, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def update_points(precision, cpst, s, r, c=None):
    """The oracle's per-point leapfrog update (reference variants, FP64/FP32)
    on independent points: returns the new (S, R, C)."""
    s = np.array(s, dtype=np.float64).ravel().copy()
    r = np.array(r, dtype=np.float64).ravel().copy()
    c = np.zeros_like(s) if c is None else np.array(c, dtype=np.float64).ravel().copy()
    rc = lib().or_update_points(int(precision), CPST[cpst] if isinstance(cpst, str) else int(cpst), s.size,
                                _dp(s), _dp(r), _dp(c))
    if rc:
        raise ValueError("or_update_points: %d" % rc)
    return s, r, c


def store16_points(precision, cpst, t):
    """The oracle's 16-bit store of FP32 sums t (BF16/FP16; cpst 0, K or K8):
    returns (stored value, K residual value or K8 code)."""
    t = np.array(t, dtype=np.float64).ravel().copy()
    hi = np.zeros_like(t)
    res = np.zeros_like(t)
    rc = lib().or_store16_points(int(precision), CPST[cpst] if isinstance(cpst, str) else int(cpst), t.size,
                                 _dp(t), _dp(hi), _dp(res))
    if rc:
        raise ValueError("or_store16_points: %d" % rc)
    return hi, res


def k8_decode_points(precision, hi, q):
    """Residual value of K8 codes q stored beside 16-bit values hi."""
    hi = np.array(hi, dtype=np.float64).ravel().copy()
    q = np.array(q, dtype=np.float64).ravel().copy()
    out = np.zeros_like(hi)
    rc = lib().or_k8_decode_points(int(precision), hi.size, _dp(hi), _dp(q), _dp(out))
    if rc:
        raise ValueError("or_k8_decode_points: %d" % rc)
    return out


def set_threads(n):
    lib().or_set_threads(int(n))


def field_shape(cfg, field):
    r, c = ctypes.c_int(), ctypes.c_int()
    e = lib().or_field_shape(ctypes.byref(cfg), field, ctypes.byref(r), ctypes.byref(c))
    if e:
        raise ValueError("field %d not in this physics" % field)
    return r.value, c.value


def locate(cfg, grid, num_x, den_x, num_y, den_y):
    ix, iy = ctypes.c_int(), ctypes.c_int()
    e = lib().or_locate(ctypes.byref(cfg), grid, num_x, den_x, num_y, den_y, ctypes.byref(ix), ctypes.byref(iy))
    return e, ix.value, iy.value


def interpolate(cfg, grid, model, num_dxm, den_dxm):
    model = np.ascontiguousarray(model, dtype=np.float64)
    nxm, nym = model.shape[0] - 1, model.shape[1] - 1
    rows = cfg.mx + (1 if grid in (GRID_NN, GRID_NM) else 0)
    cols = cfg.my + (1 if grid in (GRID_NN, GRID_MN) else 0)
    out = np.zeros((rows, cols))
    e = lib().or_interpolate(ctypes.byref(cfg), grid, _dp(model), nxm, nym, num_dxm, den_dxm, _dp(out))
    if e:
        raise ValueError(lib().or_error().decode())
    return out


def ricker_increment(cfg, f0, t0, it, half):
    return lib().or_ricker_increment(ctypes.byref(cfg), f0, t0, it, half)


class Oracle:
    """Stateful CPU solver with the same entry points as the tw_* C-ABI."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.h = lib().or_create(ctypes.byref(cfg))
        if not self.h:
            raise ValueError("oracle create failed: " + lib().or_error().decode())
        self.fields = [VX, VY, SXX, SYY, SXY] if cfg.physics == ELASTIC else [VX, VY, P]
        self.nrcv = 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def shape(self, field):
        return field_shape(self.cfg, field)

    def set_model(self, rho, vp, vs, num_dxm, den_dxm):
        rho, vp, vs = (np.ascontiguousarray(a, dtype=np.float64) for a in (rho, vp, vs))
        e = lib().or_set_model(self.h, _dp(rho), _dp(vp), _dp(vs), rho.shape[0] - 1, rho.shape[1] - 1,
                               num_dxm, den_dxm)
        if e:
            raise ValueError(lib().or_error().decode())

    def set_coef(self, k, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        if lib().or_set_coef(self.h, k, _dp(v)):
            raise ValueError(lib().or_error().decode())

    def get_coef(self, k, shape):
        out = np.zeros(shape)
        if lib().or_get_coef(self.h, k, _dp(out)):
            raise ValueError("coef %d" % k)
        return out

    def set_source(self, grid, ix, iy, increments):
        inc = np.ascontiguousarray(increments, dtype=np.float64)
        e = lib().or_set_source(self.h, grid, ix, iy, _dp(inc), len(inc))
        if e:
            raise ValueError("source rejected (%d)" % e)

    def set_source_ricker(self, grid, ix, iy, f0, t0):
        e = lib().or_set_source_ricker(self.h, grid, ix, iy, f0, t0)
        if e:
            raise ValueError("source rejected (%d)" % e)

    def set_receivers(self, fields, ixs, iys):
        f = np.ascontiguousarray(fields, dtype=np.int32)
        a = np.ascontiguousarray(ixs, dtype=np.int32)
        b = np.ascontiguousarray(iys, dtype=np.int32)
        e = lib().or_set_receivers(self.h, len(f), _ip(f), _ip(a), _ip(b))
        if e:
            raise ValueError("receivers rejected (%d)" % e)
        self.nrcv = len(f)

    def step(self, n):
        lib().or_step(self.h, n)

    def receivers(self):
        nt = ctypes.c_int()
        lib().or_read_receivers(self.h, None, ctypes.byref(nt))
        out = np.zeros((self.nrcv, nt.value))
        lib().or_read_receivers(self.h, _dp(out), ctypes.byref(nt))
        return out

    def energy(self):
        nt = ctypes.c_int()
        lib().or_read_energy(self.h, None, None, None, ctypes.byref(nt))
        ek, ep0, ep1 = (np.zeros(nt.value) for _ in range(3))
        lib().or_read_energy(self.h, _dp(ek), _dp(ep0), _dp(ep1), ctypes.byref(nt))
        return ek, ep0, ep1

    def get_field(self, field):
        out = np.zeros(self.shape(field))
        lib().or_get_field(self.h, field, _dp(out))
        return out

    def set_field(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.shape == self.shape(field)
        lib().or_set_field(self.h, field, _dp(v))

    def get_companion(self, field):
        out = np.zeros(self.shape(field))
        lib().or_get_companion(self.h, field, _dp(out))
        return out

    def table(self, which):
        out = np.zeros(32)
        n = lib().or_table(self.h, which, _dp(out))
        return out[:n]


def acoustic_reference_config(precision=FP64, nt_max=60000, multiple=1, cpst="0"):
    """The single-core acoustic reference run (singlecore/src/main_cpp/cpu_acoustic.cpp:72-87):
    600*multiple cells per axis, periodic, dx = (1/(5*2.5))/(10*multiple), dt = 1e-4,
    rho = vp = 1."""
    ppw = 10 * multiple
    dx = (1.0 / (5.0 * 2.5)) / ppw
    n = 60 * ppw
    return make_config(physics=ACOUSTIC, mx=n, my=n, dx=dx, num_dx=0, den_dx=0, dt=1e-4,
                       bc="PPPP", precision=precision, cpst=cpst, stencil_form=1,
                       energy_form=1, nt_max=nt_max)
