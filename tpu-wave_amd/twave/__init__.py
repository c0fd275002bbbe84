"""twave -- Python binding of the MI355X staggered SBP-SAT wave solver.

A thin ctypes layer over the C-ABI in include/tw.h (libtwave.so, built for
gfx950).  It mirrors the reference solver's entry points: grid/material
setup, source and receivers, time stepping, receiver and energy readout.
There is no CPU fallback: every call runs the HIP library, and a missing
library is an error.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtwave.so")

ELASTIC, ACOUSTIC = 0, 1
FP64, FP32, BF16, FP16 = 0, 1, 2, 3
PRECISION = {"fp64": FP64, "fp32": FP32, "bf16": BF16, "fp16": FP16}
CPST = {"0": 0, "3R": 1, "-3R": 2, "6R": 3, "3C": 4, "-3C": 5, "6C": 6, "K": 7, "K8": 8, "K8S": 9}
VX, VY, SXX, SYY, SXY, P = 0, 1, 2, 3, 4, 5
FIELD_NAMES = {VX: "vx", VY: "vy", SXX: "sxx", SYY: "syy", SXY: "sxy", P: "p"}
GRID_NN, GRID_MN, GRID_NM, GRID_MM = 0, 1, 2, 3
C_BINV_VX, C_BINV_VY, C_MU, C_LAMBDA, C_TWOMU, C_KAPPA = 0, 1, 2, 3, 4, 5

ERRORS = {1: "TW_EINVAL", 2: "TW_ERANGE", 3: "TW_ENOMEM", 4: "TW_EDEVICE", 5: "TW_ENONFINITE", 6: "TW_ESTATE"}


class TwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, str(code)), msg))
        self.code = code


class Config(ctypes.Structure):
    """tw_config (include/tw.h)."""
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


class Input(ctypes.Structure):
    """tw_input (include/tw.h): the reference's InputFile.txt parameters."""
    _fields_ = [
        ("model_m", ctypes.c_int32 * 2),
        ("grid_m", ctypes.c_int32 * 2),
        ("num_dx_model", ctypes.c_int64),
        ("den_dx_model", ctypes.c_int64),
        ("num_dx", ctypes.c_int64),
        ("den_dx", ctypes.c_int64),
        ("bc", ctypes.c_char * 4),
        ("nt", ctypes.c_int32),
        ("dt_max", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("central_f", ctypes.c_double),
        ("time_delay", ctypes.c_double),
        ("device", ctypes.c_int32),
        ("energy", ctypes.c_char),
        ("medium", ctypes.c_char * 64),
    ]


class SrcRcvEntry(ctypes.Structure):
    """tw_srcrcv_entry (include/tw.h): one Src_ or Rcv_ line of SrcRcv.txt."""
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("src_index", ctypes.c_int32),
        ("rcv_index", ctypes.c_int32),
        ("grid", ctypes.c_char * 2),
        ("num", ctypes.c_int64 * 2),
        ("den", ctypes.c_int64 * 2),
    ]


_lib = None

# (name, restype, argtypes)
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_cp = ctypes.POINTER(Config)
_SIGS = [
    ("tw_config_default", None, [_cp]),
    ("tw_config_check", ctypes.c_int, [_cp]),
    ("tw_field_shape", ctypes.c_int, [_cp, ctypes.c_int32, _ip, _ip]),
    ("tw_locate", ctypes.c_int, [_cp, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                 ctypes.c_int64, _ip, _ip]),
    ("tw_interpolate_model", ctypes.c_int, [_cp, ctypes.c_int32, _dp, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int64, ctypes.c_int64, _dp]),
    ("tw_interpolate_model_adjoint", ctypes.c_int, [_cp, ctypes.c_int32, _dp, ctypes.c_int32, ctypes.c_int32,
                                                    ctypes.c_int64, ctypes.c_int64, _dp]),
    ("tw_ricker_increment", ctypes.c_double, [_cp, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                              ctypes.c_int32]),
    ("tw_global_error", ctypes.c_char_p, []),
    ("tw_version", ctypes.c_char_p, []),
    ("tw_create", ctypes.c_int, [_cp, ctypes.POINTER(_vp)]),
    ("tw_destroy", ctypes.c_int, [_vp]),
    ("tw_last_error", ctypes.c_char_p, [_vp]),
    ("tw_set_model", ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                    ctypes.c_int64]),
    ("tw_set_coef", ctypes.c_int, [_vp, ctypes.c_int32, _dp]),
    ("tw_get_coef", ctypes.c_int, [_vp, ctypes.c_int32, _dp]),
    ("tw_set_source_ricker", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_double, ctypes.c_double]),
    ("tw_set_source_wavelet", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp,
                                             ctypes.c_int32]),
    ("tw_clear_source", ctypes.c_int, [_vp]),
    ("tw_set_receivers", ctypes.c_int, [_vp, ctypes.c_int32, _ip, _ip, _ip]),
    ("tw_get_receivers", ctypes.c_int, [_vp, ctypes.c_int32, _ip, _ip, _ip, _ip]),
    ("tw_step", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("tw_synchronize", ctypes.c_int, [_vp]),
    ("tw_current_step", ctypes.c_int, [_vp, _ip]),
    ("tw_read_receivers", ctypes.c_int, [_vp, _dp, _ip]),
    ("tw_read_energy", ctypes.c_int, [_vp, _dp, _dp, _dp, _ip]),
    ("tw_get_field", ctypes.c_int, [_vp, ctypes.c_int32, _dp]),
    ("tw_set_field", ctypes.c_int, [_vp, ctypes.c_int32, _dp]),
    ("tw_get_companion", ctypes.c_int, [_vp, ctypes.c_int32, _dp]),
    ("tw_get_field_compensated", ctypes.c_int, [_vp, ctypes.c_int32, _dp]),
    ("tw_reset", ctypes.c_int, [_vp]),
    ("tw_checkpoint_size", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    ("tw_checkpoint_save", ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    ("tw_checkpoint_load", ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    ("tw_set_stream", ctypes.c_int, [_vp, _vp]),
    ("tw_check_finite", ctypes.c_int, [_vp]),
    ("tw_check_finite_noalloc", ctypes.c_int, [_vp]),
    ("tw_stream_copy", ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp]),
    ("tw_get_config", ctypes.c_int, [_vp, _cp]),
    ("tw_input_default", None, [ctypes.POINTER(Input)]),
    ("tw_read_input_file", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(Input)]),
    ("tw_input_option", ctypes.c_int, [ctypes.POINTER(Input), ctypes.c_char_p, ctypes.c_char_p]),
    ("tw_input_to_config", ctypes.c_int, [ctypes.POINTER(Input), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _cp]),
    ("tw_read_srcrcv_file", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(SrcRcvEntry), ctypes.c_int32, _ip]),
    ("tw_read_model_file", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p, _dp]),
    ("tw_setup_shot", ctypes.c_int, [_vp, ctypes.POINTER(Input), ctypes.POINTER(SrcRcvEntry), ctypes.c_int32,
                                     ctypes.c_int32]),
    ("tw_write_snapshots", ctypes.c_int, [_vp, ctypes.c_char_p]),
    ("tw_write_results", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(SrcRcvEntry), ctypes.c_int32,
                                        ctypes.c_int32]),
    ("tw_result_folder", ctypes.c_int, [ctypes.POINTER(Input), _cp, ctypes.c_char_p, ctypes.c_char_p,
                                        ctypes.c_int32]),
    ("tw_data_misfit", ctypes.c_int, [_vp, _dp, ctypes.c_int32, _dp, _dp]),
    ("tw_profile_begin", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("tw_profile_end", ctypes.c_int, [_vp, _dp, _dp, _ip]),
    ("tw_adjoint_forward", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    ("tw_adjoint_run", ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
    ("tw_adjoint_run_coefs", ctypes.c_int, [_vp, _dp, ctypes.POINTER(_dp)]),
    ("tw_batch_create", ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int32, ctypes.POINTER(_vp)]),
    ("tw_batch_destroy", ctypes.c_int, [_vp]),
    ("tw_batch_last_error", ctypes.c_char_p, [_vp]),
    ("tw_batch_step", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("tw_batch_set_rows", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("tw_batch_elapsed", ctypes.c_int, [_vp, _dp]),
    ("tw_batch_autotune", ctypes.c_int, [_vp, ctypes.c_int32, _ip]),
]


def lib():
    """Load libtwave.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtwave.so not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return [s[0] for s in _SIGS]


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def config(physics=ELASTIC, mx=64, my=64, dx=1.0 / 125, num_dx=1, den_dx=125, dt=1e-4, bc="FFPP",
           precision=FP32, coef16=0, cpst="0", stencil_form=0, energy_form=0, record_energy=1, nt_max=1000,
           source_scaling=1.0, device=0):
    c = Config()
    c.physics = physics
    c.mx, c.my = mx, my
    c.dx = dx
    c.num_dx, c.den_dx = num_dx, den_dx
    c.dt = dt
    c.bc = bc.encode() if isinstance(bc, str) else bc
    c.precision = PRECISION[precision] if isinstance(precision, str) else precision
    c.coef16 = coef16
    c.cpst = CPST[cpst] if isinstance(cpst, str) else cpst
    c.stencil_form = stencil_form
    c.energy_form = energy_form
    c.record_energy = record_energy
    c.nt_max = nt_max
    c.source_scaling = source_scaling
    c.device = device
    return c


def check_config(cfg):
    e = lib().tw_config_check(ctypes.byref(cfg))
    if e:
        raise TwError(e, lib().tw_global_error().decode())


def field_shape(cfg, field):
    r, c = ctypes.c_int32(), ctypes.c_int32()
    e = lib().tw_field_shape(ctypes.byref(cfg), field, ctypes.byref(r), ctypes.byref(c))
    if e:
        raise TwError(e, lib().tw_global_error().decode())
    return r.value, c.value


def locate(cfg, grid, num_x, den_x, num_y, den_y):
    """Rational physical location -> (ix, iy) on a grid (forward.cpp:30-54)."""
    ix, iy = ctypes.c_int32(), ctypes.c_int32()
    e = lib().tw_locate(ctypes.byref(cfg), grid, num_x, den_x, num_y, den_y, ctypes.byref(ix), ctypes.byref(iy))
    if e:
        raise TwError(e, lib().tw_global_error().decode())
    return ix.value, iy.value


def interpolate_model(cfg, grid, model, num_dxm, den_dxm):
    model = np.ascontiguousarray(model, dtype=np.float64)
    rows = cfg.mx + (1 if grid in (GRID_NN, GRID_NM) else 0)
    cols = cfg.my + (1 if grid in (GRID_NN, GRID_MN) else 0)
    out = np.zeros((rows, cols))
    e = lib().tw_interpolate_model(ctypes.byref(cfg), grid, _d(model), model.shape[0] - 1, model.shape[1] - 1,
                                   num_dxm, den_dxm, _d(out))
    if e:
        raise TwError(e, lib().tw_global_error().decode())
    return out


def interpolate_model_adjoint(cfg, grid, values, nxm, nym, num_dxm, den_dxm):
    """Transpose of interpolate_model: values on one staggered grid -> the
    (nxm+1) x (nym+1) model nodes (tw_interpolate_model_adjoint)."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    out = np.zeros((nxm + 1, nym + 1))
    e = lib().tw_interpolate_model_adjoint(ctypes.byref(cfg), grid, _d(v), nxm, nym, num_dxm, den_dxm, _d(out))
    if e:
        raise TwError(e, lib().tw_global_error().decode())
    return out


# grid of each material coefficient
COEF_GRID = {C_KAPPA: GRID_MM, C_LAMBDA: GRID_MM, C_TWOMU: GRID_MM, C_MU: GRID_NN, C_BINV_VX: GRID_NM,
             C_BINV_VY: GRID_MN}


def model_gradient(cfg, grads, rho, vp, vs, num_dxm, den_dxm):
    """Gradients with respect to the coefficients ({coefficient id: gradient
    on its grid}, as adjoint_run_coefs returns them) -> gradients with respect
    to the model of set_model on its nodes: (d/d rho, d/d vp, d/d vs), the
    last None for acoustic.  set_model forms the moduli on the nodes and
    interpolates them (kappa = rho vp^2; lambda = rho vp^2 - 2 rho vs^2,
    2 mu = 2 rho vs^2, mu = rho vs^2) and takes 1/rho as one over the
    interpolated density, so the chain rule runs through the transposed
    interpolation (the coefficient rounding of the storage type is taken as
    the identity)."""
    rho, vp = np.asarray(rho, np.float64), np.asarray(vp, np.float64)
    vs = None if vs is None else np.asarray(vs, np.float64)
    nxm, nym = rho.shape[0] - 1, rho.shape[1] - 1

    def back(k, v):
        return interpolate_model_adjoint(cfg, COEF_GRID[k], v, nxm, nym, num_dxm, den_dxm)

    g_rho, g_vp = np.zeros_like(rho), np.zeros_like(rho)
    g_vs = None if vs is None else np.zeros_like(rho)
    if C_KAPPA in grads:
        G = back(C_KAPPA, grads[C_KAPPA])
        g_rho += G * vp ** 2
        g_vp += G * 2 * rho * vp
    if C_LAMBDA in grads:
        G = back(C_LAMBDA, grads[C_LAMBDA])
        g_rho += G * (vp ** 2 - 2 * vs ** 2)
        g_vp += G * 2 * rho * vp
        g_vs += G * (-4 * rho * vs)
    if C_TWOMU in grads:
        G = back(C_TWOMU, grads[C_TWOMU])
        g_rho += G * 2 * vs ** 2
        g_vs += G * 4 * rho * vs
    if C_MU in grads:
        G = back(C_MU, grads[C_MU])
        g_rho += G * vs ** 2
        g_vs += G * 2 * rho * vs
    for k in (C_BINV_VX, C_BINV_VY):
        if k in grads:
            ri = interpolate_model(cfg, COEF_GRID[k], rho, num_dxm, den_dxm)
            g_rho += back(k, -np.asarray(grads[k]) / ri ** 2)
    return g_rho, g_vp, g_vs


def ricker_increment(cfg, f0, t0, it, half):
    return lib().tw_ricker_increment(ctypes.byref(cfg), f0, t0, it, half)


class Solver:
    """One solver handle (tw_create ... tw_destroy) on a HIP device."""

    def __init__(self, cfg):
        self.cfg = cfg
        h = ctypes.c_void_p()
        e = lib().tw_create(ctypes.byref(cfg), ctypes.byref(h))
        if e:
            raise TwError(e, lib().tw_global_error().decode())
        self.h = h
        self.nrcv = 0
        self._adj_steps = 0
        self.fields = [VX, VY, SXX, SYY, SXY] if cfg.physics == ELASTIC else [VX, VY, P]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().tw_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, e):
        if e:
            raise TwError(e, lib().tw_last_error(self.h).decode())

    def shape(self, field):
        return field_shape(self.cfg, field)

    def set_model(self, rho, vp, vs, num_dxm, den_dxm):
        rho, vp, vs = (np.ascontiguousarray(a, dtype=np.float64) for a in (rho, vp, vs))
        if rho.ndim != 2 or vp.shape != rho.shape or (self.cfg.physics == ELASTIC and vs.shape != rho.shape):
            raise ValueError("rho, vp (and vs) must be 2-D node models of one shape")
        self._chk(lib().tw_set_model(self.h, _d(rho), _d(vp), _d(vs), rho.shape[0] - 1, rho.shape[1] - 1,
                                     num_dxm, den_dxm))

    def set_coef(self, k, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        want = self.get_coef(k).shape
        if v.shape != want:
            raise ValueError("coefficient shape %s, expected %s" % (v.shape, want))
        self._chk(lib().tw_set_coef(self.h, k, _d(v)))

    def get_coef(self, k):
        grid = {C_BINV_VX: GRID_NM, C_BINV_VY: GRID_MN, C_MU: GRID_NN}.get(k, GRID_MM)
        rows = self.cfg.mx + (1 if grid in (GRID_NN, GRID_NM) else 0)
        cols = self.cfg.my + (1 if grid in (GRID_NN, GRID_MN) else 0)
        out = np.zeros((rows, cols))
        self._chk(lib().tw_get_coef(self.h, k, _d(out)))
        return out

    def set_source_ricker(self, grid, ix, iy, f0, t0):
        self._chk(lib().tw_set_source_ricker(self.h, grid, ix, iy, f0, t0))

    def set_source(self, grid, ix, iy, increments):
        inc = np.ascontiguousarray(increments, dtype=np.float64)
        self._chk(lib().tw_set_source_wavelet(self.h, grid, ix, iy, _d(inc), len(inc)))

    def clear_source(self):
        self._chk(lib().tw_clear_source(self.h))

    def set_receivers(self, fields, ixs, iys):
        f = np.ascontiguousarray(fields, dtype=np.int32)
        a = np.ascontiguousarray(ixs, dtype=np.int32)
        b = np.ascontiguousarray(iys, dtype=np.int32)
        self._chk(lib().tw_set_receivers(self.h, len(f), _i(f), _i(a), _i(b)))
        self.nrcv = len(f)

    def get_receivers(self):
        """Registered receivers in record order: (fields, ix, iy)."""
        n = ctypes.c_int32()
        self._chk(lib().tw_get_receivers(self.h, 0, ctypes.byref(n), None, None, None))
        out = [np.zeros(n.value, dtype=np.int32) for _ in range(3)]
        if n.value:
            self._chk(lib().tw_get_receivers(self.h, n.value, ctypes.byref(n), *(_i(a) for a in out)))
        self.nrcv = n.value
        return tuple(out)

    def step(self, n):
        self._chk(lib().tw_step(self.h, n))

    def synchronize(self):
        self._chk(lib().tw_synchronize(self.h))

    def current_step(self):
        it = ctypes.c_int32()
        self._chk(lib().tw_current_step(self.h, ctypes.byref(it)))
        return it.value

    def receivers(self):
        nt = ctypes.c_int32()
        self._chk(lib().tw_read_receivers(self.h, None, ctypes.byref(nt)))
        out = np.zeros((self.nrcv, nt.value))
        if self.nrcv:
            self._chk(lib().tw_read_receivers(self.h, _d(out), ctypes.byref(nt)))
        return out

    def energy(self):
        nt = ctypes.c_int32()
        self._chk(lib().tw_read_energy(self.h, None, None, None, ctypes.byref(nt)))
        ek, ep0, ep1 = (np.zeros(nt.value) for _ in range(3))
        self._chk(lib().tw_read_energy(self.h, _d(ek), _d(ep0), _d(ep1), ctypes.byref(nt)))
        return ek, ep0, ep1

    def get_field(self, field):
        out = np.zeros(self.shape(field))
        self._chk(lib().tw_get_field(self.h, field, _d(out)))
        return out

    def set_field(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != self.shape(field):
            raise ValueError("field shape %s, expected %s" % (v.shape, self.shape(field)))
        self._chk(lib().tw_set_field(self.h, field, _d(v)))

    def get_field_compensated(self, field):
        """Stored value plus its compensation (decoded K8 residual, K/R/C
        companion): tw_get_field_compensated."""
        out = np.zeros(self.shape(field))
        self._chk(lib().tw_get_field_compensated(self.h, field, _d(out)))
        return out

    def get_companion(self, field):
        out = np.zeros(self.shape(field))
        self._chk(lib().tw_get_companion(self.h, field, _d(out)))
        return out

    def reset(self):
        self._chk(lib().tw_reset(self.h))

    def checkpoint(self):
        n = ctypes.c_size_t()
        self._chk(lib().tw_checkpoint_size(self.h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._chk(lib().tw_checkpoint_save(self.h, ctypes.cast(buf, ctypes.c_void_p), n.value))
        return buf.raw

    def restore(self, blob):
        buf = ctypes.create_string_buffer(blob, len(blob))
        self._chk(lib().tw_checkpoint_load(self.h, ctypes.cast(buf, ctypes.c_void_p), len(blob)))

    def set_stream(self, stream_ptr):
        self._chk(lib().tw_set_stream(self.h, ctypes.c_void_p(stream_ptr)))

    def adjoint_forward(self, nsteps, interval=0):
        """Advance nsteps keeping sqrt(N) checkpoints for adjoint_run /
        adjoint_run_coefs (FP64/FP32 storage, cpst 0)."""
        self._chk(lib().tw_adjoint_forward(self.h, int(nsteps), int(interval)))
        self._adj_steps = int(nsteps)

    def adjoint_run(self, rcv_adjoint):
        """dL/d(coefficients) for rcv_adjoint[n_rcv][nsteps] = dL/d(trace) of
        the last adjoint_forward run.  Returns (d/d kappa on MM, d/d(1/rho) on
        NM, d/d(1/rho) on MN)."""
        w = np.ascontiguousarray(rcv_adjoint, dtype=np.float64)
        if self._adj_steps and (w.ndim != 2 or w.shape[0] != self.nrcv or w.shape[1] != self._adj_steps):
            raise ValueError("rcv_adjoint shape %s, expected (%d, %d)" % (w.shape, self.nrcv, self._adj_steps))
        gk = np.zeros(self.get_coef(C_KAPPA).shape)
        gx = np.zeros(self.get_coef(C_BINV_VX).shape)
        gy = np.zeros(self.get_coef(C_BINV_VY).shape)
        self._chk(lib().tw_adjoint_run(self.h, _d(w), _d(gk), _d(gx), _d(gy)))
        return gk, gx, gy

    def adjoint_run_coefs(self, rcv_adjoint):
        """dL/d(coefficient) for every coefficient of the physics
        (tw_adjoint_run_coefs): {coefficient id: gradient on its grid}; elastic
        C_LAMBDA, C_TWOMU (MM), C_MU (NN), C_BINV_VX (NM), C_BINV_VY (MN)."""
        w = np.ascontiguousarray(rcv_adjoint, dtype=np.float64)
        if self._adj_steps and (w.ndim != 2 or w.shape[0] != self.nrcv or w.shape[1] != self._adj_steps):
            raise ValueError("rcv_adjoint shape %s, expected (%d, %d)" % (w.shape, self.nrcv, self._adj_steps))
        ks = (C_BINV_VX, C_BINV_VY, C_MU, C_LAMBDA, C_TWOMU) if self.cfg.physics == ELASTIC else \
            (C_KAPPA, C_BINV_VX, C_BINV_VY)
        out = {k: np.zeros(self.get_coef(k).shape) for k in ks}
        ptrs = (_dp * 6)()
        for k, g in out.items():
            ptrs[k] = _d(g)
        self._chk(lib().tw_adjoint_run_coefs(self.h, _d(w), ptrs))
        return out

    def check_finite(self):
        e = lib().tw_check_finite(self.h)
        return e == 0 if e in (0, 5) else self._chk(e)

    def check_finite_noalloc(self):
        """check_finite without device allocation (tw_check_finite_noalloc)."""
        e = lib().tw_check_finite_noalloc(self.h)
        return e == 0 if e in (0, 5) else self._chk(e)

    def data_misfit(self, data):
        """(per-record sum of (trace - data)^2, total) over the first
        data.shape[1] recorded steps (tw_data_misfit)."""
        d = np.ascontiguousarray(data, dtype=np.float64)
        if d.ndim != 2 or d.shape[0] != self.nrcv or d.shape[1] > self.current_step() or \
                d.shape[1] > self.cfg.nt_max:
            raise ValueError("data shape %s: expected (%d, nt) with nt <= the recorded steps" % (d.shape, self.nrcv))
        per = np.zeros(self.nrcv)
        tot = ctypes.c_double()
        self._chk(lib().tw_data_misfit(self.h, _d(d), d.shape[1], _d(per), ctypes.byref(tot)))
        return per, tot.value

    def profile_begin(self, max_steps):
        """Time the core launches of the next max_steps steps (HIP events on
        the solver's stream)."""
        self._chk(lib().tw_profile_begin(self.h, int(max_steps)))

    def profile_end(self):
        """(ms of the velocity core launches, ms of the stress core launches,
        profiled steps), summed over the profiled steps."""
        v, st, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int32()
        self._chk(lib().tw_profile_end(self.h, ctypes.byref(v), ctypes.byref(st), ctypes.byref(n)))
        return v.value, st.value, n.value


class Batch:
    """Independent solvers advanced together (tw_batch_*): one launch per half
    step for all of them, the time loop captured in a hipGraph.  Keeps
    references to the solvers; batched steps record no energy or receivers."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (_vp * len(self.solvers))(*[s.h for s in self.solvers])
        self.h = _vp()
        e = lib().tw_batch_create(arr, len(self.solvers), ctypes.byref(self.h))
        if e:
            raise TwError(e, lib().tw_global_error().decode())

    def _chk(self, e):
        if e:
            raise TwError(e, lib().tw_batch_last_error(self.h).decode())

    def step(self, n):
        self._chk(lib().tw_batch_step(self.h, int(n)))

    def elapsed_ms(self):
        """Device time of the last step() graph launch (waits for it)."""
        ms = ctypes.c_double()
        self._chk(lib().tw_batch_elapsed(self.h, ctypes.byref(ms)))
        return ms.value

    def set_rows(self, rows):
        self._chk(lib().tw_batch_set_rows(self.h, int(rows)))

    def autotune(self, trial_steps=20):
        best = ctypes.c_int32()
        self._chk(lib().tw_batch_autotune(self.h, int(trial_steps), ctypes.byref(best)))
        return best.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().tw_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
