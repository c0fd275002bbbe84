"""Case builders shared by the parity tests: identical seeded inputs for the
HIP solver (twave) and the CPU oracle."""
import numpy as np

import oracle as O
import twave as T

PREC = {"fp64": 0, "fp32": 1, "bf16": 2, "fp16": 3}


def cfg_kwargs(physics="elastic", mx=40, my=48, bc="FFPP", precision="fp32", cpst="0", coef16=0,
               nt=40, dx=1.0 / 125, num_dx=1, den_dx=125, dt=None, stencil_form=0, energy_form=0,
               source_scaling=1.0):
    ph = 0 if physics == "elastic" else 1
    if dt is None:
        dt = 0.5 * dx / (np.sqrt(2.0) * 2.0)  # CFL 0.5 at vmax 2 (cuda/src/cpp/namespace_input.hpp CFL form)
    return dict(physics=ph, mx=mx, my=my, dx=dx, num_dx=num_dx, den_dx=den_dx, dt=dt, bc=bc,
                precision=PREC[precision], coef16=coef16, cpst=cpst, stencil_form=stencil_form,
                energy_form=energy_form, record_energy=1, nt_max=nt, source_scaling=source_scaling)


def make_pair(**kw):
    ck = cfg_kwargs(**kw)
    return T.Solver(T.config(**ck)), O.Oracle(O.make_config(**ck))


def random_material(solver, orc, seed, physics="elastic"):
    rng = np.random.default_rng(seed)
    if physics == "elastic":
        ks = [T.C_BINV_VX, T.C_BINV_VY, T.C_MU, T.C_LAMBDA, T.C_TWOMU]
    else:
        ks = [T.C_BINV_VX, T.C_BINV_VY, T.C_KAPPA]
    for k in ks:
        shape = solver.get_coef(k).shape
        if k in (T.C_BINV_VX, T.C_BINV_VY):
            v = rng.uniform(0.7, 1.3, shape)
        elif k == T.C_LAMBDA:
            v = rng.uniform(1.5, 2.5, shape)
        else:
            v = rng.uniform(0.8, 1.2, shape) * (2.0 if k == T.C_TWOMU else 1.0)
        solver.set_coef(k, v)
        orc.set_coef(k, v)


def random_fields(solver, orc, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    for f in solver.fields:
        v = amp * rng.standard_normal(solver.shape(f))
        solver.set_field(f, v)
        orc.set_field(f, v)


def receivers_all(solver, orc, seed, n_per_field=3):
    rng = np.random.default_rng(seed)
    fs, xs, ys = [], [], []
    for f in solver.fields:
        r, c = solver.shape(f)
        for _ in range(n_per_field):
            fs.append(f)
            xs.append(int(rng.integers(0, r)))
            ys.append(int(rng.integers(0, c)))
    solver.set_receivers(fs, xs, ys)
    orc.set_receivers(fs, xs, ys)
    return fs, xs, ys


def compare_state(solver, orc, exact=True, rtol=0.0, companions=False):
    """State arrays of every field; with companions=True also the
    compensation companion (R or C array, the K residual, the K8 byte codes;
    zeros on both sides where the variant keeps none)."""
    for f in solver.fields:
        pairs = [("field", solver.get_field(f), orc.get_field(f))]
        if companions:
            pairs.append(("companion", solver.get_companion(f), orc.get_companion(f)))
        for what, a, b in pairs:
            msg = "%s %s" % (what, T.FIELD_NAMES[f])
            if exact:
                np.testing.assert_array_equal(a, b, err_msg=msg)
            else:
                np.testing.assert_allclose(a, b, rtol=rtol, atol=0, err_msg=msg)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)
