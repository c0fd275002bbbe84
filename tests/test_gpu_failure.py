"""Failure detection forced from the host (GPU): a non-finite value injected
mid-run is reported by the finiteness check and shows up in the energy
record from that step on; a time step above the stability limit makes the
recorded energy grow without bound until the check trips; every kernel still
runs to completion (no trap, no abort), and the solver keeps answering."""
import numpy as np
import pytest

import common as C
import twave as T

pytestmark = pytest.mark.gpu


def solver(dt_scale=1.0, nt=80, precision="fp32", physics="elastic"):
    kw = C.cfg_kwargs(physics=physics, mx=40, my=48, bc="FCCC", precision=precision, nt=nt,
                      stencil_form=1 if physics == "acoustic" else 0)
    kw["dt"] *= dt_scale
    s = T.Solver(T.config(**kw))
    rng = np.random.default_rng(7)
    for f in s.fields:
        s.set_field(f, 1e-3 * rng.standard_normal(s.shape(f)))
    return s


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_nan_injected_mid_run_is_detected(precision):
    print("case failure nan", precision, flush=True)
    s = solver(precision=precision)
    s.step(20)
    assert s.check_finite()
    v = s.get_field(T.SXX)
    v[20, 24] = np.nan
    s.set_field(T.SXX, v)
    s.step(5)
    assert not s.check_finite()
    ek, ep0, ep1 = s.energy()
    assert np.all(np.isfinite(ek[:20])) and np.all(np.isfinite(ep0[:20]))
    assert not np.isfinite(ek[20:25]).all()
    s.step(3)  # keeps stepping; nothing aborts
    assert s.current_step() == 28
    s.close()


@pytest.mark.parametrize("physics", ["elastic", "acoustic"])
def test_unstable_time_step_is_caught(physics):
    print("case failure cfl", physics, flush=True)
    stable = solver(1.0, nt=300, physics=physics)
    stable.step(300)
    e_ok = np.array(stable.energy()[0])
    assert stable.check_finite() and np.all(np.isfinite(e_ok))
    stable.close()
    s = solver(3.0, nt=300, physics=physics)  # CFL 1.5: above the limit of the scheme
    s.step(300)
    ek = np.array(s.energy()[0])
    grown = np.nanmax(np.where(np.isfinite(ek), ek, np.nan)) / ek[1]
    assert (not s.check_finite()) or grown > 1e6, grown
    assert grown > 1e3
    s.close()


def test_bad_material_rejected():
    """Material input is validated on the host before anything reaches the
    device: non-finite coefficients, non-positive density or velocity."""
    kw = C.cfg_kwargs(physics="elastic", mx=32, my=40, bc="FCCC", precision="bf16", cpst="K8", coef16=2, nt=4)
    s = T.Solver(T.config(**kw))
    v = s.get_coef(T.C_LAMBDA)
    v[3, 5] = np.nan
    with pytest.raises(T.TwError):
        s.set_coef(T.C_LAMBDA, v)
    n = (33, 41)
    rho, vp, vs = np.ones(n), np.full(n, 2.0), np.ones(n)
    rho[0, 0] = -1.0
    with pytest.raises(T.TwError):
        s.set_model(rho, vp, vs, 1, 125)
    rho[0, 0] = 1.0
    vs[2, 2] = np.inf
    with pytest.raises(T.TwError):
        s.set_model(rho, vp, vs, 1, 125)
    vs[2, 2] = 1.0
    s.set_model(rho, vp, vs, 1, 125)
    s.step(4)
    assert s.check_finite()
    s.close()
