"""GPU FP64 path vs the oracle (serial host loop of THOR's kernels).

Two builds are compared:
  * parity build ("pm"): GPU kernels and oracle share the portable elementary
    functions of portable_math.h, so the restated arithmetic must agree to the
    last bit (tolerance 0 up to a few ulp of accumulated difference: 1e-14
    relative);
  * production build: the GPU uses ocml's pow/log/sin/cos, the oracle glibc's;
    the step must match to 1e-12 relative per field (max |gpu - oracle| /
    max |oracle|) for Mh, Rho and pressure. W and Wh, whose values are the
    small residual of hydrostatic balance, carry the libm ulp differences
    amplified by that cancellation; their bound is 1e-11.
"""
import numpy as np
import pytest

from tests.common import STATE, earth_params, grid, rel_err

pytestmark = pytest.mark.gpu
TOL_STEP = 1e-12
TOL_PM = 1e-14


def _pair(capi, glevel, nv, variant="", ns=6, dt=1000.0, tnoise=1e-3, wnoise=0.0, seed=1000, **kw):
    from oracle.oracle import Oracle
    g = grid(capi, glevel, nv)
    p = earth_params(capi, glevel, dt=dt, ns=ns, **kw)
    st = capi.ic_rest(g, p, seed=seed, temperature_noise=tnoise, wind_noise=wnoise)
    d = capi.Dyn(g, p, variant=variant)
    d.set_state(st)
    o = Oracle(g, p, variant=variant)
    o.set_state(st)
    return g, p, st, d, o


TOL_VERTICAL = 1e-11


def _compare(d, o, tol, label, tol_vertical=None):
    gs = d.get_state()
    os_ = o.get_state()
    errs = {k: rel_err(gs[k], os_[k]) for k in STATE}
    print(label, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        t = tol_vertical if (tol_vertical is not None and k in ("W", "Wh")) else tol
        assert v <= t, f"{label}: {k} relative error {v:.3e} > {t}"
    return errs


def test_pm_thor_step_bitwise_g4(built):
    print("case: parity build, g4 L32, one Thor step, HS IC + 1e-3 K noise", flush=True)
    g, p, st, d, o = _pair(built, 4, 32, variant="pm")
    d.thor()
    d.sync()
    o.thor()
    _compare(d, o, TOL_PM, "pm thor g4")


def test_pm_profx_thor_steps_g4(built):
    print("case: parity build, g4 L32, 3 x (ProfX HS + Thor)", flush=True)
    g, p, st, d, o = _pair(built, 4, 32, variant="pm")
    for i in range(3):
        d.run(1)
        d.sync()
        o.steps(1)
        _compare(d, o, TOL_PM, f"pm step {i + 1}")


def test_native_step_with_winds_g4(built):
    print("case: production build, g4 L32, one step from a state with 1 m/s wind noise", flush=True)
    g, p, st, d, o = _pair(built, 4, 32, wnoise=1.0)
    d.run(1)
    d.sync()
    o.steps(1)
    _compare(d, o, TOL_STEP, "native wind g4", TOL_VERTICAL)


@pytest.mark.parametrize("order", [6, 8])
def test_pm_higher_order_hyperdiffusion_g4(built, order):
    """HyDiffOrder / 2 - 2 intermediate Laplacian passes between the first and
    the last (each reading the previous pass's whole output)."""
    print(f"case: parity build, g4 L20, HyDiffOrder {order}, one Thor step with 1 m/s winds", flush=True)
    g, p, st, d, o = _pair(built, 4, 20, variant="pm", wnoise=1.0, HyDiffOrder=order)
    d.thor()
    d.sync()
    o.thor()
    _compare(d, o, TOL_PM, f"pm HyDiffOrder {order}")


def test_native_order_six_hyperdiffusion_g4(built):
    print("case: production build, g4 L20, HyDiffOrder 6, 2 x (ProfX HS + Thor) with 1 m/s winds", flush=True)
    g, p, st, d, o = _pair(built, 4, 20, wnoise=1.0, HyDiffOrder=6)
    d.run(2)
    d.sync()
    o.steps(2)
    _compare(d, o, TOL_STEP, "native HyDiffOrder 6", TOL_VERTICAL)


@pytest.mark.parametrize("deep", [1, 0])
def test_pm_vertical_hyperdiffusion_g4(built, deep):
    """VertHyDiff (order 6) on top of the horizontal hyperdiffusion: the
    vertical terms enter the slow modes and the pressure update."""
    print(f"case: parity build, g4 L20, VertHyDiff order 6, DeepModel {deep}, 2 x (ProfX HS + Thor)", flush=True)
    g, p, st, d, o = _pair(built, 4, 20, variant="pm", wnoise=1.0, DeepModel=deep, VertHyDiff=1,
                           VertHyDiffOrder=6, Diffc_v=2e-3)
    for i in range(2):
        d.run(1)
        d.sync()
        o.steps(1)
        _compare(d, o, TOL_PM, f"pm VertHyDiff step {i + 1}")


def test_native_vertical_hyperdiffusion_g4(built):
    print("case: production build, g4 L20, VertHyDiff order 6, 2 x (ProfX HS + Thor)", flush=True)
    g, p, st, d, o = _pair(built, 4, 20, wnoise=1.0, VertHyDiff=1, VertHyDiffOrder=6, Diffc_v=2e-3)
    d.run(2)
    d.sync()
    o.steps(2)
    _compare(d, o, TOL_STEP, "native VertHyDiff", TOL_VERTICAL)


@pytest.mark.parametrize("order,sponge_order", [(4, 2), (4, 4), (6, 2), (6, 4)])
def test_pm_diffusive_sponge_g4(built, order, sponge_order):
    """The diffusive sponge from a third of the model top: added to the last
    hyperdiffusion pass (order_diff_sponge below HyDiffOrder, reading the
    right pass's output) or folded into Kdh4 (equal orders)."""
    print(f"case: parity build, g4 L20, HyDiffOrder {order}, order_diff_sponge {sponge_order}, "
          f"2 x (ProfX HS + Thor)", flush=True)
    g, p, st, d, o = _pair(built, 4, 20, variant="pm", wnoise=1.0, HyDiffOrder=order, DiffSponge=1,
                           order_diff_sponge=sponge_order, Dv_sponge=0.05, ns_diff_sponge=0.33)
    for i in range(2):
        d.run(1)
        d.sync()
        o.steps(1)
        _compare(d, o, TOL_PM, f"pm sponge step {i + 1}")
