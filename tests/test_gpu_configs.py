"""The GPU path against the oracle at each workload's own sizes.

Workloads (grids, physics, seeds as the benchmark configurations define them):
  0  Held-Suarez Earth g4 L32, dt 1000 s, 6 substeps, rest + 1e-3 K noise,
     100 steps of ProfX + Thor (serial host loop as the checker);
  1  hot-Jupiter Newtonian relaxation g5 L40, dt 300 s, +-0.1 m/s wind noise;
  2  Held-Suarez Earth g6 L48, dt 300 s, 8 substeps, rest + 1e-3 K noise;
  3  hot Jupiter with double-gray RT, g5 L40 (the surrogate's own test at
     L40 is in test_gpu_mlp.py);
  4  hot Jupiter g7 L48, dt 120 s, double-gray RT and Rayleigh sponge.

The parity build (portable elementary functions on both sides) must equal
the oracle bit for bit at every step checked. The production build differs
from the glibc oracle only through libm ulps; its per-step error is measured
from a common start state (the oracle's state before each step), so errors
do not accumulate across steps.
"""
from __future__ import annotations

import numpy as np
import pytest

from common import (CACHE, STATE, earth_params, grid, hip_runtime, hj_grid, hj_params, max_rel,
                    momentum_scale_errs, nr_params, rt_params, HIP_D2H, HIP_H2D)

pytestmark = pytest.mark.gpu

# per-step tolerance of the production FP64 build against the glibc oracle
TOL_STEP = 1e-12


def _errs(gs, os_):
    return {k: max_rel(gs[k], os_[k]) for k in STATE}


def _fmt(errs):
    return {k: f"{v:.2e}" for k, v in errs.items()}


def _momentum_scale_errs(gs, os_, p):
    return momentum_scale_errs(gs, os_, p)


def test_cfg0_hundred_steps_parity_build_bitwise_with_globdiag(built):
    """Config 0, 100 steps of the parity build against the oracle: bit for
    bit after every step; mass, total energy and angular momentum (x, y, z)
    from the GPU's reduction equal the oracle's at steps 0, 50 and 100, and
    their drift over the run is reported (mass to round-off)."""
    from oracle.oracle import Oracle
    capi = built
    print("case: config 0 (HS g4 L32, dt 1000 s, ns 6, seed 1000), parity build, 100 steps", flush=True)
    g = grid(capi, 4, 32)
    p = earth_params(capi, 4, dt=1000.0, ns=6)
    st = capi.ic_rest(g, p, seed=1000, temperature_noise=1e-3)
    d = capi.Dyn(g, p, variant="pm")
    o = Oracle(g, p, variant="pm")
    d.set_state(st)
    o.set_state(st)
    diag = {}
    worst = {k: 0.0 for k in STATE}
    for step in range(101):
        if step in (0, 50, 100):
            gd, od = d.globdiag(), o.globdiag()
            diag[step] = gd
            print(f"step {step}: mass {gd[0]:.17e} energy {gd[1]:.17e} AM {gd[2]:.6e} {gd[3]:.6e} {gd[4]:.17e}",
                  flush=True)
            assert np.array_equal(gd, od), (step, gd, od)
        if step == 100:
            break
        d.run(1)
        d.sync()
        o.steps(1)
        e = _errs(d.get_state(), o.get_state())
        for k in STATE:
            worst[k] = max(worst[k], e[k])
        assert max(e.values()) == 0.0, (step + 1, _fmt(e))
    m0, e0, amz0 = diag[0][0], diag[0][1], diag[0][4]
    drift = {"mass": diag[100][0] / m0 - 1, "energy": diag[100][1] / e0 - 1, "AMz": diag[100][4] / amz0 - 1,
             "AMx/AMz": diag[100][2] / amz0, "AMy/AMz": diag[100][3] / amz0}
    print("100-step drift:", {k: f"{v:.3e}" for k, v in drift.items()}, "worst per-step diff", _fmt(worst),
          flush=True)
    # flux-form continuity: mass changes only by round-off (about 1e-15 per step)
    assert abs(drift["mass"]) < 1e-13
    d.close()


def test_cfg0_hundred_steps_production_per_step(built):
    """Config 0 with the production build: each of the 100 steps starts from
    the oracle's state before it (so errors do not carry over). Density and
    pressure within 1e-12 relative per step; the momenta within 1e-12 of the
    flow's momentum scale rho c_s (relative to their own maximum they are
    near-rest residuals, where a one-ulp pow difference shows up at 1e-10 or
    so). Free-running drift between the two over the 100 steps is reported."""
    from oracle.oracle import Oracle
    capi = built
    print("case: config 0 (HS g4 L32), production build, 100 steps from the oracle's states", flush=True)
    g = grid(capi, 4, 32)
    p = earth_params(capi, 4, dt=1000.0, ns=6)
    st = capi.ic_rest(g, p, seed=1000, temperature_noise=1e-3)
    d = capi.Dyn(g, p)
    free = capi.Dyn(g, p)
    o = Oracle(g, p)
    o.set_state(st)
    free.set_state(st)
    worst, worst_m = {k: 0.0 for k in STATE}, {"Mh": 0.0, "Wh": 0.0, "W": 0.0}
    for step in range(100):
        d.set_state(o.get_state())
        d.run(1)
        d.sync()
        o.steps(1)
        os_ = o.get_state()
        gs = d.get_state()
        e, em = _errs(gs, os_), _momentum_scale_errs(gs, os_, p)
        for k in STATE:
            worst[k] = max(worst[k], e[k])
        for k in em:
            worst_m[k] = max(worst_m[k], em[k])
        assert e["Rho"] <= TOL_STEP and e["pressure"] <= TOL_STEP, (step + 1, _fmt(e))
        assert max(em.values()) <= TOL_STEP, (step + 1, _fmt(em))
    free.run(100)
    free.sync()
    drift = _errs(free.get_state(), o.get_state())
    print("worst per-step relative", _fmt(worst), "vs rho c_s", _fmt(worst_m), flush=True)
    print("free-running after 100 steps", _fmt(drift), flush=True)
    d.close()
    free.close()


def _hj_cfg(capi, glevel, nv, dt, **kw):
    from thor_amd import surrogate as sg
    g = hj_grid(capi, glevel, nv)
    p = sg.hot_jupiter_params(glevel, dt=dt, ns=6, **kw)
    return g, p


def test_cfg1_newtonian_g5L40_bitwise(built):
    """Config 1: hot-Jupiter Newtonian relaxation at g5 L40, dt 300 s, 6
    substeps, +-0.1 m/s wind noise, two steps: the parity build equals the
    oracle; the production build within 1e-12 relative for density and
    pressure and within 1e-12 of the momentum scale rho c_s for the momenta."""
    from oracle.oracle import Oracle
    capi = built
    print("case: config 1 (hot Jupiter Newtonian g5 L40, dt 300 s, seed 1001), 2 steps, both builds", flush=True)
    g, p = _hj_cfg(capi, 5, 40, 300.0)
    st = capi.ic_rest(g, p, seed=1001, temperature_noise=0.0, wind_noise=0.1)
    for variant, tol in (("pm", 0.0), ("", TOL_STEP)):
        d = capi.Dyn(g, p, variant=variant)
        o = Oracle(g, p, variant=variant)
        for x in (d, o):
            x.set_state(st)
            x.set_newtonian(nr_params(capi))
            x.run(2, capi.FORCING_NEWTONIAN)
        d.sync()
        gs, os_ = d.get_state(), o.get_state()
        e = _errs(gs, os_)
        em = _momentum_scale_errs(gs, os_, p)
        print(f"config 1 [{variant or 'production'}]", _fmt(e), "momenta vs rho c_s", _fmt(em), flush=True)
        if tol == 0.0:
            assert max(e.values()) == 0.0
        else:
            assert e["Rho"] <= tol and e["pressure"] <= tol
            assert max(em.values()) <= tol
        d.close()


def test_cfg2_g6L48_parity_build_bitwise(built):
    """Config 2: Held-Suarez g6 L48, dt 300 s, 8 substeps, rest + 1e-3 K
    noise, one ProfX + Thor step. The parity build equals the oracle bit for
    bit; the production build's differences (1e-10 relative in Mh and Wh,
    whose maxima are near-rest residuals) therefore come from libm alone, and
    stay within 1e-12 of the momentum scale rho c_s."""
    from oracle.oracle import Oracle
    capi = built
    print("case: config 2 (HS g6 L48, dt 300 s, ns 8, seed 1002), 1 step, both builds", flush=True)
    g = grid(capi, 6, 48)
    p = earth_params(capi, 6, dt=300.0, ns=8)
    st = capi.ic_rest(g, p, seed=1002, temperature_noise=1e-3)
    for variant in ("pm", ""):
        d = capi.Dyn(g, p, variant=variant)
        o = Oracle(g, p, variant=variant)
        for x in (d, o):
            x.set_state(st)
        d.run(1)
        d.sync()
        o.steps(1)
        gs, os_ = d.get_state(), o.get_state()
        e, em = _errs(gs, os_), _momentum_scale_errs(gs, os_, p)
        print(f"config 2 [{variant or 'production'}] relative", _fmt(e), "vs rho c_s", _fmt(em), flush=True)
        if variant == "pm":
            assert max(e.values()) == 0.0
        else:
            assert e["Rho"] <= TOL_STEP and e["pressure"] <= TOL_STEP
            assert max(em.values()) <= TOL_STEP
        d.close()
        del o


def test_cfg3_double_gray_g5L40_bitwise(built):
    """Config 3's dynamics: hot Jupiter g5 L40 with the double-gray RT (kappa_sw
    4e-4, kappa_lw 1e-3, Tint 100 K, sigma (2000 K)^4): one ProfX alone (heating,
    insolation, ASR, OLR) and then two ProfX + Thor steps, parity build
    against the oracle, bit for bit."""
    from oracle.oracle import Oracle
    capi = built
    print("case: config 3 (hot Jupiter double-gray g5 L40, dt 300 s, seed 1003), ProfX then 2 steps", flush=True)
    g, p = _hj_cfg(capi, 5, 40, 300.0)
    st = capi.ic_rest(g, p, seed=1003, temperature_noise=1.0, wind_noise=0.1)
    d = capi.Dyn(g, p, variant="pm")
    o = Oracle(g, p, variant="pm")
    for x in (d, o):
        x.set_state(st)
        x.set_rt(rt_params(capi))
    d.profx(capi.FORCING_DOUBLE_GRAY)
    d.sync()
    o.profx_forcing(capi.FORCING_DOUBLE_GRAY)
    q_g, q_o = d.get_field("profx_Qheat"), o.get("profx_Qheat").reshape(g.P, g.nv)
    assert np.array_equal(q_g, q_o)
    for name in ("insol", "ASR", "OLR"):
        assert np.array_equal(d.rt_column(name), o.get(name)), name
    for x in (d, o):
        x.set_state(st)
    d.run(2, capi.FORCING_DOUBLE_GRAY)
    d.sync()
    o.run(2, capi.FORCING_DOUBLE_GRAY)
    e = _errs(d.get_state(), o.get_state())
    print("config 3 [pm]", _fmt(e), flush=True)
    assert max(e.values()) == 0.0
    d.close()


def test_cfg4_g7L48_rt_sponge_bitwise(built):
    """Config 4's full step at its own size: hot Jupiter g7 L48 (163 842
    columns), dt 120 s, 6 substeps, double-gray RT and the Rayleigh sponge
    (top quarter, 1e-4 1/s, 20 bands), +-0.1 m/s wind noise, one ProfX + Thor
    step, parity build against the oracle, bit for bit."""
    from oracle.oracle import Oracle
    from thor_amd import surrogate as sg
    capi = built
    print("case: config 4 (hot Jupiter g7 L48, dt 120 s, RT + sponge, seed 1004), 1 step, parity build", flush=True)
    g, p = _hj_cfg(capi, 7, 48, 120.0)
    st = capi.ic_rest(g, p, seed=1004, temperature_noise=0.0, wind_noise=0.1)
    d = capi.Dyn(g, p, variant="pm")
    o = Oracle(g, p, variant="pm")
    for x in (d, o):
        x.set_state(st)
        x.set_rt(rt_params(capi))
        x.set_sponge(sg.hot_jupiter_sponge())
    d.run(1, capi.FORCING_DOUBLE_GRAY)
    d.sync()
    o.run(1, capi.FORCING_DOUBLE_GRAY)
    e = _errs(d.get_state(), o.get_state())
    print("config 4 [pm]", _fmt(e), flush=True)
    assert max(e.values()) == 0.0
    d.close()


@pytest.mark.parametrize("deep,nonhydro,ghv", [(1, 0, 0), (1, 1, 1), (0, 1, 0), (0, 0, 0), (1, 0, 1)])
def test_switches_parity_build_bitwise(built, deep, nonhydro, ghv):
    """DeepModel, NonHydro and GravHeightVar away from their defaults (each
    changes the vertical solve's coefficients, the right-hand side or the
    radius factors): two steps of the parity build with 1 m/s winds equal the
    oracle bit for bit."""
    from oracle.oracle import Oracle
    capi = built
    print(f"case: parity build g4 L20, DeepModel {deep}, NonHydro {nonhydro}, GravHeightVar {ghv}, 2 steps",
          flush=True)
    g = grid(capi, 4, 20)
    p = earth_params(capi, 4, dt=600.0, ns=6, DeepModel=deep, NonHydro=nonhydro, GravHeightVar=ghv)
    st = capi.ic_rest(g, p, seed=1011, temperature_noise=1e-3, wind_noise=1.0)
    d = capi.Dyn(g, p, variant="pm")
    o = Oracle(g, p, variant="pm")
    for x in (d, o):
        x.set_state(st)
    for step in (1, 2):
        d.run(1)
        d.sync()
        o.steps(1)
        e = _errs(d.get_state(), o.get_state())
        print(f"step {step}", _fmt(e), flush=True)
        assert max(e.values()) == 0.0
    d.close()


def test_state_on_device_round_trip(built):
    """The boundary with the caller's state in device memory (on_device = 1,
    THOR's layout): set_state from device buffers, two steps, get_state into
    device buffers. Equal bit for bit to the same run through host buffers."""
    import ctypes as C
    capi = built
    print("case: FP64 g4 L20, state handed over in device buffers, 2 steps", flush=True)
    hip = hip_runtime()
    g = grid(capi, 4, 20)
    p = earth_params(capi, 4, dt=600.0, ns=6)
    st = capi.ic_rest(g, p, seed=1012, temperature_noise=1e-3, wind_noise=1.0)
    ref = capi.Dyn(g, p)
    ref.set_state(st)
    ref.run(2)
    ref.sync()
    want = ref.get_state()
    ref.close()
    bufs = {}
    try:
        for k in STATE:
            a = np.ascontiguousarray(st[k])
            ptr = C.c_void_p()
            assert hip.hipMalloc(C.byref(ptr), a.nbytes) == 0
            bufs[k] = (ptr.value, a.nbytes, a.shape)
            assert hip.hipMemcpy(ptr.value, a.ctypes.data, a.nbytes, HIP_H2D) == 0
        d = capi.Dyn(g, p)
        dp = capi._dp
        s = capi.State(1, *(C.cast(C.c_void_p(bufs[k][0]), dp) for k in STATE))
        d._chk(d.L.thor_dyn_set_state(d._h, C.byref(s)))
        # clear the caller's buffers: get_state must rewrite them completely
        for k in STATE:
            assert hip.hipMemsetAsync(bufs[k][0], 0xff, bufs[k][1], None) == 0
        assert hip.hipDeviceSynchronize() == 0
        d.run(2)
        d._chk(d.L.thor_dyn_get_state(d._h, C.byref(s)))
        d.sync()
        for k in STATE:
            out = np.empty(bufs[k][2])
            assert hip.hipMemcpy(out.ctypes.data, bufs[k][0], bufs[k][1], HIP_D2H) == 0
            assert np.array_equal(out, want[k]), k
        d.close()
    finally:
        for ptr, _, _ in bufs.values():
            hip.hipFree(ptr)
