"""The tracer module (THOR's chemistry tracers, thor_chemistry.h, driven as
chemistry.cu:333-502 drives them) on the GPU against the oracle: a consumer
of the physics hooks and the RK-array registry end to end.

Two tracers (a smooth and a noisy mixing ratio) over two ProfX + Thor steps
with 1 m/s winds: the parity build equals the oracle bit for bit in the
state and the tracers, with and without hyperdiffusion and in the shallow
model; the production build stays within 1e-12; the mixed form stays close.
"""
from __future__ import annotations

import numpy as np
import pytest

from common import STATE, earth_params, grid, max_rel, momentum_scale_errs

pytestmark = pytest.mark.gpu


def _tracers(g, st, seed):
    rng = np.random.default_rng(seed)
    lat = np.asarray(g.lonlat).reshape(g.P, 2)[:, 1]
    q0 = 1.0 + 0.3 * np.sin(lat)[:, None] * np.ones((1, g.nv))
    q1 = 1.0 + 0.1 * rng.uniform(-1, 1, (g.P, g.nv))
    return np.stack([st["Rho"] * q0, st["Rho"] * q1], axis=2)


@pytest.mark.parametrize("variant,deep,hydiff,tol", [("pm", 1, 1, 0.0), ("pm", 0, 1, 0.0), ("pm", 1, 0, 0.0),
                                                     ("", 1, 1, 1e-12)])
def test_tracers_through_hooks_match_oracle(built, variant, deep, hydiff, tol):
    from oracle.oracle import Oracle
    capi = built
    print(f"case: tracers x2, g4 L20, build {variant or 'production'}, DeepModel {deep}, HyDiff {hydiff}, 2 steps",
          flush=True)
    g = grid(capi, 4, 20)
    p = earth_params(capi, 4, dt=600.0, ns=6, DeepModel=deep, HyDiff=hydiff)
    st = capi.ic_rest(g, p, seed=1017, temperature_noise=1e-3, wind_noise=1.0)
    tr0 = _tracers(g, st, 3)
    d = capi.Dyn(g, p, variant=variant)
    d.set_state(st)
    t = capi.Tracers(d, 2)
    t.set(tr0)
    o = Oracle(g, p, variant=variant)
    o.set_state(st)
    o.set_tracers(2)
    o.set("tracer", tr0)
    for step in (1, 2):
        d.run(1)
        d.sync()
        o.steps(1)
        gs, os_ = d.get_state(), o.get_state()
        e = {k: max_rel(v, os_[k]) for k, v in gs.items()}
        e["tracer"] = max_rel(t.get().ravel(), o.get("tracer"))
        em = momentum_scale_errs(gs, os_, p)
        print(f"step {step}", {k: f"{v:.2e}" for k, v in e.items()},
              "momenta vs rho c_s", {k: f"{v:.2e}" for k, v in em.items()}, flush=True)
        if tol == 0.0:
            assert max(e.values()) == 0.0
        else:
            # production build: libm ulps; momenta against the momentum scale
            assert max(e[k] for k in ("Rho", "pressure", "tracer")) <= tol
            assert max(em.values()) <= tol
    # the tracers moved (advection and diffusion did something)
    assert max_rel(t.get(), tr0) > 1e-9
    t.close()
    d.close()


def test_mixed_tracers_stay_close(built):
    from oracle.oracle import Oracle
    capi = built
    print("case: tracers x2, g4 L20, mixed form vs the FP64 oracle, 2 steps", flush=True)
    g = grid(capi, 4, 20)
    p = earth_params(capi, 4, dt=600.0, ns=6, precision=capi.PRECISION_MIXED)
    st = capi.ic_rest(g, p, seed=1018, temperature_noise=1e-3, wind_noise=1.0)
    tr0 = _tracers(g, st, 4)
    d = capi.Dyn(g, p)
    d.set_state(st)
    t = capi.Tracers(d, 2)
    t.set(tr0)
    p64 = capi.Params.from_buffer_copy(p)
    p64.precision = capi.PRECISION_FP64
    o = Oracle(g, p64)
    o.set_state(st)
    o.set_tracers(2)
    o.set("tracer", tr0)
    d.run(2)
    d.sync()
    o.steps(2)
    e = max_rel(t.get().ravel(), o.get("tracer"))
    print(f"tracer mixed vs FP64: {e:.2e}", flush=True)
    assert e < 1e-6
    t.close()
    d.close()
