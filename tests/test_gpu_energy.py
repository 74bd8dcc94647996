"""The total-energy form of the fast modes (thermo_equation = energy,
esp.cu:696-710) on the GPU against the oracle: Compute_Temperature_H_Pt_Geff's
energies (thor_auxiliary.h:143-161, 212-230, 281-287), the internal-energy
hyperdiffusion field (thor_diff.h:196-202), the slow pressure tendency
(thor_slowmodes.h:375-381) and the total-energy density/pressure update
(thor_fastmodes.h:352-533), inside whole Held-Suarez steps.

  * parity build: bit for bit after every step, the total energy of the last
    acoustic substep included, over the switches that change the energies
    (DeepModel, GravHeightVar) and the solve (NonHydro), with the vertical
    hyperdiffusion on in one case;
  * production build: per-step error from a common start below 1e-12
    (W and Wh 1e-11, as for the entropy form);
The mixed-precision form has no energy variant (its creation fails with
THOR_ERR_UNSUPPORTED; tests/test_capi.py).
"""
from __future__ import annotations

import numpy as np
import pytest

from common import STATE, earth_params, grid, max_rel

pytestmark = pytest.mark.gpu


def _setup(capi, variant, **kw):
    from oracle.oracle import Oracle
    g = grid(capi, 4, 20)
    p = earth_params(capi, 4, dt=600.0, ns=6, thermo_equation=capi.THERMO_ENERGY, **kw)
    st = capi.ic_rest(g, p, seed=1017, temperature_noise=1e-3, wind_noise=1.0)
    d = capi.Dyn(g, p, variant=variant)
    o = Oracle(g, p, variant=variant)
    for x in (d, o):
        x.set_state(st)
    return g, p, d, o


@pytest.mark.parametrize("deep,nonhydro,ghv,vhd", [(1, 1, 0, 0), (0, 1, 0, 0), (1, 0, 1, 0), (1, 1, 0, 1)])
def test_energy_parity_build_bitwise(built, deep, nonhydro, ghv, vhd):
    capi = built
    print(f"case: energy equation, parity build g4 L20, DeepModel {deep}, NonHydro {nonhydro}, "
          f"GravHeightVar {ghv}, VertHyDiff {vhd}, 3 steps", flush=True)
    extra = dict(VertHyDiff=1, VertHyDiffOrder=6, Diffc_v=2e-3) if vhd else {}
    g, p, d, o = _setup(capi, "pm", DeepModel=deep, NonHydro=nonhydro, GravHeightVar=ghv, **extra)
    for step in (1, 2, 3):
        d.run(1)
        d.sync()
        o.steps(1)
        gs, os_ = d.get_state(), o.get_state()
        e = {k: max_rel(gs[k], os_[k]) for k in STATE}
        et = max_rel(d.get_field("Etotal").ravel(), o.get("Etotal"))
        print(f"step {step}", {k: f"{v:.2e}" for k, v in e.items()}, f"Etotal {et:.2e}", flush=True)
        assert max(e.values()) == 0.0 and et == 0.0
    d.close()


def test_energy_production_per_step(built):
    capi = built
    print("case: energy equation, production build g4 L20, 3 steps from common starts", flush=True)
    g, p, d, o = _setup(capi, "")
    for step in (1, 2, 3):
        d.set_state(o.get_state())
        d.run(1)
        d.sync()
        o.steps(1)
        gs, os_ = d.get_state(), o.get_state()
        e = {k: max_rel(gs[k], os_[k]) for k in STATE}
        print(f"step {step}", {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
        for k, v in e.items():
            assert v <= (1e-11 if k in ("W", "Wh") else 1e-12), (k, v)
    d.close()


def test_energy_and_entropy_forms_differ(built):
    """The switch reaches the step: one step of each form from the same state
    gives different pressures (the forms discretise different equations),
    both close to the same physical state (within 1e-3 of the pressure)."""
    capi = built
    print("case: energy vs entropy equation, parity build g4 L20, 1 step", flush=True)
    g, p, d, o = _setup(capi, "pm")
    st0 = o.get_state()
    pe = earth_params(capi, 4, dt=600.0, ns=6)
    de = capi.Dyn(g, pe, variant="pm")
    de.set_state(st0)
    for x in (d, de):
        x.run(1)
        x.sync()
    a, b = d.get_state()["pressure"], de.get_state()["pressure"]
    diff = float(np.abs(a - b).max())
    step = float(np.abs(b - st0["pressure"]).max())
    print(f"max |p_energy - p_entropy| {diff:.3e} Pa, step change {step:.3e} Pa", flush=True)
    assert diff > 0.0 and diff < 1e-3 * float(np.abs(b).max())
    d.close()
    de.close()
