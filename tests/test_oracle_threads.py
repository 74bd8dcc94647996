"""The oracle's block and column loops on several host threads give the same
bits as the serial loop (each output element is written by one iteration,
with the same arithmetic). Covers the step with the optional paths that have
loops of their own: HyDiffOrder 6, VertHyDiff, the Rayleigh sponge, the
double-gray RT and the energy equation."""
from __future__ import annotations

import numpy as np
import pytest

from common import STATE, earth_params, grid, hj_grid, rt_params


def _run(capi, g, p, st, threads, forcing=None, rt=None, sponge=None):
    from oracle import oracle as orc
    orc.set_threads(threads, "pm")
    o = orc.Oracle(g, p, variant="pm")
    o.set_state(st)
    if rt is not None:
        o.set_rt(rt)
    if sponge is not None:
        o.set_sponge(sponge)
    if forcing is None:
        o.steps(2)
    else:
        o.run(2, forcing)
    out = o.get_state()
    out["diag"] = o.globdiag()
    return out


@pytest.mark.parametrize("case", ["hs_switches", "energy", "rt_sponge"])
def test_threads_bitwise(built, case):
    import os
    from oracle import oracle as orc
    from thor_amd import surrogate as sg
    capi = built
    nthr = max(2, min(8, os.cpu_count() or 2))
    if case == "rt_sponge":
        g = hj_grid(capi, 4, 20)
        p = sg.hot_jupiter_params(4, dt=300.0, ns=6)
        st = capi.ic_rest(g, p, seed=7, temperature_noise=1.0, wind_noise=0.1)
        kw = dict(forcing=capi.FORCING_DOUBLE_GRAY, rt=rt_params(capi), sponge=sg.hot_jupiter_sponge())
    else:
        g = grid(capi, 4, 20)
        extra = dict(HyDiffOrder=6, VertHyDiff=1) if case == "hs_switches" else dict(thermo_equation=1)
        p = earth_params(capi, 4, dt=600.0, ns=6, **extra)
        st = capi.ic_rest(g, p, seed=5, temperature_noise=1e-3, wind_noise=1.0)
        kw = {}
    try:
        a = _run(capi, g, p, st, 1, **kw)
        b = _run(capi, g, p, st, nthr, **kw)
    finally:
        orc.set_threads(0, "pm")
    for k in STATE + ("diag",):
        assert np.array_equal(a[k], b[k]), (case, k)
