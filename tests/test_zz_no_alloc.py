"""thor_dyn_globdiag_noalloc (device scratch, nothing allocated) against
thor_dyn_globdiag (the handle's own partials buffer and a host sum), through
the C ABI: two calls in a row, the second on a smaller handle, each equal bit
for bit to the allocating entry point; a handle above the scratch's limit is
refused with THOR_ERR_UNSUPPORTED."""
from __future__ import annotations

import numpy as np
import pytest

from common import earth_params, grid, hj_grid, hj_params, nr_params

pytestmark = pytest.mark.gpu


def _handle(capi, g, p, st, steps, forcing):
    d = capi.Dyn(g, p)
    d.set_state(st)
    if forcing == capi.FORCING_NEWTONIAN:
        d.set_newtonian(nr_params(capi))
    d.run(steps, forcing)
    d.sync()
    return d


def test_noalloc_globdiag_equals_globdiag_twice(built):
    capi = built
    print("case: globdiag_noalloc vs globdiag, hot Jupiter g5 L40 (2 steps), then Held-Suarez g4 L20 (1 step)",
          flush=True)
    g1 = hj_grid(capi, 5, 40)
    p1 = hj_params(capi, 5, dt=300.0, ns=6)
    st1 = capi.ic_rest(g1, p1, seed=1070, temperature_noise=0.0, wind_noise=0.1)
    d1 = _handle(capi, g1, p1, st1, 2, capi.FORCING_NEWTONIAN)
    want1 = d1.globdiag()
    got1 = d1.globdiag_noalloc()
    print("call 1 (g5 L40)", want1, got1, flush=True)
    assert np.array_equal(got1, want1)
    d1.close()

    g2 = grid(capi, 4, 20)
    p2 = earth_params(capi, 4, dt=600.0, ns=6)
    st2 = capi.ic_rest(g2, p2, seed=1071, temperature_noise=1e-3, wind_noise=1.0)
    d2 = _handle(capi, g2, p2, st2, 1, capi.FORCING_HELD_SUAREZ)
    want2 = d2.globdiag()
    got2 = d2.globdiag_noalloc()
    print("call 2 (g4 L20)", want2, got2, flush=True)
    assert np.array_equal(got2, want2)
    # the smaller call's totals are its own, not the first call's partials
    assert not np.array_equal(got2, got1)
    # repeated calls on one handle give the same bits
    assert np.array_equal(d2.globdiag_noalloc(), want2)
    d2.close()


def test_noalloc_globdiag_refuses_above_its_limit(built):
    capi = built
    print("case: globdiag_noalloc on g7 L56 (9 175 152 cells > 8 388 608): refused", flush=True)
    g = grid(capi, 7, 56)
    p = earth_params(capi, 7, dt=120.0, ns=6)
    st = capi.ic_rest(g, p, seed=1072, temperature_noise=1e-3)
    d = capi.Dyn(g, p)
    d.set_state(st)
    with pytest.raises(capi.ThorError) as e:
        d.globdiag_noalloc()
    assert "8388608" in str(e.value)
    assert np.isfinite(d.globdiag()).all()
    d.close()
