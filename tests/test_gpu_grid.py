"""Spring dynamics of the grid generator on the GPU against the host loop.

Both run the iteration of Icogrid::spring_dynamics (grid.cu:1670-1815) from
the same lattice; the GPU kernel uses the device's acos and sqrt, so its
positions may differ from the host loop's in the last bits, and a force
maximum sitting on the stopping threshold could move the converging
iteration by one. The finished grids (after the centroid relocation) must
agree to round-off.
"""
from __future__ import annotations

import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _build(capi, glevel, host):
    old = os.environ.get("THOR_GRID_HOST")
    os.environ["THOR_GRID_HOST"] = "1" if host else "0"
    try:
        t0 = time.perf_counter()
        g = capi.Grid(glevel, 4, 1.0, 1.0, spring=True, beta=1.15, cache_dir=None)
        return g, time.perf_counter() - t0
    finally:
        if old is None:
            del os.environ["THOR_GRID_HOST"]
        else:
            os.environ["THOR_GRID_HOST"] = old


@pytest.mark.parametrize("glevel", [4, 5])
def test_spring_dynamics_device_matches_host(built, glevel):
    capi = built
    gh, th = _build(capi, glevel, host=True)
    gd, td = _build(capi, glevel, host=False)
    print(f"g{glevel}: host {gh.spring_iterations} iterations {th:.2f} s, "
          f"device {gd.spring_iterations} iterations {td:.2f} s", flush=True)
    assert gh.spring_iterations > 100
    assert abs(gd.spring_iterations - gh.spring_iterations) <= 1
    assert np.abs(gd.xyz - gh.xyz).max() < 1e-11
    assert np.abs(gd.areasT - gh.areasT).max() < 1e-10 * gh.areasT.max()
