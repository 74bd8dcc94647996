"""The thor_esp driver on the GPU: THOR's command line, configuration file,
output files and restart (esp.cu:164-1470, esp_output.cu:136-486,
esp_initial.cu:1140-1283).

Earth Held-Suarez g4 L32 from rest (tests/golden/earth_hs_g4.thr, n_out 2):
  * a 4-step run writes the grid and planet files and outputs 0, 1, 2 with
    the steps and times THOR gives them; its state equals the library's own
    ProfX + Thor sequence bit for bit, and the global sums equal
    thor_dyn_globdiag's;
  * continuing from output 1 (-c) writes output 2 equal bit for bit to the
    uninterrupted run's.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from common import CACHE, earth_params, grid

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "thor_amd", "thor_esp")
CFG = os.path.join(ROOT, "tests", "golden", "earth_hs_g4.thr")


def _esp(*args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (python -m thor_amd.build)")
    os.makedirs(CACHE, exist_ok=True)
    r = subprocess.run([EXE, CFG, "--grid-cache", CACHE, *args], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    return r.stdout


def test_driver_run_restart_and_parity(built, tmp_path):
    from thor_amd import io
    print("case: thor_esp, Earth HS g4 L32 from rest, 4 steps (n_out 2), then -c from output 1", flush=True)
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    _esp("-o", a, "-N", "4")
    names = sorted(os.listdir(a))
    assert names == ["esp_output_Earth_0.h5", "esp_output_Earth_1.h5", "esp_output_Earth_2.h5",
                     "esp_output_grid_Earth.h5", "esp_output_planet_Earth.h5"], names
    out = lambda d, i: os.path.join(d, f"esp_output_Earth_{i}.h5")
    for i, (ns, t) in enumerate(((0, 0.0), (2, 2000.0), (4, 4000.0))):
        assert io.h5_read(out(a, i), "nstep")[0] == ns and io.h5_read(out(a, i), "simulation_time")[0] == t
        assert io.h5_is_int(out(a, i), "nstep")
    # the library's own sequence from the same rest state
    capi = built
    g = grid(capi, 4, 32)
    p = earth_params(capi, 4, dt=1000.0, ns=6)
    d = capi.Dyn(g, p)
    d.set_state(capi.ic_rest(g, p, seed=0, temperature_noise=0.0, wind_noise=0.0))
    d.run(4)
    d.sync()
    st = d.get_state()
    gd = d.globdiag()
    for name, key in (("Rho", "Rho"), ("Pressure", "pressure"), ("Mh", "Mh"), ("Wh", "Wh")):
        assert np.array_equal(io.h5_read(out(a, 2), name), st[key].ravel()), name
    for i, name in enumerate(("GlobalMass", "GlobalE", "GlobalAMx", "GlobalAMy", "GlobalAMz")):
        assert io.h5_read(out(a, 2), name)[0] == gd[i], name
    assert io.h5_attr(out(a, 2), "Mh", "units") == "kg m^-2 s^-1"
    grid_file = os.path.join(a, "esp_output_grid_Earth.h5")
    assert np.array_equal(io.h5_read(grid_file, "lonlat"), np.asarray(g.lonlat).ravel())
    assert io.h5_is_int(grid_file, "pntloc")
    d.close()
    # continue from output 1 (step 2): THOR reads the planet file beside it
    os.makedirs(b)
    _esp("-c", out(a, 1), "-o", b, "-N", "4")
    assert sorted(os.listdir(b)) == ["esp_output_Earth_2.h5"]
    for name in ("nstep", "simulation_time", "Rho", "Pressure", "Mh", "Wh", "GlobalMass", "GlobalE"):
        assert np.array_equal(io.h5_read(out(b, 2), name), io.h5_read(out(a, 2), name)), name
