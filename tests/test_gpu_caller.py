"""The drop-in boundary from a compiled C++ caller (tests/c/esp_loop.cpp).

The program links against the library through include/thor_dyn.h only, owns
the state and THOR's forcing arrays in device memory in THOR's layout, and
calls thor_dyn_step once per time step on those arrays in place. Its result
after the loop is compared with the oracle run from the same initial state
with the same forcing: bit for bit with the parity build, within the usual
libm distance with the production build; the mixed form within its usual
distance of FP64.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from common import CACHE, STATE, earth_params, grid, max_rel

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
NSTEPS = 3


def _read(path):
    out = []
    with open(path, "rb") as fh:
        while True:
            n = np.fromfile(fh, dtype=np.uint64, count=1)
            if n.size == 0:
                break
            out.append(np.fromfile(fh, dtype=np.float64, count=int(n[0])))
    names = ("Qheat", "dMh", "dWh", *STATE)
    assert len(out) == len(names)
    return dict(zip(names, out))


def _run(exe, precision, tmp_path):
    path = os.path.join(str(tmp_path), f"{os.path.basename(exe)}_{precision}.bin")
    exe = os.path.join(HERE, "c", exe)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (make -C tests/c)")
    os.makedirs(CACHE, exist_ok=True)
    r = subprocess.run([exe, CACHE, str(precision), str(NSTEPS), path], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    return _read(path)


def _oracle(capi, variant, forcing):
    from oracle.oracle import Oracle
    g = grid(capi, 4, 20)
    p = earth_params(capi, 4, dt=600.0, ns=6)
    st = capi.ic_rest(g, p, seed=1016, temperature_noise=1e-3, wind_noise=1.0)
    o = Oracle(g, p, variant=variant)
    o.set_state(st)
    o.set("profx_Qheat", forcing["Qheat"])
    o.set("profx_dMh", forcing["dMh"])
    o.set("profx_dWh", forcing["dWh"])
    for _ in range(NSTEPS):
        o.thor()
    return {k: o.get(k) for k in STATE}


@pytest.mark.parametrize("exe,variant,tol", [("esp_loop_pm", "pm", 0.0), ("esp_loop", "", 1e-11)])
def test_compiled_caller_fp64_matches_oracle(built, tmp_path, exe, variant, tol):
    print(f"case: {exe}: g4 L20 FP64, {NSTEPS} x thor_dyn_step on caller-owned device arrays with an "
          f"external forcing, vs the oracle", flush=True)
    out = _run(exe, 0, tmp_path)
    want = _oracle(built, variant, out)
    errs = {k: max_rel(out[k], want[k]) for k in STATE}
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= tol


def test_compiled_caller_mixed_stays_close(built, tmp_path):
    print(f"case: esp_loop: g4 L20 mixed precision, {NSTEPS} x thor_dyn_step, vs the FP64 oracle", flush=True)
    out = _run("esp_loop", 1, tmp_path)
    want = _oracle(built, "", out)
    for k, t in {"Mh": 1e-4, "Wh": 1e-3, "W": 1e-3, "Rho": 1e-6, "pressure": 1e-6}.items():
        err = max_rel(out[k], want[k])
        print(f"  mixed vs FP64 {k}: {err:.2e}", flush=True)
        assert err < t, k
