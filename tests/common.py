"""Shared helpers of the test suite: grids, parameters and error norms."""
from __future__ import annotations

import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CACHE = os.path.join(ROOT, "tests", "_cache")

STATE = ("Mh", "Wh", "W", "Rho", "pressure")


def earth_params(capi, glevel, dt=1000.0, ns=6, **kw):
    """Held-Suarez Earth parameters (ifile/earth_hstest.thr)."""
    return capi.default_params(glevel=glevel, timestep=dt, n_substeps=ns, **kw)


def grid(capi, glevel, nv, top=36000.0, radius=6.371e6):
    os.makedirs(CACHE, exist_ok=True)
    return capi.Grid(glevel, nv, radius, top, spring=True, beta=1.15, cache_dir=CACHE)


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a - b| / max |b| (0 if both are zero)."""
    den = np.abs(b).max()
    num = np.abs(a - b).max()
    if den == 0:
        return float(num)
    return float(num / den)
