"""Shared helpers of the test suite: grids, parameters and error norms."""
from __future__ import annotations

import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# grid constructions shared with bench.py (g5-g7 take seconds to minutes to build)
CACHE = os.path.join(ROOT, ".thor_cache")

STATE = ("Mh", "Wh", "W", "Rho", "pressure")


def earth_params(capi, glevel, dt=1000.0, ns=6, **kw):
    """Held-Suarez Earth parameters (ifile/earth_hstest.thr)."""
    return capi.default_params(glevel=glevel, timestep=dt, n_substeps=ns, **kw)


def grid(capi, glevel, nv, top=36000.0, radius=6.371e6):
    os.makedirs(CACHE, exist_ok=True)
    return capi.Grid(glevel, nv, radius, top, spring=True, beta=1.15, cache_dir=CACHE)


def momentum_scale_errs(gs: dict, os_: dict, p) -> dict:
    """Errors of the momenta (Mh, Wh, W) relative to the flow's own momentum
    scale rho c_s (max rho times the isothermal sound speed at Tmean). Near
    rest the maximum of a momentum field is itself a round-off residual of
    hydrostatic balance, so an error relative to it says nothing about the
    step's accuracy."""
    cs = np.sqrt(p.Cp / (p.Cp - p.Rd) * p.Rd * p.Tmean)
    scale = np.abs(os_["Rho"]).max() * cs
    return {k: float(np.abs(gs[k] - os_[k]).max() / scale) for k in ("Mh", "Wh", "W")}


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a - b| / max |b| (0 if both are zero)."""
    den = np.abs(b).max()
    num = np.abs(a - b).max()
    if den == 0:
        return float(num)
    return float(num / den)


# Hot-Jupiter set-up of the Newtonian-relaxation and double-gray runs
HJ_TOP = 7.4e6  # [m]: isothermal H = Rd T / g ~ 6.42e5 m, ln(1e7 / 1e2) ~ 11.5 scale heights


def hj_params(capi, glevel, dt=300.0, ns=6, **kw):
    """Hot Jupiter: R = 1e8 m, g = 9.42, Rd = 3779, Cp = 13226, Omega = 2.06e-5,
    P0 = 1e7 Pa, isothermal 1600 K start."""
    return capi.default_params(glevel=glevel, timestep=dt, n_substeps=ns, A=1e8, Omega=2.06e-5, Gravit=9.42,
                               P_Ref=1e7, Rd=3779.0, Cp=13226.0, Tmean=1600.0, **kw)


def hj_grid(capi, glevel, nv):
    return grid(capi, glevel, nv, top=HJ_TOP, radius=1e8)


def nr_params(capi):
    """Day-night relaxation: 1600 +- 600 K above 1e5 Pa, vanishing at 1e7 Pa;
    tau_rad 1e4 s at 1e2 Pa to 1e7 s at 1e7 Pa."""
    return capi.NewtonianParams(T_mean=1600.0, dT=600.0, lon_sub=0.0, p_top=1e2, p_bottom=1e7, p_contrast=1e5,
                                tau_top=1e4, tau_bottom=1e7)


def rt_params(capi, **kw):
    """Double-gray RT: kappa_sw 4e-4, kappa_lw 1e-3 m2/kg, Tint 100 K, incident
    flux sigma * (2000 K)^4 at the substellar point, f_lw 0.5, n_sw 1, n_lw 2,
    diffusivity 0.5, no albedo."""
    q = dict(kappa_sw=4e-4, kappa_lw=1e-3, n_sw=1.0, n_lw=2.0, f_lw=0.5, diff_ang=0.5, Tint=100.0, albedo=0.0,
             incflx=5.6704e-8 * 2000.0 ** 4, ps0=1e7, lon_sub=0.0)
    q.update(kw)
    return capi.RTParams(**q)


def hip_runtime():
    """The HIP runtime the library itself links, for device buffers a test
    owns (the caller's side of the boundary)."""
    import ctypes as C
    hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipMemsetAsync.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


HIP_H2D, HIP_D2H, HIP_D2D = 1, 2, 3


def max_rel(a: np.ndarray, b: np.ndarray) -> float:
    return rel_err(a, b)
