"""Radiative-heating surrogate: training data from the exact double-gray RT,
normalisation, GPU training, and attaching the network to a dynamical core.

Inputs per column are x = [T_k (nv), ln p_k (nv), cos z] and targets the
exact heating Qheat_k (nv) [W/m3] of the double-gray ProfX
(src/physics/modules/inc/profx_RT.h:296-587). Everything numeric runs in the
library (thor_amd/csrc/mlp.hip); this module only moves arrays.
"""
from __future__ import annotations

import time

import numpy as np

from . import capi


# Hot-Jupiter set-up of the double-gray runs: R = 1e8 m, g = 9.42, Rd = 3779,
# Cp = 13226, Omega = 2.06e-5, P0 = 1e7 Pa, isothermal 1600 K start; top at
# 7.4e6 m (isothermal H = Rd T / g ~ 6.42e5 m, ln(1e7 / 1e2) ~ 11.5 scale
# heights, so the top layer sits near 1e2 Pa).
HJ_TOP = 7.4e6


def hot_jupiter(glevel: int, nv: int, dt: float = 300.0, ns: int = 6, precision: int = 0, cache_dir=None, **kw):
    """(grid, params) of the hot-Jupiter set-up. Hyperdiffusion and divergence
    damping as THOR's deep hot-Jupiter benchmark (ifile/deephj.thr: Diffc 0.02,
    DivDampc 0.03); kw overrides any parameter."""
    g = capi.Grid(glevel, nv, 1e8, HJ_TOP, spring=True, beta=1.15, cache_dir=cache_dir)
    return g, hot_jupiter_params(glevel, dt, ns, precision, **kw)


def hot_jupiter_params(glevel: int, dt: float = 300.0, ns: int = 6, precision: int = 0, **kw):
    """The parameters of hot_jupiter() alone."""
    q = dict(glevel=glevel, timestep=dt, n_substeps=ns, precision=precision, A=1e8, Omega=2.06e-5, Gravit=9.42,
             P_Ref=1e7, Rd=3779.0, Cp=13226.0, Tmean=1600.0, Diffc=0.02, DivDampc=0.03)
    q.update(kw)
    return capi.default_params(**q)


def double_gray(**kw) -> "capi.RTParams":
    """Double-gray RT of the hot-Jupiter runs: kappa_sw 4e-4, kappa_lw 1e-3
    m2/kg, Tint 100 K, incident flux sigma (2000 K)^4 at the substellar point,
    f_lw 0.5, n_sw 1, n_lw 2, diffusivity 0.5, no albedo."""
    q = dict(kappa_sw=4e-4, kappa_lw=1e-3, n_sw=1.0, n_lw=2.0, f_lw=0.5, diff_ang=0.5, Tint=100.0, albedo=0.0,
             incflx=5.6704e-8 * 2000.0 ** 4, ps0=1e7, lon_sub=0.0)
    q.update(kw)
    return capi.RTParams(**q)


def hot_jupiter_sponge() -> "capi.SpongeParams":
    """Rayleigh sponge of THOR's double-gray hot-Jupiter example
    (ifile/wasp43b_ex.thr): top quarter of the model, 1e-4 1/s for the winds,
    damping towards the zonal means over 20 latitude bands."""
    return capi.SpongeParams(0.75, 1e-4, 1e-4, 20, 1, 1)


def cos_zenith(grid, lon_sub: float = 0.0) -> np.ndarray:
    """cos z of a synchronous rotator with the substellar point at lon_sub, clipped at 0."""
    lon, lat = np.asarray(grid.lonlat).reshape(grid.P, 2).T
    return np.maximum(np.cos(lat) * np.cos(lon - lon_sub), 0.0)


def features(T: np.ndarray, p: np.ndarray, mu: np.ndarray) -> np.ndarray:
    """[T_k, ln p_k, cos z] per column (rows = columns)."""
    return np.concatenate([T, np.log(p), mu[:, None]], axis=1)


def normalise_inputs(X: np.ndarray, norm: dict) -> np.ndarray:
    """The normalised inputs as the surrogate kernels form them: the
    difference from the mean in FP64 rounded to FP32, times the FP32
    reciprocal of the scale."""
    inv = np.float32(1.0) / norm["in_std"].astype(np.float32)
    return (X - norm["in_mean"].astype(np.float64)).astype(np.float32) * inv


def collect(dyn: "capi.Dyn", grid, nsteps: int, every: int, rng=None, max_columns=None, lon_sub=0.0):
    """Runs nsteps of (double-gray ProfX + Thor) and samples (inputs, heating)
    of the columns after every `every`-th ProfX. Returns X [n, 2 nv + 1],
    Y [n, nv] as float64."""
    mu = cos_zenith(grid, lon_sub)
    xs, ys = [], []
    for s in range(nsteps):
        dyn.profx(capi.FORCING_DOUBLE_GRAY)
        if s % every == 0:
            T = dyn.get_field("temperature")
            p = dyn.get_field("pressure")
            Q = dyn.get_field("profx_Qheat")
            X = features(T, p, mu)
            if max_columns is not None and rng is not None and max_columns < len(X):
                keep = rng.choice(len(X), max_columns, replace=False)
                X, Q = X[keep], Q[keep]
            xs.append(X)
            ys.append(Q)
        dyn.thor()
    dyn.sync()
    return np.concatenate(xs), np.concatenate(ys)


def normalisation(X: np.ndarray, Y: np.ndarray, margin: float = 0.5) -> dict:
    """Per-feature mean / std of the inputs and outputs, and the envelope
    [min - margin, max + margin] of the normalised training inputs. The input
    scale is at least 1e-3 of the feature's typical magnitude, so that the
    FP32 rounding of a normalised input stays far below its BF16 resolution."""
    in_mean = X.mean(axis=0)
    in_std = np.maximum(X.std(axis=0), 1e-3 * np.maximum(np.abs(in_mean), 1.0))
    Xn = (X - in_mean) / in_std
    out_mean = Y.mean(axis=0)
    out_std = Y.std(axis=0)
    out_std[out_std == 0] = 1.0
    return {"in_mean": in_mean.astype(np.float32), "in_std": in_std.astype(np.float32),
            "in_lo": (Xn.min(axis=0) - margin).astype(np.float32),
            "in_hi": (Xn.max(axis=0) + margin).astype(np.float32),
            "out_mean": out_mean.astype(np.float32), "out_std": out_std.astype(np.float32)}


def apply_norm(X: np.ndarray, Y: np.ndarray, norm: dict):
    return normalise_inputs(X, norm), ((Y - norm["out_mean"]) / norm["out_std"]).astype(np.float32)


def train(mlp: "capi.MLP", Xn: np.ndarray, Yn: np.ndarray, epochs: int, batch: int = 512, lr: float = 1e-3,
          seed: int = 0, lr_decay: float = 0.93, log=None) -> list:
    """Adam epochs on the GPU with an exponentially decaying learning rate;
    returns the per-epoch training losses (normalised units)."""
    mlp.set_data(Xn, Yn)
    losses = []
    t0 = time.perf_counter()
    for e in range(epochs):
        losses.append(mlp.train_epoch(batch, lr * lr_decay ** e, seed + e))
        if log:
            log(f"  epoch {e:3d}: loss {losses[-1]:.4e} ({time.perf_counter() - t0:.1f} s)")
    return losses


def heating_error(mlp: "capi.MLP", X: np.ndarray, Y: np.ndarray, norm: dict) -> float:
    """RMS error of the network's heating against Y, relative to the RMS of Y."""
    Xn, _ = apply_norm(X, Y, norm)
    Yp = mlp.forward(Xn) * norm["out_std"] + norm["out_mean"]
    return float(np.sqrt(((Yp - Y) ** 2).mean()) / np.sqrt((Y ** 2).mean()))
