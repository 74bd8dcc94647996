"""Boundary-layer module, RayleighHS type (boundary_layer.cu:441-490), in the
oracle (CPU), pinned by what the drag's construction fixes:

  * it only rescales the momentum of each cell, by a factor that depends on
    the pressures alone (two different momentum fields get the same factors);
  * no drag where sigma = p / ps <= bl_sigma;
  * ps, extrapolated from the two lowest layers, is the surface pressure of a
    hydrostatic column (the rest state's P_Ref, to the discretisation);
  * at the lowest layer the factor is 1 / (1 + kv dt) with
    kv = surf_drag (sigma - bl_sigma) / (1 - bl_sigma) for that sigma;
  * GravHeightVar scales the extrapolation's gravity by (A / (A + z0))^2.
"""
import numpy as np
import pytest

from common import earth_params, grid


def _factors(capi, g, p, st, bl, seed):
    from oracle.oracle import Oracle
    o = Oracle(g, p)
    o.set_state(st)
    o.set_boundary_layer(bl)
    rng = np.random.default_rng(seed)
    mh = rng.uniform(0.5, 2.0, (g.P, g.nv, 3)) * rng.choice([-1.0, 1.0], (g.P, g.nv, 3))
    o.set("Mh", mh.ravel())
    o.kernel("boundary_layer")
    out = o.get("Mh").reshape(g.P, g.nv, 3)
    return out / mh


@pytest.mark.parametrize("ghv", [0, 1])
def test_rayleigh_boundary_layer_closed_forms(built, ghv):
    capi = built
    g = grid(capi, 4, 20)
    p = earth_params(capi, 4, dt=600.0, GravHeightVar=ghv)
    st = capi.ic_rest(g, p, seed=3, temperature_noise=0.0)
    bl = capi.BLParams(surf_drag=1.0 / 86400.0, bl_sigma=0.7)
    c1 = _factors(capi, g, p, st, bl, 1)
    c2 = _factors(capi, g, p, st, bl, 2)
    # one factor per cell, the same for every component and any momentum
    assert np.allclose(c1, c1[:, :, :1], rtol=1e-15, atol=0)
    assert np.allclose(c1, c2, rtol=1e-15, atol=0)
    c = c1[:, :, 0]
    assert (c > 0).all() and (c <= 1).all()
    pr = st["pressure"]
    rho = st["Rho"]
    z0, z1 = g.Altitude[0], g.Altitude[1]
    # the extrapolated surface pressure of the hydrostatic rest column is P_Ref
    # to the extrapolation's error over 1.5 layers ((z / H)^2 ~ 0.4 % at 1.8 km
    # layers; a wrong sign in the extrapolation is off by ~20 %)
    gfac = (p.A / (p.A + z0)) ** 2 if ghv else 1.0
    ps = 0.5 * (pr[:, 0] + pr[:, 1] + rho[:, 0] * p.Gravit * gfac * (z0 + z1))
    assert np.abs(ps / p.P_Ref - 1).max() < 1e-2
    sigma = pr / ps[:, None]
    # no drag above bl_sigma; drag below it, growing towards the surface
    assert (c[sigma <= 0.7] == 1.0).all()
    assert (c[sigma > 0.7 + 1e-3] < 1.0).all()
    assert (np.diff(c, axis=1)[sigma[:, 1:] > 0.7] >= 0).all()
    kv = bl.surf_drag * (sigma[:, 0] - 0.7) / (1 - 0.7)
    assert np.allclose(c[:, 0], 1.0 / (1.0 + kv * p.timestep), rtol=1e-13, atol=0)
