"""Vertical hyperdiffusion (THOR's VertHyDiff, vertical_diff,
thor_diff.h:1198-1426) in the oracle, pinned by a closed form (CPU).

In the shallow model (r = 1) on the uniform vertical grid, away from the two
ghost levels, one application of the operator is the second difference
(a[l+1] - 2 a[l] + a[l-1]) / dz^2. cos(k z) is its eigenfunction with
eigenvalue lam = -(4 / dz^2) sin^2(k dz / 2). With rho = 1 the order-6
operator (two applications, the scaling by sign Kdv6 rho, a third
application) maps a field a = cos(k z) to sign Kdv6 lam^3 a at every level at
least three layers from both ends, sign = (-1)^(6/2 + 1) = +1 and
Kdv6 = Diffc_v dz^6 / dt; the T result is also multiplied by Rd. The
momentum result of a purely tangential wind keeps its direction (the radial
part removed by Correct_Horizontal is zero). An index shift, a dropped scaling
or a wrong sign fails this. THOR's own kernel text is compared with the
oracle on a deep-model state in test_ref_kernels.py.
"""
from __future__ import annotations

import numpy as np

from common import earth_params, grid


def test_cosine_is_damped_by_the_closed_form(built):
    from oracle.oracle import Oracle
    capi = built
    g = grid(capi, 4, 24)
    P, nv = g.P, g.nv
    dt, Dv, order = 600.0, 2e-3, 6
    p = earth_params(capi, 4, dt=dt, DeepModel=0, VertHyDiff=1, VertHyDiffOrder=order, Diffc_v=Dv)
    o = Oracle(g, p)
    alt, alth = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    dz = alth[1] - alth[0]
    assert np.allclose(np.diff(alth), dz) and np.allclose(alt, alth[:-1] + dz / 2)
    k = 2 * np.pi / (7.3 * dz)
    a = np.cos(k * alt)
    lam = -(4 / dz**2) * np.sin(k * dz / 2) ** 2
    Kv = Dv * dz**order / dt
    factor = (+1.0) * Kv * lam**3
    np.testing.assert_allclose(o.get("Kdv6"), Kv, rtol=1e-15)

    lon, lat = g.lonlat.reshape(P, 2).T
    e_lon = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)], axis=1)
    amp = 1.0 + np.arange(P) / P  # a different amplitude in every column
    col = amp[:, None] * a[None, :]
    o.set("Rhok", np.ones(P * nv))
    o.set("temperature", (250.0 + col).reshape(-1))  # a constant offset is annihilated
    o.set("Wk", col.reshape(-1))
    o.set("Mhk", (col[:, :, None] * e_lon[:, None, :]).reshape(-1))
    o.kernel("vertical_diff")
    inner = slice(3, nv - 3)
    want = factor * col[:, inner]
    scale = np.abs(want).max()
    got_w = o.get("diffwv").reshape(P, nv)[:, inner]
    got_p = o.get("diffprv").reshape(P, nv)[:, inner]
    got_m = o.get("diffmv").reshape(P, nv, 3)[:, inner, :]
    err_w = np.abs(got_w - want).max() / scale
    err_p = np.abs(got_p - p.Rd * want).max() / (p.Rd * scale)
    err_m = np.abs(got_m - want[:, :, None] * e_lon[:, None, :]).max() / scale
    print(f"closed form vs oracle (levels 3..{nv - 4}): w {err_w:.1e}, R T {err_p:.1e}, momentum {err_m:.1e}; "
          f"damping per step at this wavelength {abs(factor) * dt:.2e}")
    assert err_w < 1e-9 and err_p < 1e-9 and err_m < 1e-9
    assert factor < 0  # damping
    np.testing.assert_array_equal(o.get("diffrv"), 0.0)


def test_unsupported_vertical_order_is_rejected(built):
    capi = built
    g = grid(capi, 4, 8)
    import pytest
    with pytest.raises(capi.ThorError) as e:
        capi.Dyn(g, earth_params(capi, 4, VertHyDiff=1, VertHyDiffOrder=3))
    assert e.value.code == capi.THOR_ERR_UNSUPPORTED
