"""THOR's diffusive sponge (DiffSponge, thor_diff.h:453-509, strengths
esp_initial.cu:1297-1321) in the oracle, pinned by closed forms (CPU).

* Strength: Kdh2 = Dv_sponge sin^2(pi/2 (n - ns) / (1 - ns)) dbar^2 / dt above
  n = Altitude / Top = ns and zero below, dbar = sqrt(2 pi / 5) A / 2^glevel;
  with order_diff_sponge = HyDiffOrder the same strength (times dbar^order)
  is added to Kdh4 instead.
* Second-order sponge with fourth-order hyperdiffusion: the last pass's
  momentum and w results change by exactly (-1)^2 rho Kdh2 times the pass
  input at the point (the first pass's Laplacian of the velocity), and the
  density and R T results do not change.
THOR's own Diffusion_Op with the sponge is compared with the oracle in
test_ref_kernels.py.
"""
from __future__ import annotations

import numpy as np

from common import earth_params, grid


def _oracle(capi, g, **kw):
    from oracle.oracle import Oracle
    p = earth_params(capi, 4, dt=600.0, **kw)
    return Oracle(g, p), p


def test_sponge_strength_profile(built):
    capi = built
    g = grid(capi, 4, 16)
    ns, Dv = 0.5, 0.01
    o, p = _oracle(capi, g, DiffSponge=1, order_diff_sponge=2, Dv_sponge=Dv, ns_diff_sponge=ns)
    o4, _ = _oracle(capi, g, DiffSponge=1, order_diff_sponge=4, Dv_sponge=Dv, ns_diff_sponge=ns)
    off, _ = _oracle(capi, g)
    alt, top = np.asarray(g.Altitude), g.Altitudeh[g.nv]
    n = alt / top
    ks = np.where(n > ns, Dv * np.sin(0.5 * np.pi * (n - ns) / (1 - ns)) ** 2, 0.0)
    dbar = np.sqrt(2 * np.pi / 5) * p.A / 2 ** p.glevel
    np.testing.assert_allclose(o.get("Kdh2"), ks * dbar**2 / p.timestep, rtol=1e-13, atol=0)
    np.testing.assert_array_equal(o.get("Kdh4"), off.get("Kdh4"))
    np.testing.assert_allclose(o4.get("Kdh4") - off.get("Kdh4"), ks * dbar**4 / p.timestep, rtol=1e-12,
                               atol=1e-12 * off.get("Kdh4").max())
    np.testing.assert_array_equal(o4.get("Kdh2"), 0.0)
    assert (ks[n <= ns] == 0).all() and ks[-1] > 0.9 * Dv


def test_second_order_sponge_adds_rho_kdh2_times_first_pass(built):
    capi = built
    g = grid(capi, 4, 16)
    P, nv = g.P, g.nv
    kw = dict(order_diff_sponge=2, Dv_sponge=0.05, ns_diff_sponge=0.25)
    rng = np.random.default_rng(11)
    D = rng.standard_normal((P, nv, 6))
    rho = 1.0 + 0.1 * rng.random((P, nv))
    out = {}
    for on in (0, 1):
        o, p = _oracle(capi, g, DiffSponge=on, **kw)
        o.set("Rhok", rho.reshape(-1))
        o.set("diff", D.reshape(-1))
        o.kernel("diffusion", 1, 1)
        out[on] = {k: o.get(k) for k in ("diffmh", "diffw", "diffrh", "diffpr")}
        if on:
            k2 = o.get("Kdh2")
    want = rho[:, :, None] * k2[None, :, None] * D[:, :, 1:5]
    dm = (out[1]["diffmh"] - out[0]["diffmh"]).reshape(P, nv, 3)
    dw = (out[1]["diffw"] - out[0]["diffw"]).reshape(P, nv)
    scale = np.abs(out[0]["diffmh"]).max()
    err = max(np.abs(dm - want[:, :, :3]).max(), np.abs(dw - want[:, :, 3]).max()) / scale
    print(f"sponge increment vs rho Kdh2 (pass input): {err:.1e} of the hyperdiffusion scale; "
          f"sponge / hyperdiffusion at the top {np.abs(want[:, -1, :3]).max() / scale:.2e}")
    assert err < 1e-14  # one rounding of (lap + sponge)
    assert np.abs(want).max() > 1e-3 * scale
    np.testing.assert_array_equal(out[1]["diffrh"], out[0]["diffrh"])
    np.testing.assert_array_equal(out[1]["diffpr"], out[0]["diffpr"])


def test_unsupported_sponge_order_is_rejected(built):
    import pytest
    capi = built
    g = grid(capi, 4, 8)
    for order in (3, 6, 0):
        with pytest.raises(capi.ThorError) as e:
            capi.Dyn(g, earth_params(capi, 4, DiffSponge=1, order_diff_sponge=order))
        assert e.value.code == capi.THOR_ERR_UNSUPPORTED
