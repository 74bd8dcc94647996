"""Oracle forcings pinned by closed forms.

Double-gray RT (profx_RT.h:133-587): with longwave off the shortwave flux is
Beer-Lambert through the layers' optical depths; in an isothermal atmosphere
with no shortwave and no internal heat the longwave fluxes are the exact
two-stream solution B (1 - exp(-tau/mu)); and in every column the heating
integrates to the net flux entering it (ASR - OLR + sigma Tint^4 with the
shortwave absorbed before the bottom). Newtonian relaxation: Teq is a fixed
point and T - Teq shrinks by exactly 1 / (1 + dt / tau) with tau at the
profile's end points.
"""
from __future__ import annotations

import numpy as np
import pytest

from common import hj_grid, hj_params, nr_params, rt_params

BC = 5.677036E-8


@pytest.fixture(scope="module")
def hj(built):
    capi = built
    g = hj_grid(capi, 4, 20)
    p = hj_params(capi, 4)
    st = capi.ic_rest(g, p, seed=1001, temperature_noise=1.0)
    return capi, g, p, st


def _oracle(capi, g, p, st):
    from oracle.oracle import Oracle
    o = Oracle(g, p)
    o.set_state(st)
    return o


def _top_pressure(pr, alt, alth):
    nv = pr.shape[1]
    pp = pr[:, nv - 2] + (pr[:, nv - 1] - pr[:, nv - 2]) / (alt[nv - 1] - alt[nv - 2]) * (
        2 * alth[nv] - alt[nv - 1] - alt[nv - 2])
    return 0.5 * (pr[:, nv - 1] + np.maximum(pp, 0.0))


def test_rt_shortwave_is_beer_lambert(hj):
    capi, g, p, st = hj
    q = rt_params(capi, kappa_lw=0.0, Tint=0.0)
    o = _oracle(capi, g, p, st)
    o.set_rt(q)
    o.profx_forcing(capi.FORCING_DOUBLE_GRAY)
    P, nv = g.P, g.nv
    alt, alth = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    pr, rho = st["pressure"], st["Rho"]
    lon, lat = g.lonlat.reshape(P, 2).T
    mu = np.maximum(np.cos(lat) * np.cos(lon), 0.0)
    day = mu > 0
    np.testing.assert_allclose(o.get("insol"), q.incflx * mu, rtol=1e-14, atol=1e-9)
    tau = q.kappa_sw * rho * (alth[1:] - alth[:-1])[None, :]               # n_sw = 1
    above = np.cumsum(tau[:, ::-1], axis=1)[:, ::-1]                      # optical depth above layer bottoms
    tau_top = q.kappa_sw / p.Gravit * _top_pressure(pr, alt, alth)
    F = np.zeros((P, nv + 1))
    F[day, nv] = q.incflx * mu[day] * np.exp(-tau_top[day] / mu[day])
    F[day, :nv] = q.incflx * mu[day, None] * np.exp(-(tau_top[day, None] + above[day]) / mu[day, None])
    Q = (F[:, 1:] - F[:, :-1]) / (alth[1:] - alth[:-1])[None, :]
    got = o.get("profx_Qheat").reshape(P, nv)
    err = np.abs(got - Q).max() / np.abs(Q).max()
    print(f"shortwave heating vs Beer-Lambert: {err:.2e}")
    assert err < 1e-12
    r2 = ((p.A + alth[nv]) / p.A) ** 2  # ASR and OLR are per top-of-atmosphere area (deep model)
    np.testing.assert_allclose(o.get("ASR"), F[:, nv] * g.areasT * r2, rtol=1e-12, atol=1e-6)


def test_rt_isothermal_longwave_is_two_stream_solution(hj):
    capi, g, p, st = hj
    T0 = 1500.0
    P, nv = g.P, g.nv
    iso = dict(st)
    iso["pressure"] = st["Rho"] * p.Rd * T0
    q = rt_params(capi, kappa_sw=0.0, incflx=0.0, Tint=0.0)
    o = _oracle(capi, g, p, iso)
    o.set_rt(q)
    o.profx_forcing(capi.FORCING_DOUBLE_GRAY)
    alth = np.asarray(g.Altitudeh)
    pr, rho = iso["pressure"], iso["Rho"]
    tau = q.kappa_lw * (1 + q.n_lw * (1 - q.f_lw) / q.f_lw * pr / q.ps0) * rho * (alth[1:] - alth[:-1])[None, :]
    B = BC * T0 ** 4
    above = np.concatenate([np.cumsum(tau[:, ::-1], axis=1)[:, ::-1], np.zeros((P, 1))], axis=1)
    below = np.concatenate([np.zeros((P, 1)), np.cumsum(tau, axis=1)], axis=1)
    Fd = B * (1 - np.exp(-above / q.diff_ang))
    Fu = B + (Fd[:, :1] - B) * np.exp(-below / q.diff_ang)
    net = Fu - Fd
    Q = -(net[:, :-1] - net[:, 1:]) / (alth[:-1] - alth[1:])[None, :]
    got = o.get("profx_Qheat").reshape(P, nv)
    err = np.abs(got - Q).max() / np.abs(Q).max()
    print(f"isothermal longwave heating vs two-stream solution: {err:.2e}")
    assert err < 1e-9
    r2 = ((p.A + alth[nv]) / p.A) ** 2
    np.testing.assert_allclose(o.get("OLR"), Fu[:, nv] * g.areasT * r2, rtol=1e-12)


def test_rt_column_heating_balances_boundary_fluxes(hj):
    capi, g, p, st = hj
    q = rt_params(capi)
    o = _oracle(capi, g, p, st)
    o.set_rt(q)
    o.profx_forcing(capi.FORCING_DOUBLE_GRAY)
    P, nv = g.P, g.nv
    alth = np.asarray(g.Altitudeh)
    Q = o.get("profx_Qheat").reshape(P, nv)
    col = (Q * (alth[1:] - alth[:-1])[None, :]).sum(axis=1) * g.areasT
    # shortwave fully absorbed (tau_sw to the bottom ~ 400); the bottom gets Tint
    r2 = ((p.A + alth[nv]) / p.A) ** 2  # ASR, OLR are per top-of-atmosphere area
    want = (o.get("ASR") - o.get("OLR")) / r2 + BC * q.Tint ** 4 * g.areasT
    err = np.abs(col - want).max() / np.abs(o.get("OLR")).max()
    print(f"column heating vs boundary fluxes: {err:.2e}")
    assert err < 1e-10


def test_newtonian_fixed_point_and_decay(hj):
    capi, g, p, st = hj
    q = nr_params(capi)
    P, nv = g.P, g.nv
    lon, lat = g.lonlat.reshape(P, 2).T
    pr = st["pressure"]
    lnp = np.log(pr)
    w = np.clip((np.log(q.p_bottom) - lnp) / (np.log(q.p_bottom) - np.log(q.p_contrast)), 0, 1)
    Teq = q.T_mean + q.dT * w * (np.cos(lon) * np.cos(lat))[:, None]
    fr = np.clip((lnp - np.log(q.p_top)) / (np.log(q.p_bottom) - np.log(q.p_top)), 0, 1)
    tau = np.exp(np.log(q.tau_top) + fr * (np.log(q.tau_bottom) - np.log(q.tau_top)))
    # Teq is a fixed point (p kept, rho set so that T = Teq(p))
    at_eq = dict(st)
    at_eq["Rho"] = pr / (p.Rd * Teq)
    o = _oracle(capi, g, p, at_eq)
    o.set_newtonian(q)
    o.profx_forcing(capi.FORCING_NEWTONIAN)
    T1 = o.get("temperature").reshape(P, nv)
    assert np.abs(T1 - Teq).max() < 1e-9
    # a uniform 1500 K column relaxes by 1 / (1 + dt / tau)
    cold = dict(st)
    cold["Rho"] = pr / (p.Rd * 1500.0)
    o = _oracle(capi, g, p, cold)
    o.set_newtonian(q)
    o.profx_forcing(capi.FORCING_NEWTONIAN)
    T1 = o.get("temperature").reshape(P, nv)
    Teq_c = q.T_mean + q.dT * np.clip((np.log(q.p_bottom) - np.log(cold["pressure"])) /
                                      (np.log(q.p_bottom) - np.log(q.p_contrast)), 0, 1) * \
        (np.cos(lon) * np.cos(lat))[:, None]
    fr_c = np.clip((np.log(cold["pressure"]) - np.log(q.p_top)) / (np.log(q.p_bottom) - np.log(q.p_top)), 0, 1)
    tau_c = np.exp(np.log(q.tau_top) + fr_c * (np.log(q.tau_bottom) - np.log(q.tau_top)))
    ratio = (T1 - Teq_c) / (1500.0 - Teq_c)
    np.testing.assert_allclose(ratio, 1.0 / (1.0 + p.timestep / tau_c), rtol=1e-8)
    # end points of the profile: the deepest level sits near p_bottom with tau ~ tau_bottom
    assert tau_c.max() <= q.tau_bottom * (1 + 1e-12) and tau_c.min() >= q.tau_top * (1 - 1e-12)
