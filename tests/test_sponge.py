"""The oracle's Rayleigh sponge (THOR's RayleighSponge, implicit mode,
profx_driver.cu:92-195, profx_sponge.cu:50-437) pinned by closed forms (CPU).

* Latitude bins (grid.cu:3215-3262): every point is in exactly one bin, the
  indices within a bin run 0..count-1 in point order, bins follow latitude,
  and the slot count is the largest bin rounded up to a power of two.
* Damping towards zero: in the local east/north frame the winds become
  u / (1 + k dt) with k = R sin^2(pi/2 (n - ns) / (1 - ns)) above n = ns and
  are unchanged below; W likewise, and Wh at interior interfaces is the
  linear interpolation of the new W.
* Damping towards the zonal mean with a rate so large that the relaxation is
  complete: the winds of a point in a polar half-band become the arithmetic
  mean of its band (numpy's mean over the band's points); a wind field that
  is the same everywhere is left as it is.
* Temperature sponge (RayleighSpongeT): with a rate so large that the
  relaxation is complete, the temperature p / (Rd rho) of a point in a polar
  half-band becomes the arithmetic mean of its band; a temperature that is
  the same in every column of a level is left as it is.
The Newtonian forcing drives ProfX; the sponge tests read the state right
after the sponge by running the forcing with a zero rate where it matters.
"""
from __future__ import annotations

import numpy as np
import pytest

from common import hj_grid, hj_params, nr_params


@pytest.fixture(scope="module")
def hj(built):
    capi = built
    g = hj_grid(capi, 4, 20)
    p = hj_params(capi, 4)
    st = dict(capi.ic_rest(g, p, seed=1006, temperature_noise=1.0, wind_noise=5.0))
    # a vertical wind too: random Wh at the interior interfaces
    Wh = st["Wh"].reshape(g.P, g.nv + 1).copy()
    Wh[:, 1:g.nv] = 1e-3 * np.random.default_rng(1).standard_normal((g.P, g.nv - 1))
    st["Wh"] = Wh.reshape(st["Wh"].shape)
    return capi, g, p, st


def _frame(g):
    lon, lat = g.lonlat.reshape(g.P, 2).T
    lon = np.where(lon < 0, lon + 2 * np.pi, lon)
    e_lon = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)], axis=1)
    e_lat = np.stack([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)], axis=1)
    return lat, e_lon, e_lat


def _run(capi, g, p, st, q):
    from oracle.oracle import Oracle
    o = Oracle(g, p)
    o.set_state(st)
    o.set_newtonian(nr_params(capi))
    o.set_sponge(q)
    o.profx_forcing(capi.FORCING_NEWTONIAN)
    return o


def _w_of(g, Wh):
    """Recompute_W: W at layer centres from Wh."""
    alt, alth = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    intt = (alt - alth[1:]) / (alth[:-1] - alth[1:])
    intl = (alt - alth[:-1]) / (alth[1:] - alth[:-1])
    return Wh[:, :-1] * intt + Wh[:, 1:] * intl


def test_zonal_table(hj):
    capi, g, p, st = hj
    q = capi.SpongeParams(0.75, 1e-4, 1e-4, 20, 1, 1)
    o = _run(capi, g, p, st, q)
    t, mc = o.zonal_table()
    b, cnt, idx = t.T
    lat = g.lonlat.reshape(g.P, 2)[:, 1]
    counts = np.bincount(b, minlength=20)
    assert counts.sum() == g.P and (cnt == counts[b]).all()
    for j in range(20):
        ids = np.nonzero(b == j)[0]
        assert (idx[ids] == np.arange(len(ids))).all()  # point order within the bin
    assert (np.diff(b[np.argsort(lat, kind="stable")]) >= 0).all()
    assert mc == 1 << int(np.ceil(np.log2(counts.max())))
    assert counts[0] > 0 and counts[-1] > 0  # poles at the ends


def test_damping_to_zero_is_implicit_relaxation(hj):
    capi, g, p, st = hj
    ns, R, Rw, dt = 0.5, 3e-4, 2e-4, p.timestep
    o = _run(capi, g, p, st, capi.SpongeParams(ns, R, Rw, 12, 0, 0))
    P, nv = g.P, g.nv
    alt, alth = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    rho = st["Rho"].reshape(P, nv)
    M0 = st["Mh"].reshape(P, nv, 3)
    W0 = _w_of(g, st["Wh"].reshape(P, nv + 1))
    lat, e_lon, e_lat = _frame(g)
    u = np.einsum("pk,plk->pl", e_lon, M0) / rho
    v = np.einsum("pk,plk->pl", e_lat, M0) / rho
    n = alt / alth[nv]
    s2 = np.where(n >= ns, np.sin(0.5 * np.pi * (n - ns) / (1 - ns)) ** 2, 0.0)
    fu, fw = 1 / (1 + R * s2 * dt), 1 / (1 + Rw * s2 * dt)
    M = (u * fu)[:, :, None] * e_lon[:, None, :] + (v * fu)[:, :, None] * e_lat[:, None, :]
    M *= rho[:, :, None]
    W = W0 * fw
    got_M = o.get("Mh").reshape(P, nv, 3)
    got_W = o.get("W").reshape(P, nv)
    errM = np.abs(got_M - M).max() / np.abs(M0).max()
    errW = np.abs(got_W - W).max() / np.abs(W0).max()
    print(f"sponge to zero: Mh {errM:.2e}, W {errW:.2e}; damped fraction at the top {1 - fu[-1]:.3f}")
    assert errM < 1e-13 and errW < 1e-13
    assert fu[-1] < 0.99 and (fu[n < ns] == 1).all()
    # interior Wh interpolates the new W; the boundary interfaces are untouched
    Wh = o.get("Wh").reshape(P, nv + 1)
    xi, xim, xip = alth[1:nv], alt[:-1], alt[1:]
    want = W[:, :-1] * ((xi - xip) / (xim - xip)) + W[:, 1:] * ((xi - xim) / (xip - xim))
    assert np.abs(Wh[:, 1:nv] - want).max() <= 1e-13 * np.abs(W0).max()
    np.testing.assert_array_equal(Wh[:, [0, nv]], st["Wh"].reshape(P, nv + 1)[:, [0, nv]])


def test_full_damping_to_the_zonal_mean(hj):
    capi, g, p, st = hj
    nlat, P, nv = 10, g.P, g.nv
    o = _run(capi, g, p, st, capi.SpongeParams(0.0, 1e9, 1e9, nlat, 1, 1))
    t, _ = o.zonal_table()
    b = t[:, 0]
    rho = st["Rho"].reshape(P, nv)
    M0 = st["Mh"].reshape(P, nv, 3)
    W0 = _w_of(g, st["Wh"].reshape(P, nv + 1))
    lat, e_lon, e_lat = _frame(g)
    u = np.einsum("pk,plk->pl", e_lon, M0) / rho
    w = W0 / rho
    got_u = np.einsum("pk,plk->pl", e_lon, o.get("Mh").reshape(P, nv, 3)) / rho
    got_w = o.get("W").reshape(P, nv) / rho
    des = np.pi / nlat
    polar = (lat <= -np.pi / 2 + 0.5 * des) | (lat >= -np.pi / 2 + 0.5 * des + des * (nlat - 1))
    lev = nv - 1  # n = Altitude / Top close to 1: k = 1e9 s^-1, relaxation complete to ~1e-12
    for j in (0, nlat - 1):
        ids = np.nonzero(polar & (b == j))[0]
        assert len(ids) > 0
        mu, mw = u[b == j, lev].mean(), w[b == j, lev].mean()
        eu = np.abs(got_u[ids, lev] - mu).max() / np.abs(u[:, lev]).max()
        ew = np.abs(got_w[ids, lev] - mw).max() / np.abs(w[:, lev]).max()
        print(f"band {j}: {len(ids)} polar points, u vs band mean {eu:.1e}, w {ew:.1e}")
        assert eu < 1e-9 and ew < 1e-9


def test_uniform_wind_is_left_alone(hj):
    capi, g, p, st = hj
    P, nv = g.P, g.nv
    lat, e_lon, e_lat = _frame(g)
    rho = st["Rho"].reshape(P, nv)
    u0, v0 = 30.0, -4.0
    M = rho[:, :, None] * (u0 * e_lon[:, None, :] + v0 * e_lat[:, None, :])
    s2 = dict(st)
    s2["Mh"] = M.reshape(st["Mh"].shape)
    o = _run(capi, g, p, s2, capi.SpongeParams(0.5, 1e-3, 1e-3, 20, 1, 1))
    got = o.get("Mh").reshape(P, nv, 3)
    err = np.abs(got - M).max() / np.abs(M).max()
    print(f"uniform (u, v) = ({u0}, {v0}) after the sponge: max rel change {err:.2e}")
    assert err < 1e-12


def _sponge_only(capi, g, p, st, q):
    """The oracle state right after the sponge: the sponge runs at the start
    of ProfX; the Newtonian forcing after it gets zero rates."""
    from oracle.oracle import Oracle
    o = Oracle(g, p)
    o.set_state(st)
    nr = nr_params(capi)
    nr.tau_top = nr.tau_bottom = 1e300  # relaxation times so long that T is left as it is
    o.set_newtonian(nr)
    o.set_sponge(q)
    o.profx_forcing(capi.FORCING_NEWTONIAN)
    return o


def test_full_temperature_damping_to_the_zonal_mean(hj):
    capi, g, p, st = hj
    nlat, P, nv = 10, g.P, g.nv
    o = _sponge_only(capi, g, p, st, capi.SpongeParams(0.0, 0.0, 0.0, nlat, 0, 0, 1, 1e9))
    t, _ = o.zonal_table()
    b = t[:, 0]
    lat = g.lonlat.reshape(P, 2)[:, 1]
    T0 = (st["pressure"] / (p.Rd * st["Rho"])).reshape(P, nv)
    T1 = (o.get("pressure") / (p.Rd * o.get("Rho"))).reshape(P, nv)
    des = np.pi / nlat
    polar = (lat <= -np.pi / 2 + 0.5 * des) | (lat >= -np.pi / 2 + 0.5 * des + des * (nlat - 1))
    lev = nv - 1
    for j in (0, nlat - 1):
        ids = np.nonzero(polar & (b == j))[0]
        assert len(ids) > 0
        mt = T0[b == j, lev].mean()
        e = np.abs(T1[ids, lev] - mt).max() / mt
        print(f"band {j}: {len(ids)} polar points, T vs band mean {e:.1e} (spread before {np.ptp(T0[b == j, lev]):.2f} K)")
        assert e < 1e-12
    # the wind rates are zero: the winds are rebuilt unchanged (to rounding)
    M0 = st["Mh"].reshape(-1)
    assert np.abs(o.get("Mh") - M0).max() < 1e-13 * np.abs(M0).max()


def test_uniform_temperature_is_left_alone(hj):
    capi, g, p, st = hj
    P, nv = g.P, g.nv
    s2 = dict(st)
    Tl = 1500.0 + 10.0 * np.arange(nv)
    s2["pressure"] = (p.Rd * st["Rho"].reshape(P, nv) * Tl[None, :]).reshape(st["pressure"].shape)
    o = _sponge_only(capi, g, p, s2, capi.SpongeParams(0.3, 0.0, 0.0, 20, 0, 0, 1, 1e-3))
    p0 = np.asarray(s2["pressure"]).reshape(-1)
    err = np.abs(o.get("pressure") - p0).max() / np.abs(p0).max()
    print(f"level-uniform T after the temperature sponge: max rel change of p {err:.2e}")
    assert err < 1e-12


def test_explicit_sponge_drag_towards_zero(hj):
    """raysp_calc_mode exp1 damping to zero: the drag of stage 0 is
    profx_dMh = -k_uv (M - (M.r) r) (the horizontal momentum), profx_dWh the
    interpolation of -k_w W at the interior interfaces (0 at the ends), and
    the heating of the temperature sponge Rd rho (-k_T (T - Tbar)) vanishes for
    a temperature that is the same in every column of a level."""
    from oracle.oracle import Oracle
    capi, g, p, st = hj
    P, nv = g.P, g.nv
    ns, R, Rw, RT = 0.4, 3e-4, 2e-4, 5e-4
    s2 = dict(st)
    Tl = 1500.0 + 10.0 * np.arange(nv)
    s2["pressure"] = (p.Rd * st["Rho"].reshape(P, nv) * Tl[None, :]).reshape(np.shape(st["pressure"]))
    o = Oracle(g, p)
    o.set_state(s2)
    o.set_sponge(capi.SpongeParams(ns, R, Rw, 12, 0, 0, 1, RT, 1))
    for a, b in (("Mhk", "Mh"), ("Rhok", "Rho"), ("pressurek", "pressure")):
        o.set(a, o.get(b))
    o.set("Wk", _w_of(g, st["Wh"].reshape(P, nv + 1)).reshape(-1))  # the fixture's random vertical wind
    q0 = o.get("profx_Qheat").copy()
    o.kernel("sponge_explicit", 0)
    alt, alth = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    n = alt / alth[nv]
    s2n = np.where(n >= ns, np.sin(0.5 * np.pi * (n - ns) / (1 - ns)) ** 2, 0.0)
    M = o.get("Mhk").reshape(P, nv, 3)
    r = g.func_r.reshape(P, 3)
    Mt = M - np.einsum("plk,pk->pl", M, r)[:, :, None] * r[:, None, :]
    want = -(R * s2n)[None, :, None] * Mt
    got = o.get("profx_dMh").reshape(P, nv, 3)
    err = np.abs(got - want).max() / np.abs(want).max()
    W = o.get("Wk").reshape(P, nv)
    dW = -(Rw * s2n)[None, :] * W
    xi, xim, xip = alth[1:nv], alt[:-1], alt[1:]
    dWh = o.get("profx_dWh").reshape(P, nv + 1)
    wantWh = dW[:, :-1] * ((xi - xip) / (xim - xip)) + dW[:, 1:] * ((xi - xim) / (xip - xim))
    errW = np.abs(dWh[:, 1:nv] - wantWh).max() / np.abs(wantWh).max()
    dq = np.abs(o.get("profx_Qheat") - q0).max()
    print(f"explicit sponge drag vs closed form: dMh {err:.1e}, dWh {errW:.1e}; level-uniform T heating {dq:.1e}")
    assert err < 1e-12 and errW < 1e-12
    np.testing.assert_array_equal(dWh[:, [0, nv]], 0.0)
    assert dq < 1e-12 * np.abs(p.Rd * o.get("Rhok") * 1600 * RT).max()


def test_shrinking_sponge_scales_the_rates(hj):
    """shrink_sponge with t_shrink = 0: at step 1 every rate carries
    Rv_fac = exp(-(1 - 0) dt / (1000 dt)) = exp(-1e-3), so damping to zero
    gives u / (1 + Rv_fac k dt)."""
    capi, g, p, st = hj
    ns, R, dt = 0.5, 3e-4, p.timestep
    q = capi.SpongeParams(ns, R, R, 12, 0, 0, 0, 0.0, 0, 1, 0)
    o = _run(capi, g, p, st, q)
    P, nv = g.P, g.nv
    alt, alth = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    n = alt / alth[nv]
    s2 = np.where(n >= ns, np.sin(0.5 * np.pi * (n - ns) / (1 - ns)) ** 2, 0.0)
    rv = np.exp(-1e-3)
    fu = 1 / (1 + rv * R * s2 * dt)
    rho = st["Rho"].reshape(P, nv)
    lat, e_lon, e_lat = _frame(g)
    M0 = st["Mh"].reshape(P, nv, 3)
    u0 = np.einsum("pk,plk->pl", e_lon, M0) / rho
    u1 = np.einsum("pk,plk->pl", e_lon, o.get("Mh").reshape(P, nv, 3)) / rho
    err = np.abs(u1 - u0 * fu).max() / np.abs(u0).max()
    plain = 1 / (1 + R * s2 * dt)
    sep = np.abs(u0 * (fu - plain)).max() / np.abs(u0).max()
    print(f"shrinking sponge: u vs closed form {err:.1e}; the factor changes u by {sep:.1e}")
    assert err < 1e-13 and sep > 1e3 * max(err, 1e-16)
