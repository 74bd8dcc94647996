"""Pins of the oracle's fast-mode and column kernels by what the mathematics
fixes (CPU only). Each check builds the expected result in numpy by another
route than the oracle's code, so a dropped term, a wrong sign, an index shift
or a transposed operand in the restatement fails it:

  Vertical_Eq (thor_vertical_int.h:52-563): the solve is the elimination of
    three substep equations; here they are composed as operators (layer
    continuity, layer pressure tendency, interface momentum) and solved
    densely, then compared with the oracle's Whs;
  Density_Pressure_Eqs (thor_fastmodes.h:255-847): with no fluxes the
    potential-temperature form is an exact round trip, p' advances by
    dt (diffpr + Rd/Cv Q) and rho' by dt SlowRho; W' = c / r^2 is free of
    vertical divergence in the deep shell; horizontal fluxes through the
    point_local neighbour table (not the tile maps) give the density change
    and the pressure in closed form;
  Momentum_Eq (thor_fastmodes.h:54-253) and Prepare_Implicit_Vertical
    (thor_vertical_int.h:566-771): the gradient and divergence sums through
    point_local, the radial projection, the R/Cv factor;
  Compute_Temperature_H_Pt_Geff (thor_auxiliary.h:55-309): h = Cp T, the
    textbook potential temperature, interface values exact for fields linear
    in height (end extrapolations included), gtilh = g on a hydrostatic
    column;
  UpdateRK2 (thor_auxiliary.h:420-529): W at layer centres exact for Wh
    linear in height, the zero-flux boundaries.

The grids are THOR's uniform vertical grids (Altitudeh = k Top / nv), on
which the reference's pairing of interpolation weights with layer
thicknesses reduces to plain averages.
"""
from __future__ import annotations

import numpy as np
import pytest

from common import earth_params, grid

DT_MS = 100000  # substep length: timestep 600 s / 6 substeps = 100 s, exact in ms
DT = DT_MS * 1e-3


@pytest.fixture(scope="module")
def g4(built):
    g = grid(built, 4, 20)
    dz = np.diff(g.Altitudeh)
    assert np.allclose(dz, dz[0], rtol=1e-12, atol=0)  # the uniform grid these checks assume
    return built, g


def _oracle(capi, g, **kw):
    from oracle.oracle import Oracle
    p = earth_params(capi, 4, dt=600.0, ns=6, **kw)
    st = capi.ic_rest(g, p, seed=1015, temperature_noise=1.0, wind_noise=2.0)
    o = Oracle(g, p)
    o.set_state(st)
    o.steps(1)  # every intermediate array allocated and realistic
    return p, o


def _nbr(g):
    """Neighbour index of stencil slot j (0 = the point itself, 1..6 = the ring
    in THOR's order) from point_local; pentagon and pole slots without a
    neighbour point back to the point (their weights are zero)."""
    P = g.P
    nb = np.empty((P, 7), dtype=np.int64)
    nb[:, 0] = np.arange(P)
    pl = np.asarray(g.point_local, dtype=np.int64)
    nb[:, 1:] = np.where(pl >= 0, pl, np.arange(P)[:, None])
    pent = np.asarray(g.pent_ind)
    nb[pent, 6] = nb[pent, 0]
    nb[P - 2:, 6] = nb[P - 2:, 0]
    return nb


def _weights(g, name):
    """Operator weights [P][7][3] with the slots that have no neighbour set to
    zero (pentagons and poles have five neighbours; the grid may store
    anything there)."""
    w = np.array(getattr(g, name), dtype=np.float64)
    w[np.asarray(g.pent_ind), 6, :] = 0.0
    w[g.P - 2:, 6, :] = 0.0
    return w


def _stencil_vec(w, nb, F):
    """sum_j sum_k w[p, j, k] F[nb[p, j], lev, k]  ->  [P][nv]"""
    return np.einsum("pjk,pjlk->pl", w, F[nb])


def _stencil_sca(w, nb, f):
    """sum_j w[p, j, k] f[nb[p, j], lev]  ->  [P][nv][3]"""
    return np.einsum("pjk,pjl->plk", w, f[nb])


def _rscale(g, p):
    return p.A / (np.asarray(g.Altitude) + p.A) if p.DeepModel else np.ones(g.nv)


def _tangential(v, fr):
    return v - np.einsum("plk,pk->pl", v, fr)[:, :, None] * fr[:, None, :]


# ------------------------------------------------------------- Vertical_Eq
@pytest.mark.parametrize("deep,nonhydro,ghv", [(1, 1, 0), (0, 1, 0), (1, 0, 0), (0, 0, 0), (1, 1, 1), (0, 1, 1)])
def test_vertical_eq_solves_the_composed_substep_equations(g4, deep, nonhydro, ghv):
    capi, g = g4
    p, o = _oracle(capi, g, DeepModel=deep, NonHydro=nonhydro, GravHeightVar=ghv)
    P, nv = g.P, g.nv
    rng = np.random.default_rng(100 + 4 * deep + 2 * nonhydro + ghv)
    R, Cp, grav, A = p.Rd, p.Cp, p.Gravit, p.A
    Cv = Cp - R
    shp, shi = (P, nv), (P, nv + 1)
    hh = Cp * 280.0 * (1 + 0.2 * rng.uniform(-1, 1, shi))
    gtilh = grav * (1 + 0.3 * rng.uniform(-1, 1, shi))
    Sp = rng.uniform(-5, 5, shp)
    Sd = rng.uniform(-1e-4, 1e-4, shp)
    ps = rng.uniform(-50, 50, shp)
    rhos = rng.uniform(-1e-3, 1e-3, shp)
    srh = rng.uniform(-1e-3, 1e-3, shi)
    wold = rng.uniform(-1e-2, 1e-2, shi)
    wold[:, 0] = wold[:, nv] = 0.0
    for name, a in (("hh", hh), ("gtilh", gtilh), ("Sp", Sp), ("Sd", Sd), ("pressures", ps), ("Rhos", rhos),
                    ("SlowWh", srh), ("Whs", wold)):
        o.set(name, a)
    o.kernel("vertical_eq", DT_MS, 0)
    got = o.get("Whs").reshape(shi)
    ws_got = o.get("Ws").reshape(shp)

    alt, alth = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    dzk = alth[1:] - alth[:-1]
    rh = alth + A if deep else np.ones(nv + 1)
    gl = grav * (A / (A + alt)) ** 2 if ghv else np.full(nv, grav)
    cw = 1.0 if nonhydro else 0.0
    i = np.arange(1, nv)          # interior interfaces
    lo, up = i - 1, i             # the layers below and above interface i
    dzh = 1.0 / (alt[up] - alt[lo])
    wl, wu = (alt[up] - alth[i]) * dzh, (alth[i] - alt[lo]) * dzh  # linear interpolation to alth_i
    gimp = wl * gl[lo] + wu * gl[up]                               # g of the implicit coupling
    gexp = gl[up] if ghv else np.full(nv - 1, grav)                # g of the explicit terms

    def s(k):  # 1 / r^2 at layer k as a first-order expansion about interface i (deep)
        if not deep:
            return np.ones(nv - 1)
        return 1.0 / rh[i] ** 2 - 2.0 * (alt[k] - alth[i]) / rh[i] ** 3

    def E(W):
        """Left side of the interface momentum equations for new W (interfaces
        0..nv, zero at both ends): W + dt^2 d/dz(P(W)) + dt^2 g rho'(D(W))."""
        F = rh ** 2 * hh * W
        Ak = (F[:, 1:] - F[:, :-1]) / dzk                     # layer flux divergence of r^2 h W
        Bk = 0.5 * (gtilh[:, :-1] * W[:, :-1] + gtilh[:, 1:] * W[:, 1:])  # layer mean of gtil W
        Dk = -(W[:, 1:] - W[:, :-1]) / dzk                    # layer continuity (flux form)
        Pu = -(R / Cv) * (s(up) * Ak[:, up] + Bk[:, up])
        Pl = -(R / Cv) * (s(lo) * Ak[:, lo] + Bk[:, lo])
        return cw * W[:, i] + DT ** 2 * dzh * (Pu - Pl) + DT ** 2 * gimp * (wl * Dk[:, lo] + wu * Dk[:, up])

    M = np.empty((P, nv - 1, nv - 1))
    for j in range(1, nv):
        e = np.zeros(shi)
        e[:, j] = 1.0
        M[:, :, j - 1] = E(e)
    rhs = (cw * wold[:, i] + DT * srh[:, i]
           - DT * dzh * ((ps[:, up] + DT * Sp[:, up]) - (ps[:, lo] + DT * Sp[:, lo]))
           - DT * gexp * (wl * rhos[:, lo] + wu * rhos[:, up] + DT * (wl * Sd[:, lo] + wu * Sd[:, up])))
    want = np.linalg.solve(M, rhs[:, :, None])[:, :, 0]
    err = np.abs(got[:, 1:nv] - want).max() / np.abs(want).max()
    print(f"DeepModel {deep} NonHydro {nonhydro} GravHeightVar {ghv}: max |Whs| {np.abs(want).max():.3e}, "
          f"rel diff {err:.2e}")
    assert np.all(got[:, 0] == 0) and np.all(got[:, nv] == 0)
    assert err < 1e-11
    # W at the layer centres: linear interpolation of the new Whs
    wb = (alth[1:] - alt) / dzk
    ws_want = wb * got[:, :-1] + (1 - wb) * got[:, 1:]
    assert np.abs(ws_got - ws_want).max() <= 1e-14 * np.abs(ws_want).max()


# ------------------------------------------------- Density_Pressure_Eqs
def _dp_setup(o, g, p, rng):
    P, nv = g.P, g.nv
    shp = (P, nv)
    st = {
        "pressurek": 1e5 * np.exp(-np.asarray(g.Altitude) / 8000.0)[None, :] * (1 + 0.01 * rng.uniform(-1, 1, shp)),
        "Rhok": 1.2 * np.exp(-np.asarray(g.Altitude) / 8000.0)[None, :] * (1 + 0.01 * rng.uniform(-1, 1, shp)),
        "pressures": rng.uniform(-30, 30, shp), "Rhos": rng.uniform(-1e-4, 1e-4, shp),
        "SlowRho": rng.uniform(-1e-6, 1e-6, shp), "diffpr": rng.uniform(-1e-3, 1e-3, shp),
        "diffprv": rng.uniform(-1e-3, 1e-3, shp), "profx_Qheat": rng.uniform(-1e-2, 1e-2, shp),
        "pt": 300.0 + 50 * rng.uniform(0, 1, shp), "pth": 300.0 + 50 * rng.uniform(0, 1, (P, nv + 1)),
        "Mhs": np.zeros((P, nv, 3)), "Mhk": np.zeros((P, nv, 3)), "Whs": np.zeros((P, nv + 1)),
        "Whk": np.zeros((P, nv + 1)),
    }
    for k, v in st.items():
        o.set(k, v)
    return st


def _dp_run(o, g):
    o.kernel("density_pressure", DT_MS, 0)
    return o.get("Rhos").reshape(g.P, g.nv), o.get("pressures").reshape(g.P, g.nv)


@pytest.mark.parametrize("deep", [1, 0])
def test_density_pressure_without_fluxes_is_a_round_trip(g4, deep):
    """aux = theta(p, rho) rho and p = P0 (Rd aux / P0)^(Cp/Cv) undo each other:
    with no fluxes p' advances by dt (diffpr + diffprv + Rd/Cv Qheat) and rho'
    by dt SlowRho."""
    capi, g = g4
    p, o = _oracle(capi, g, DeepModel=deep)
    st = _dp_setup(o, g, p, np.random.default_rng(7))
    rs, ps = _dp_run(o, g)
    Cv = p.Cp - p.Rd
    want_p = st["pressures"] + DT * (st["diffpr"] + st["diffprv"]) + DT * p.Rd / Cv * st["profx_Qheat"]
    want_r = st["Rhos"] + DT * st["SlowRho"]
    ptot = st["pressurek"] + st["pressures"]
    e_p = float(np.abs(ps - want_p).max() / ptot.max())
    e_r = float(np.abs(rs - want_r).max() / np.abs(want_r).max())
    print(f"deep {deep}: p' error {e_p:.2e} of p, rho' error {e_r:.2e}")
    assert e_p < 2e-15 and e_r < 1e-15


def test_density_pressure_inverse_square_vertical_flux_is_divergence_free(g4):
    """Deep shell: the flux W' = c / r^2 at the interfaces has no vertical
    divergence in the cell's finite volume, so rho' and p' change only by
    their sources (layers above the first, whose bottom flux the reference
    takes as zero)."""
    capi, g = g4
    p, o = _oracle(capi, g, DeepModel=1)
    st = _dp_setup(o, g, p, np.random.default_rng(8))
    c = 3.0 * p.A ** 2
    r = np.asarray(g.Altitudeh) + p.A
    whs = np.tile(c / r ** 2, (g.P, 1))
    o.set("Whs", whs)
    st["pth"] = np.full((g.P, g.nv + 1), 310.0)
    o.set("pth", st["pth"])
    rs, ps = _dp_run(o, g)
    Cv = p.Cp - p.Rd
    alth = np.asarray(g.Altitudeh)
    one_term = DT * c / ((r[1:] ** 3 - r[:-1] ** 3) / 3.0)  # size of one face's contribution
    want_r = st["Rhos"] + DT * st["SlowRho"]
    e_r = np.abs(rs - want_r)[:, 1:] / one_term[None, 1:]
    want_p = st["pressures"] + DT * (st["diffpr"] + st["diffprv"]) + DT * p.Rd / Cv * st["profx_Qheat"]
    ptot = st["pressurek"] + st["pressures"]
    e_p = np.abs(ps - want_p)[:, 1:] / ptot[:, 1:]
    # the bottom layer sees only its top face: -dt c r_1^2 / r_1^2 / volume factor
    bottom = st["Rhos"][:, 0] + DT * st["SlowRho"][:, 0] - DT * c / ((r[1] ** 3 - r[0] ** 3) / 3.0)
    e_b = np.abs(rs[:, 0] - bottom)[:-2] / one_term[0]  # (the poles use dz r_mid^2 as volume)
    print(f"rho' error {e_r.max():.2e} of one face flux, p' {e_p.max():.2e}, bottom {e_b.max():.2e}; "
          f"dz/r = {(alth[1] - alth[0]) / p.A:.1e}")
    assert e_r.max() < 1e-11 and e_p.max() < 1e-13 and e_b.max() < 1e-12


@pytest.mark.parametrize("deep", [1, 0])
def test_density_pressure_horizontal_fluxes(g4, deep):
    """Horizontal flux divergences through the point_local neighbour table:
    rho' = rho' + dt (SlowRho - rs div M'), and the potential-temperature
    flux of (M' + Mk) pt changes the pressure to
    P0 (Rd (theta rho - dt rs div((M' + Mk) pt)) / P0)^(Cp/Cv) - pk + sources."""
    capi, g = g4
    p, o = _oracle(capi, g, DeepModel=deep)
    rng = np.random.default_rng(9)
    st = _dp_setup(o, g, p, rng)
    P, nv = g.P, g.nv
    fr = np.asarray(g.func_r)
    ms = _tangential(rng.uniform(-1, 1, (P, nv, 3)), fr)
    mk = _tangential(rng.uniform(-20, 20, (P, nv, 3)), fr)
    o.set("Mhs", ms)
    o.set("Mhk", mk)
    rs, ps = _dp_run(o, g)
    nb, div = _nbr(g), _weights(g, "div")
    rsc = _rscale(g, p)[None, :]
    divm = rsc * _stencil_vec(div, nb, ms)
    flx = rsc * _stencil_vec(div, nb, (ms + mk) * st["pt"][:, :, None])
    want_r = st["Rhos"] + DT * (st["SlowRho"] - divm)
    Cv = p.Cp - p.Rd
    rho = st["Rhok"] + st["Rhos"]
    ptot = st["pressurek"] + st["pressures"]
    theta = p.P_Ref / (p.Rd * rho) * (ptot / p.P_Ref) ** (Cv / p.Cp)
    aux = theta * rho - DT * flx
    pnew = p.P_Ref * (p.Rd * aux / p.P_Ref) ** (p.Cp / Cv)
    want_p = pnew - st["pressurek"] + DT * (st["diffpr"] + st["diffprv"]) + DT * p.Rd / Cv * st["profx_Qheat"]
    scale_r = DT * rsc * _stencil_vec(np.abs(div), nb, np.abs(ms))
    e_r = float((np.abs(rs - want_r) / scale_r).max())
    e_p = float((np.abs(ps - want_p) / ptot).max())
    dp = float((np.abs(pnew - ptot) / ptot).max())
    print(f"deep {deep}: rho' error {e_r:.2e} of the flux terms, p' error {e_p:.2e} of p "
          f"(flux changes p by up to {dp:.1e})")
    assert e_r < 1e-13 and e_p < 1e-14 and dp > 1e-6


# ------------------------------------------- Momentum_Eq, Prepare_Implicit_Vertical
@pytest.mark.parametrize("deep", [1, 0])
def test_momentum_eq(g4, deep):
    """Mhs += dt (-rs grad p' + SlowMh + DivM), radial part removed: with
    p' = 0 exactly dt times the tangential part of SlowMh + DivM; with the
    sources zero, -dt rs times the tangential gradient of p' summed through
    point_local."""
    capi, g = g4
    p, o = _oracle(capi, g, DeepModel=deep)
    P, nv = g.P, g.nv
    rng = np.random.default_rng(10 + deep)
    fr = np.asarray(g.func_r)
    m0 = _tangential(rng.uniform(-1, 1, (P, nv, 3)), fr)
    slow, divm = rng.uniform(-1e-3, 1e-3, (P, nv, 3)), rng.uniform(-1e-3, 1e-3, (P, nv, 3))
    for name, a in (("Mhs", m0), ("SlowMh", slow), ("DivM", divm), ("pressures", np.zeros((P, nv)))):
        o.set(name, a)
    o.kernel("momentum_eq", DT_MS, 0)
    got = o.get("Mhs").reshape(P, nv, 3)
    want = m0 + _tangential(DT * (slow + divm), fr)
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()
    pr = rng.uniform(-100, 100, (P, nv))
    for name, a in (("Mhs", m0), ("SlowMh", np.zeros_like(slow)), ("DivM", np.zeros_like(divm)),
                    ("pressures", pr)):
        o.set(name, a)
    o.kernel("momentum_eq", DT_MS, 0)
    got = o.get("Mhs").reshape(P, nv, 3)
    grad = _stencil_sca(_weights(g, "grad"), _nbr(g), pr) * _rscale(g, p)[None, :, None]
    want = m0 - DT * _tangential(grad, fr)
    scale = DT * _stencil_sca(np.abs(_weights(g, "grad")), _nbr(g), np.abs(pr)).max()
    err = float(np.abs(got - want).max() / scale)
    print(f"deep {deep}: momentum from grad p' error {err:.2e} of the gradient terms")
    assert err < 1e-13


@pytest.mark.parametrize("deep", [1, 0])
def test_prepare_implicit_vertical(g4, deep):
    """Sd = -rs div Mhs + SlowRho and Sp = -(Rd/Cv) rs div(Mhs h) + Slowpressure,
    divergences summed through point_local."""
    capi, g = g4
    p, o = _oracle(capi, g, DeepModel=deep)
    P, nv = g.P, g.nv
    rng = np.random.default_rng(12 + deep)
    fr = np.asarray(g.func_r)
    ms = _tangential(rng.uniform(-1, 1, (P, nv, 3)), fr)
    h = 2.8e5 * (1 + 0.1 * rng.uniform(-1, 1, (P, nv)))
    sr, sp = rng.uniform(-1e-5, 1e-5, (P, nv)), rng.uniform(-1, 1, (P, nv))
    for name, a in (("Mhs", ms), ("h", h), ("SlowRho", sr), ("Slowpressure", sp)):
        o.set(name, a)
    o.kernel("prepare_implicit_vertical", 0, 0)
    Sd, Sp = o.get("Sd").reshape(P, nv), o.get("Sp").reshape(P, nv)
    nb, div = _nbr(g), _weights(g, "div")
    rsc = _rscale(g, p)[None, :]
    RoC = p.Rd / (p.Cp - p.Rd)
    want_d = -rsc * _stencil_vec(div, nb, ms) + sr
    want_p = -RoC * rsc * _stencil_vec(div, nb, ms * h[:, :, None]) + sp
    sc_d = rsc * _stencil_vec(np.abs(div), nb, np.abs(ms))
    sc_p = RoC * rsc * _stencil_vec(np.abs(div), nb, np.abs(ms) * h[:, :, None])
    e_d, e_p = float((np.abs(Sd - want_d) / sc_d).max()), float((np.abs(Sp - want_p) / sc_p).max())
    print(f"deep {deep}: Sd error {e_d:.2e}, Sp error {e_p:.2e} of the flux terms")
    assert e_d < 1e-13 and e_p < 1e-13


# ---------------------------------------------- Compute_Temperature_H_Pt_Geff
def test_temperature_h_pt_on_columns_linear_in_height(g4):
    """T and theta linear in height (p and rho follow): T = p / (rho Rd),
    h = Cp T, pt = T (P0 / p)^(Rd/Cp); the interface values hh and pth are
    exact for fields linear in height; hh takes the end layers' values; pth
    is extrapolated to the ground at the bottom but, at the top, to the ghost
    point 2 Altitudeh[nv] - Altitude[nv-1] above the model top (the
    reference's extrapolation factor, thor_auxiliary.h:189-190, 232-234)."""
    capi, g = g4
    p, o = _oracle(capi, g)
    P, nv = g.P, g.nv
    rng = np.random.default_rng(14)
    z, zh = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    a_t, b_t = 290 + 10 * rng.uniform(0, 1, (P, 1)), -6e-3 * (1 + 0.2 * rng.uniform(-1, 1, (P, 1)))
    a_th, b_th = 300 + 10 * rng.uniform(0, 1, (P, 1)), 4e-3 * (1 + 0.2 * rng.uniform(-1, 1, (P, 1)))
    T, th = a_t + b_t * z, a_th + b_th * z
    kappa = p.Rd / p.Cp
    pr = p.P_Ref * (T / th) ** (1.0 / kappa)
    rho = pr / (p.Rd * T)
    o.set("pressurek", pr)
    o.set("Rhok", rho)
    o.set("Whk", np.zeros((P, nv + 1)))
    o.kernel("temperature_h_pt_geff", 0, 0)
    get = lambda n, s: o.get(n).reshape(P, s)
    Tg, hg, ptg, hhg, pthg = get("temperature", nv), get("h", nv), get("pt", nv), get("hh", nv + 1), get("pth", nv + 1)
    assert np.abs(Tg / T - 1).max() < 1e-14
    assert np.abs(hg / (p.Cp * T) - 1).max() < 1e-14
    assert np.abs(ptg / (T * (p.P_Ref / pr) ** kappa) - 1).max() < 1e-13
    hh_want = p.Cp * (a_t + b_t * zh)
    hh_want[:, 0], hh_want[:, nv] = p.Cp * T[:, 0], p.Cp * T[:, nv - 1]
    assert np.abs(hhg / hh_want - 1).max() < 1e-13
    pth_want = a_th + b_th * zh
    pth_want[:, nv] = (a_th + b_th * (2 * zh[nv] - z[nv - 1]))[:, 0]
    assert np.abs(pthg / pth_want - 1).max() < 1e-12


def test_gtil_on_a_hydrostatic_linear_column(g4):
    """Constant density and p = pS - g rho z: -(1/rho_h) dp/dz = g at every
    interface, the two ends included (their ghost pressures are the column's
    mirror images); the layer term gtil is g times W interpolated to the layer
    centre, with the boundary interfaces taken as zero."""
    capi, g = g4
    p, o = _oracle(capi, g)
    P, nv = g.P, g.nv
    rng = np.random.default_rng(15)
    z, zh = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    rho0 = 1.0 + 0.1 * rng.uniform(0, 1, (P, 1))
    pr = 1e6 - p.Gravit * rho0 * z
    wh = rng.uniform(-1e-2, 1e-2, (P, nv + 1))
    o.set("pressurek", pr)
    o.set("Rhok", np.broadcast_to(rho0, (P, nv)).copy())
    o.set("Whk", wh)
    o.kernel("temperature_h_pt_geff", 0, 0)
    gtilh = o.get("gtilh").reshape(P, nv + 1)
    gtil = o.get("gtil").reshape(P, nv)
    assert np.abs(gtilh / p.Gravit - 1).max() < 1e-10
    w = wh.copy()
    w[:, 0] = w[:, nv] = 0.0
    wb = (zh[1:] - z) / (zh[1:] - zh[:-1])
    want = p.Gravit * (wb * w[:, :-1] + (1 - wb) * w[:, 1:])  # -(W / rho)(-g rho): the density cancels
    assert np.abs(gtil - want).max() < 1e-10 * np.abs(want).max()


# ---------------------------------------------------------------- UpdateRK2
def test_update_rk2_closed_forms(g4):
    """y(k) += deviation with the radial part of M removed, Wh zero at the top
    and bottom, and W exact at the layer centres for Wh linear in height."""
    capi, g = g4
    p, o = _oracle(capi, g)
    P, nv = g.P, g.nv
    rng = np.random.default_rng(16)
    fr = np.asarray(g.func_r)
    zh, z = np.asarray(g.Altitudeh), np.asarray(g.Altitude)
    a, b = rng.uniform(-1, 1, (P, 1)), rng.uniform(-1e-4, 1e-4, (P, 1))
    whk, whs = a + b * zh, np.zeros((P, nv + 1))
    mk, ms = rng.uniform(-5, 5, (P, nv, 3)), rng.uniform(-1, 1, (P, nv, 3))
    rk, rs = rng.uniform(0.5, 1, (P, nv)), rng.uniform(-1e-3, 1e-3, (P, nv))
    pk, psd = rng.uniform(1e4, 1e5, (P, nv)), rng.uniform(-10, 10, (P, nv))
    for name, v in (("Whk", whk), ("Whs", whs), ("Mhk", mk), ("Mhs", ms), ("Rhok", rk), ("Rhos", rs),
                    ("pressurek", pk), ("pressures", psd)):
        o.set(name, v)
    o.kernel("update_rk2", 0, 0)
    get = lambda n, *s: o.get(n).reshape(P, *s)
    assert np.abs(get("Mhk", nv, 3) - _tangential(mk + ms, fr)).max() < 1e-14 * np.abs(mk).max()
    assert np.array_equal(get("Rhok", nv), rk + rs) and np.array_equal(get("pressurek", nv), pk + psd)
    wh = get("Whk", nv + 1)
    assert np.all(wh[:, 0] == 0) and np.all(wh[:, nv] == 0)
    assert np.array_equal(wh[:, 1:nv], whk[:, 1:nv])
    # layers away from the ends see Wh linear in height: W = a + b z exactly
    w = get("Wk", nv)
    assert np.abs(w[:, 1:nv - 1] - (a + b * z)[:, 1:nv - 1]).max() < 1e-14


def test_update_rk_stage_deviations(g4):
    """At the start of RK stages 2 and 3 the deviations restart from
    y(n) - y(k): the tangential part of Mh - Mhk, Rho - Rhok, p - pk, Wh - Whk
    at the interior interfaces (zero at both ends) and its interpolation Ws,
    exact at the layer centres for a difference linear in height."""
    capi, g = g4
    p, o = _oracle(capi, g)
    P, nv = g.P, g.nv
    rng = np.random.default_rng(17)
    fr = np.asarray(g.func_r)
    zh, z = np.asarray(g.Altitudeh), np.asarray(g.Altitude)
    a, b = rng.uniform(-1, 1, (P, 1)), rng.uniform(-1e-4, 1e-4, (P, 1))
    whk = rng.uniform(-1, 1, (P, nv + 1))
    wh = whk + a + b * zh
    mh, mk = rng.uniform(-5, 5, (P, nv, 3)), rng.uniform(-5, 5, (P, nv, 3))
    rho, rk = rng.uniform(0.5, 1, (P, nv)), rng.uniform(0.5, 1, (P, nv))
    pr, pk = rng.uniform(1e4, 1e5, (P, nv)), rng.uniform(1e4, 1e5, (P, nv))
    for name, v in (("Wh", wh), ("Whk", whk), ("Mh", mh), ("Mhk", mk), ("Rho", rho), ("Rhok", rk),
                    ("pressure", pr), ("pressurek", pk)):
        o.set(name, v)
    o.kernel("update_rk", 0, 0)
    get = lambda n, *s: o.get(n).reshape(P, *s)
    assert np.abs(get("Mhs", nv, 3) - _tangential(mh - mk, fr)).max() < 1e-14 * np.abs(mh).max()
    assert np.array_equal(get("Rhos", nv), rho - rk) and np.array_equal(get("pressures", nv), pr - pk)
    whs = get("Whs", nv + 1)
    assert np.all(whs[:, 0] == 0) and np.all(whs[:, nv] == 0)
    assert np.array_equal(whs[:, 1:nv], wh[:, 1:nv] - whk[:, 1:nv])
    ws = get("Ws", nv)
    assert np.abs(ws[:, 1:nv - 1] - (a + b * z)[:, 1:nv - 1]).max() < 1e-12


# ------------------------------------------------------- energy equation
# thermo_equation = energy (esp.cu:696-710): Compute_Temperature_H_Pt_Geff
# forms the potential, kinetic and total energy (thor_auxiliary.h:143-161,
# 212-230, 281-287), the hyperdiffusion's fifth field becomes the internal
# energy (Cp - Rd) T (thor_diff.h:196-202), the slow pressure tendency takes
# its diffusion times Rd / Cv (thor_slowmodes.h:375-381), and the acoustic
# substeps advance the total energy with the flux of h + ep + ek and recover
# the pressure from it (thor_fastmodes.h:352-533).

def _energy_oracle(capi, g, **kw):
    return _oracle(capi, g, thermo_equation=capi.THERMO_ENERGY, **kw)


@pytest.mark.parametrize("ghv", [0, 1])
def test_energy_column_values(g4, ghv):
    """Total energy density rho g z + rho |u|^2 / 2 + rho Cv T; the flux
    variable h + ep + ek = Cp T + g z + |u|^2 / 2 at the centres; at the
    interfaces, for T and |u|^2 linear in height (g z linear too without
    GravHeightVar), the exact interface values: hh takes the end layers'
    values, ep is g z at the ground and the model top, and ek is extrapolated
    to the ground and to the ghost point 2 Altitudeh[nv] - Altitude[nv-1]."""
    capi, g = g4
    p, o = _energy_oracle(capi, g, GravHeightVar=ghv)
    P, nv = g.P, g.nv
    rng = np.random.default_rng(31)
    z, zh = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    rho0 = 1.0 + 0.2 * rng.uniform(0, 1, (P, 1))
    a_t, b_t = 290 + 10 * rng.uniform(0, 1, (P, 1)), -6e-3 * (1 + 0.2 * rng.uniform(-1, 1, (P, 1)))
    a_u, b_u = 100 + 50 * rng.uniform(0, 1, (P, 1)), 1e-2 * rng.uniform(0, 1, (P, 1))  # u^2 = a + b z
    T = a_t + b_t * z
    u = np.sqrt(a_u + b_u * z)
    rho = np.broadcast_to(rho0, (P, nv)).copy()
    # u split over the three components, and a vertical part w with u^2 = |uh|^2 + w^2
    c = rng.uniform(0.2, 1.0, (P, 1, 4))
    c /= np.sqrt((c ** 2).sum(axis=2, keepdims=True))
    mh = rho[:, :, None] * u[:, :, None] * c[:, :, :3]
    wk = rho * u * c[:, :, 3]
    o.set("pressurek", rho * p.Rd * T)
    o.set("Rhok", rho)
    o.set("Mhk", mh)
    o.set("Wk", wk)
    o.set("Whk", np.zeros((P, nv + 1)))
    o.kernel("temperature_h_pt_geff", 0, 0)
    get = lambda n, s: o.get(n).reshape(P, s)
    Cv = p.Cp - p.Rd
    gz = p.Gravit * (p.A / (p.A + z)) ** 2 * z if ghv else p.Gravit * z
    gzh = p.Gravit * (p.A / (p.A + zh)) ** 2 * zh if ghv else p.Gravit * zh
    E = rho * gz + 0.5 * rho * u ** 2 + rho * Cv * T
    assert np.abs(get("Etotal", nv) / E - 1).max() < 1e-14
    flux = get("h", nv) + get("epotential", nv) + get("ekinetic", nv)
    assert np.abs(flux / (p.Cp * T + gz + 0.5 * u ** 2) - 1).max() < 1e-14
    eph, ekh = get("epotentialh", nv + 1), get("ekinetich", nv + 1)
    assert np.abs(eph[:, [0, nv]] - gzh[[0, nv]]).max() <= 1e-15 * gzh.max()
    ek_want = 0.5 * (a_u + b_u * zh)
    ek_want[:, nv] = (0.5 * (a_u + b_u * (2 * zh[nv] - z[nv - 1])))[:, 0]
    assert np.abs(ekh / ek_want - 1).max() < 1e-12
    if not ghv:
        assert np.abs(eph[:, 1:nv] / gzh[1:nv] - 1).max() < 1e-13
        hh = get("hh", nv + 1)
        hh_want = p.Cp * (a_t + b_t * zh)
        hh_want[:, 0], hh_want[:, nv] = p.Cp * T[:, 0], p.Cp * T[:, nv - 1]
        assert np.abs(hh / hh_want - 1).max() < 1e-13


@pytest.mark.parametrize("deep,ghv", [(1, 0), (0, 0), (1, 1)])
def test_energy_density_pressure(g4, deep, ghv):
    """One acoustic substep of the energy form through the point_local
    neighbour table: E' = E + dt (diffpr + diffprv + Qheat)
    - dt (rs div((M' + Mk)(h + ep + ek)) + d(W (hh + eph + ekh))/dz), and
    p' = (Rd/Cv)(E' - rho' g z - |(M' + Mk, w)|^2 / (2 rho')) - pk with the
    updated density and w the substep's vertical momentum at the centre."""
    capi, g = g4
    p, o = _energy_oracle(capi, g, DeepModel=deep, GravHeightVar=ghv)
    rng = np.random.default_rng(32)
    st = _dp_setup(o, g, p, rng)
    P, nv = g.P, g.nv
    z, zh = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    fr = np.asarray(g.func_r)
    ms = _tangential(rng.uniform(-1, 1, (P, nv, 3)), fr)
    mk = _tangential(rng.uniform(-20, 20, (P, nv, 3)), fr)
    whk = rng.uniform(-0.5, 0.5, (P, nv + 1))
    whk[:, 0] = whk[:, nv] = 0.0
    arr = {"Mhs": ms, "Mhk": mk, "Whk": whk,
           "h": 3e5 + 1e4 * rng.uniform(0, 1, (P, nv)), "epotential": rng.uniform(0, 3e5, (P, nv)),
           "ekinetic": rng.uniform(0, 300, (P, nv)), "hh": 3e5 + 1e4 * rng.uniform(0, 1, (P, nv + 1)),
           "epotentialh": rng.uniform(0, 3e5, (P, nv + 1)), "ekinetich": rng.uniform(0, 300, (P, nv + 1))}
    Cv = p.Cp - p.Rd
    E0 = (st["Rhok"] + st["Rhos"]) * p.Gravit * z + Cv / p.Rd * (st["pressurek"] + st["pressures"])
    arr["Etotal"] = E0
    for k, v in arr.items():
        o.set(k, v)
    rs, ps = _dp_run(o, g)
    E1 = o.get("Etotal").reshape(P, nv)
    nb, div = _nbr(g), _weights(g, "div")
    rsc = _rscale(g, p)[None, :]
    flux_c = arr["h"] + arr["epotential"] + arr["ekinetic"]
    flux_h = arr["hh"] + arr["epotentialh"] + arr["ekinetich"]
    v1 = ms + mk
    flx = rsc * _stencil_vec(div, nb, v1 * flux_c[:, :, None])
    r2 = (lambda x: (x + p.A) ** 2) if deep else (lambda x: np.ones_like(x))
    wl, wt = whk[:, :-1].copy(), whk[:, 1:]
    phl, pht = flux_h[:, :-1].copy(), flux_h[:, 1:]
    wl[:, 0] = phl[:, 0] = 0.0
    dwpt = (wt * pht * r2(zh[1:]) - wl * phl * r2(zh[:-1])) / ((zh[1:] - zh[:-1]) * r2(z))
    E_want = E0 + DT * (st["diffpr"] + st["diffprv"]) + DT * st["profx_Qheat"] - DT * (flx + dwpt)
    e_E = float((np.abs(E1 - E_want) / E0).max())
    want_r = st["Rhos"] + DT * (st["SlowRho"] - rsc * _stencil_vec(div, nb, ms))
    rn = st["Rhok"] + want_r
    w = 0.5 * (wl + wt)  # the centre is the middle of a uniform layer
    gz = p.Gravit * (p.A / (p.A + z)) ** 2 * z if ghv else p.Gravit * z
    ekin = 0.5 * ((v1 ** 2).sum(axis=2) + w ** 2) / rn
    want_p = p.Rd / Cv * (E_want - rn * gz - ekin) - st["pressurek"]
    ptot = st["pressurek"] + st["pressures"]
    e_p = float((np.abs(ps - want_p) / ptot).max())
    dE = float((np.abs(E_want - E0) / E0).max())
    print(f"deep {deep} ghv {ghv}: E error {e_E:.2e} (substep changes E by up to {dE:.1e}), p' error {e_p:.2e} of p")
    assert e_E < 1e-15 and e_p < 1e-13 and dE > 1e-6
    assert float((np.abs(rs - want_r)).max()) < 1e-12 * float(np.abs(want_r).max())


_COPY = ["Mh", "Wh", "W", "Rho", "pressure", "temperature", "profx_Qheat", "profx_dMh", "profx_dWh", "Mhk", "Whk",
         "Wk", "Rhok", "pressurek", "h", "pt", "gtil", "hh", "pth", "gtilh", "Adv", "v", "diff", "diffmh", "diffw",
         "diffrh", "diffpr", "DivM", "divg_Mh", "diffmv", "diffwv", "diffrv", "diffprv"]


def _pair(capi, g):
    """An entropy-form and an energy-form oracle holding the same arrays."""
    _, oe = _oracle(capi, g)
    p, on = _energy_oracle(capi, g)
    for n in _COPY:
        on.set(n, oe.get(n))
    return p, oe, on


def test_energy_diffusion_field_is_internal_energy(g4):
    """The first hyperdiffusion pass is linear in its fields: with the energy
    equation the fifth plane is the Laplacian of Cv T instead of Rd T, so it
    equals the entropy form's times Cv / Rd; the other planes are unchanged."""
    capi, g = g4
    p, oe, on = _pair(capi, g)
    oe.kernel("diffusion", 0, 0)
    on.kernel("diffusion", 0, 0)
    P, nv = g.P, g.nv
    de = oe.get("diff").reshape(P, nv, 6)
    dn = on.get("diff").reshape(P, nv, 6)
    assert np.array_equal(de[:, :, :5], dn[:, :, :5])
    ratio = (p.Cp - p.Rd) / p.Rd
    scale = np.abs(dn[:, :, 5]).max()
    assert scale > 0 and np.abs(dn[:, :, 5] - ratio * de[:, :, 5]).max() < 1e-13 * scale


def test_energy_slow_pressure_scales_the_diffusion(g4):
    """Slow pressure tendency: the energy form adds (Rd/Cv)(diffpr + diffprv)
    where the entropy form adds diffpr + diffprv; every other slow tendency
    is the same."""
    capi, g = g4
    p, oe, on = _pair(capi, g)
    oe.kernel("slow_modes", 0, 0)
    on.kernel("slow_modes", 0, 0)
    for n in ("SlowMh", "SlowWh", "SlowRho"):
        assert np.array_equal(oe.get(n), on.get(n)), n
    d = oe.get("diffpr") + oe.get("diffprv")
    want = (p.Rd / (p.Cp - p.Rd) - 1.0) * d
    got = on.get("Slowpressure") - oe.get("Slowpressure")
    scale = np.abs(oe.get("Slowpressure")).max()
    assert np.abs(d).max() > 1e-6 * scale
    assert np.abs(got - want).max() < 1e-13 * scale
