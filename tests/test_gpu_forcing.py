"""Hot-Jupiter forcings on the GPU against the oracle.

The parity build (portable elementary functions on both sides) must
reproduce the oracle bit for bit through the Newtonian relaxation, the
double-gray RT ProfX and the dynamics steps that follow them. The mixed-
precision step under the same forcing must stay close to FP64.
"""
from __future__ import annotations

import numpy as np
import pytest

from common import hj_grid, hj_params, nr_params, rt_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hj(built):
    capi = built
    g = hj_grid(capi, 4, 20)
    p = hj_params(capi, 4)
    st = capi.ic_rest(g, p, seed=1001, temperature_noise=1.0, wind_noise=0.1)
    return capi, g, p, st


def _max_rel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def test_newtonian_steps_bitwise(hj):
    from oracle.oracle import Oracle
    capi, g, p, st = hj
    print("case: parity build, hot Jupiter g4 L20, 2 x (Newtonian ProfX + Thor)", flush=True)
    d = capi.Dyn(g, p, variant="pm")
    d.set_state(st)
    d.set_newtonian(nr_params(capi))
    o = Oracle(g, p, variant="pm")
    o.set_state(st)
    o.set_newtonian(nr_params(capi))
    d.run(2, capi.FORCING_NEWTONIAN)
    d.sync()
    o.run(2, capi.FORCING_NEWTONIAN)
    gs, os_ = d.get_state(), o.get_state()
    errs = {k: _max_rel(gs[k], os_[k]) for k in gs}
    print("newtonian", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) == 0.0
    d.close()


def test_double_gray_profx_and_steps_bitwise(hj):
    from oracle.oracle import Oracle
    capi, g, p, st = hj
    print("case: parity build, hot Jupiter g4 L20, double-gray ProfX, then 2 x (ProfX + Thor)", flush=True)
    d = capi.Dyn(g, p, variant="pm")
    d.set_state(st)
    d.set_rt(rt_params(capi))
    o = Oracle(g, p, variant="pm")
    o.set_state(st)
    o.set_rt(rt_params(capi))
    d.profx(capi.FORCING_DOUBLE_GRAY)
    d.sync()
    o.profx_forcing(capi.FORCING_DOUBLE_GRAY)
    for name in ("profx_Qheat", "temperature", "pressure"):
        e = _max_rel(d.get_field(name).ravel(), o.get(name))
        print(f"  {name}: {e:.2e}")
        assert e == 0.0, name
    for name in ("insol", "ASR", "OLR"):
        e = _max_rel(d.rt_column(name), o.get(name))
        print(f"  {name}: {e:.2e}")
        assert e == 0.0, name
    d.thor()
    o.thor()
    d.run(1, capi.FORCING_DOUBLE_GRAY)
    d.sync()
    o.run(1, capi.FORCING_DOUBLE_GRAY)
    gs, os_ = d.get_state(), o.get_state()
    errs = {k: _max_rel(gs[k], os_[k]) for k in gs}
    print("double gray", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) == 0.0
    d.close()


@pytest.mark.parametrize("forcing", ["newtonian", "double_gray"])
def test_hot_jupiter_mixed_stays_close_to_fp64(hj, forcing):
    capi, g, p64, st = hj
    code = capi.FORCING_NEWTONIAN if forcing == "newtonian" else capi.FORCING_DOUBLE_GRAY
    print(f"case: hot Jupiter g4 L20, 10 steps of {forcing}, FP64 vs mixed", flush=True)
    out, mass = {}, {}
    for prec in (0, 1):
        p = hj_params(capi, 4, precision=prec)
        d = capi.Dyn(g, p)
        d.set_state(st)
        if code == capi.FORCING_NEWTONIAN:
            d.set_newtonian(nr_params(capi))
        else:
            d.set_rt(rt_params(capi))
        m0 = d.globdiag()[0]
        d.run(10, code)
        d.sync()
        out[prec] = d.get_state()
        mass[prec] = (d.globdiag()[0] - m0) / m0
        d.close()
    for prec in (0, 1):
        print(f"  precision {prec}: relative mass change {mass[prec]:.2e}")
        assert abs(mass[prec]) < 1e-10
        assert all(np.isfinite(v).all() for v in out[prec].values())
    tol = {"Mh": 1e-3, "Rho": 1e-6, "pressure": 1e-6}
    for k, t in tol.items():
        e = _max_rel(out[1][k], out[0][k])
        print(f"  mixed vs FP64 {k}: {e:.2e}")
        assert e < t, k


@pytest.mark.parametrize("forcing,temp", [("newtonian", 0), ("double_gray", 0), ("newtonian", 1), ("double_gray", 1)])
def test_sponge_steps_bitwise(hj, forcing, temp):
    """Parity build with the Rayleigh sponge (damping to the zonal means, with
    and without the temperature sponge) in ProfX: two steps reproduce the
    oracle bit for bit."""
    from oracle.oracle import Oracle
    capi, g, p, st0 = hj
    code = capi.FORCING_NEWTONIAN if forcing == "newtonian" else capi.FORCING_DOUBLE_GRAY
    print(f"case: parity build, hot Jupiter g4 L20 with winds, sponge (temperature {temp}) + {forcing}, 2 steps",
          flush=True)
    st = capi.ic_rest(g, p, seed=1007, temperature_noise=1.0, wind_noise=20.0)
    q = capi.SpongeParams(0.6, 2e-4, 1e-4, 16, 1, 1, temp, 3e-4)
    d = capi.Dyn(g, p, variant="pm")
    o = Oracle(g, p, variant="pm")
    for x in (d, o):
        x.set_state(st)
        if code == capi.FORCING_NEWTONIAN:
            x.set_newtonian(nr_params(capi))
        else:
            x.set_rt(rt_params(capi))
        x.set_sponge(q)
    d.run(2, code)
    d.sync()
    o.run(2, code)
    gs, os_ = d.get_state(), o.get_state()
    errs = {k: _max_rel(gs[k], os_[k]) for k in gs}
    print("sponge", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) == 0.0
    # the sponge did something: the top-level winds differ from a run without it
    d2 = capi.Dyn(g, p, variant="pm")
    d2.set_state(st)
    if code == capi.FORCING_NEWTONIAN:
        d2.set_newtonian(nr_params(capi))
    else:
        d2.set_rt(rt_params(capi))
    d2.run(2, code)
    d2.sync()
    assert _max_rel(d2.get_state()["Mh"], gs["Mh"]) > 1e-6
    d.close()
    d2.close()


@pytest.mark.parametrize("forcing,mode", [("newtonian", 1), ("double_gray", 3), ("newtonian", 3)])
def test_explicit_sponge_steps_bitwise(hj, forcing, mode):
    """Parity build with the explicit Rayleigh sponge (raysp_calc_mode exp1 /
    exp3: drag and heating into profx_dMh / profx_dWh / profx_Qheat in every
    RK stage, the zonal means of the first or of every stage), winds and
    temperature towards their zonal means: two steps equal the oracle bit for
    bit, and the mixed form stays within its usual distance of FP64."""
    from oracle.oracle import Oracle
    capi, g, p, st0 = hj
    code = capi.FORCING_NEWTONIAN if forcing == "newtonian" else capi.FORCING_DOUBLE_GRAY
    print(f"case: parity build, hot Jupiter g4 L20 with winds, explicit sponge exp{mode} + {forcing}, 2 steps",
          flush=True)
    st = capi.ic_rest(g, p, seed=1017, temperature_noise=1.0, wind_noise=20.0)
    q = capi.SpongeParams(0.6, 2e-4, 1e-4, 16, 1, 1, 1, 3e-4, mode)

    def run(x):
        x.set_state(st)
        if code == capi.FORCING_NEWTONIAN:
            x.set_newtonian(nr_params(capi))
        else:
            x.set_rt(rt_params(capi))
        x.set_sponge(q)
        x.run(2, code)
        return x

    d = run(capi.Dyn(g, p, variant="pm"))
    d.sync()
    o = run(Oracle(g, p, variant="pm"))
    gs, os_ = d.get_state(), o.get_state()
    errs = {k: _max_rel(gs[k], os_[k]) for k in gs}
    print("explicit sponge", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) == 0.0
    d.close()
    pm = capi.Params.from_buffer_copy(p)
    pm.precision = 1
    dm = run(capi.Dyn(g, pm))
    dm.sync()
    gm = dm.get_state()
    tol = {"Mh": 1e-4, "Wh": 1e-3, "W": 1e-3, "Rho": 1e-6, "pressure": 1e-6}
    for k, t in tol.items():
        err = _max_rel(gm[k], os_[k])
        print(f"  mixed vs FP64 {k}: {err:.2e}")
        assert err < t, k
    dm.close()


@pytest.mark.parametrize("mode", [0, 3])
def test_shrinking_sponge_steps_bitwise(hj, mode):
    """shrink_sponge from step 2 (implicit and exp3 modes, winds and
    temperature): three steps of the parity build equal the oracle bit for bit."""
    from oracle.oracle import Oracle
    capi, g, p, st0 = hj
    print(f"case: parity build, hot Jupiter g4 L20, shrinking sponge (mode {mode}) + newtonian, 3 steps", flush=True)
    st = capi.ic_rest(g, p, seed=1019, temperature_noise=1.0, wind_noise=20.0)
    q = capi.SpongeParams(0.6, 2e-4, 1e-4, 16, 1, 1, 1, 3e-4, mode, 1, 2)
    d = capi.Dyn(g, p, variant="pm")
    o = Oracle(g, p, variant="pm")
    for x in (d, o):
        x.set_state(st)
        x.set_newtonian(nr_params(capi))
        x.set_sponge(q)
        x.run(3, capi.FORCING_NEWTONIAN)
    d.sync()
    gs, os_ = d.get_state(), o.get_state()
    errs = {k: _max_rel(gs[k], os_[k]) for k in gs}
    print("shrinking sponge", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) == 0.0
    d.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_all_switches_together_bitwise(built, mode, hj):
    """Every optional part of the step at once: sixth-order hyperdiffusion,
    vertical hyperdiffusion, the fourth-order diffusive sponge, and the
    Rayleigh sponge with its temperature part and shrinking (implicit or
    exp1), under the Newtonian forcing. Two steps of the parity build equal
    the oracle bit for bit; the mixed form stays within its usual distance
    of FP64."""
    from oracle.oracle import Oracle
    capi, g, _, _ = hj
    p = hj_params(capi, 4, HyDiffOrder=6, VertHyDiff=1, VertHyDiffOrder=6, Diffc_v=1e-3, DiffSponge=1,
                  order_diff_sponge=4, Dv_sponge=0.02, ns_diff_sponge=0.6)
    print(f"case: parity build, hot Jupiter g4 L20, all switches (sponge mode {mode}), 2 steps", flush=True)
    st = capi.ic_rest(g, p, seed=1023, temperature_noise=1.0, wind_noise=20.0)
    q = capi.SpongeParams(0.6, 2e-4, 1e-4, 16, 1, 1, 1, 3e-4, mode, 1, 2)

    def run(x):
        x.set_state(st)
        x.set_newtonian(nr_params(capi))
        x.set_sponge(q)
        x.run(2, capi.FORCING_NEWTONIAN)
        return x

    d = run(capi.Dyn(g, p, variant="pm"))
    d.sync()
    o = run(Oracle(g, p, variant="pm"))
    gs, os_ = d.get_state(), o.get_state()
    errs = {k: _max_rel(gs[k], os_[k]) for k in gs}
    print("all switches", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) == 0.0
    d.close()
    pm = capi.Params.from_buffer_copy(p)
    pm.precision = 1
    dm = run(capi.Dyn(g, pm))
    dm.sync()
    gm = dm.get_state()
    for k, t in {"Mh": 1e-4, "Wh": 1e-3, "W": 1e-3, "Rho": 1e-6, "pressure": 1e-6}.items():
        err = _max_rel(gm[k], os_[k])
        print(f"  mixed vs FP64 {k}: {err:.2e}")
        assert err < t, k
    dm.close()


@pytest.mark.parametrize("sponge_order", [4, 6])
def test_mixed_order_eight_hot_jupiter(hj, sponge_order):
    """HyDiffOrder 8 on the large planet (A = 1e8 m) in the mixed form: two
    intermediate passes scale their FP32 output by dbar^2 (alternating the two
    arrays), the last pass divides its sign by dbar^4, and a sixth-order
    diffusive sponge reads the second pass's scaled output (dsp_scale =
    dbar^-2). The FP64 parity build equals the oracle bit for bit; the mixed
    form stays within the usual mixed-vs-FP64 distance of it.
    Diffc 5e-5: with the default 4th-order strength (Diffc ~ 1e-2) the eighth-
    order operator amplifies grid-scale noise (each Laplacian pass multiplies
    it by up to ~12 / dx^2), and both precision forms blow up within two
    steps."""
    from oracle.oracle import Oracle
    capi, g, _, _ = hj
    p = hj_params(capi, 4, HyDiffOrder=8, Diffc=5e-5, DiffSponge=1, order_diff_sponge=sponge_order,
                  Dv_sponge=0.02, ns_diff_sponge=0.6)
    print(f"case: hot Jupiter g4 L20, HyDiffOrder 8, order_diff_sponge {sponge_order}, 2 steps, "
          f"parity build vs oracle and mixed vs oracle", flush=True)
    st = capi.ic_rest(g, p, seed=1029, temperature_noise=1.0, wind_noise=20.0)

    def run(x):
        x.set_state(st)
        x.set_newtonian(nr_params(capi))
        x.run(2, capi.FORCING_NEWTONIAN)
        return x

    d = run(capi.Dyn(g, p, variant="pm"))
    d.sync()
    o = run(Oracle(g, p, variant="pm"))
    gs, os_ = d.get_state(), o.get_state()
    assert all(np.isfinite(os_[k]).all() for k in os_), "the oracle's run left finite values"
    errs = {k: _max_rel(gs[k], os_[k]) for k in gs}
    print("order 8 FP64", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) == 0.0
    d.close()
    pm = capi.Params.from_buffer_copy(p)
    pm.precision = 1
    dm = run(capi.Dyn(g, pm))
    dm.sync()
    gm = dm.get_state()
    for k, t in {"Mh": 1e-4, "Wh": 1e-3, "W": 1e-3, "Rho": 1e-6, "pressure": 1e-6}.items():
        err = _max_rel(gm[k], os_[k])
        print(f"  mixed vs FP64 {k}: {err:.2e}")
        assert err < t, k
    dm.close()


def test_sponge_many_slot_blocks_bitwise(built):
    """Two latitude bands on g5 (about 5 100 points each, 8 192 slots): the
    zonal sums (winds and temperature) run over four 2 048-slot blocks per
    row, added in order. One Newtonian ProfX + Thor step of the parity build
    equals the oracle."""
    from oracle.oracle import Oracle
    capi = built
    g = hj_grid(capi, 5, 20)
    p = hj_params(capi, 5)
    print("case: parity build, hot Jupiter g5 L20 with winds, sponge over 2 bands, 1 step", flush=True)
    st = capi.ic_rest(g, p, seed=1009, temperature_noise=1.0, wind_noise=20.0)
    q = capi.SpongeParams(0.5, 3e-4, 2e-4, 2, 1, 1, 1, 2e-4)
    d = capi.Dyn(g, p, variant="pm")
    o = Oracle(g, p, variant="pm")
    for x in (d, o):
        x.set_state(st)
        x.set_newtonian(nr_params(capi))
        x.set_sponge(q)
    _, mc = o.zonal_table()
    assert mc == 8192
    d.run(1, capi.FORCING_NEWTONIAN)
    d.sync()
    o.run(1, capi.FORCING_NEWTONIAN)
    gs, os_ = d.get_state(), o.get_state()
    errs = {k: _max_rel(gs[k], os_[k]) for k in gs}
    print("sponge, 2 bands", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) == 0.0
    d.close()


@pytest.mark.parametrize("forcing,ghv", [(1, 0), (1, 1), (2, 0), (3, 0)])
def test_boundary_layer_steps_bitwise(built, forcing, ghv):
    """THOR's boundary-layer module (RayleighHS) after the ProfX forcing, on
    the pressure ProfX started from: fused into the Held-Suarez and Newtonian
    ProfX kernels, a kernel of its own before the double-gray RT. The parity
    build equals the oracle bit for bit over two ProfX + Thor steps."""
    from oracle.oracle import Oracle
    from common import earth_params, grid
    capi = built
    if forcing == 1:
        g = grid(capi, 4, 20)
        p = earth_params(capi, 4, dt=600.0, GravHeightVar=ghv)
        st = capi.ic_rest(g, p, seed=1021, temperature_noise=1.0, wind_noise=1.0)
    else:
        g = hj_grid(capi, 4, 20)
        p = hj_params(capi, 4, GravHeightVar=ghv)
        st = capi.ic_rest(g, p, seed=1022, temperature_noise=1.0, wind_noise=0.1)
    print(f"case: parity build g4 L20, forcing {forcing}, GravHeightVar {ghv}, boundary layer, 2 steps", flush=True)
    bl = capi.BLParams(surf_drag=1.0 / 86400.0, bl_sigma=0.7)
    d = capi.Dyn(g, p, variant="pm")
    o = Oracle(g, p, variant="pm")
    for x in (d, o):
        x.set_state(st)
        x.set_boundary_layer(bl)
        if forcing == 2:
            x.set_newtonian(nr_params(capi))
        if forcing == 3:
            x.set_rt(rt_params(capi))
    d.run(2, forcing)
    d.sync()
    o.run(2, forcing)
    gs, os_ = d.get_state(), o.get_state()
    errs = {k: _max_rel(gs[k], os_[k]) for k in gs}
    print("boundary layer", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) == 0.0
    # the module acted: the same steps without it differ
    d2 = capi.Dyn(g, p, variant="pm")
    d2.set_state(st)
    if forcing == 2:
        d2.set_newtonian(nr_params(capi))
    if forcing == 3:
        d2.set_rt(rt_params(capi))
    d2.run(2, forcing)
    d2.sync()
    assert _max_rel(d2.get_state()["Mh"], gs["Mh"]) > 1e-8
    d.close()
    d2.close()
