"""Mixed-precision path: operator weights and one step against FP64.

The mixed path applies the hyperdiffusion Laplacian and the divergence-damping
operator as FP32 per-point weights derived on the GPU from the FP64 operators.
Applied in FP64 with numpy, those weights must reproduce the oracle's own
kernels (which are pinned bit for bit to THOR's Diffusion_Op and DivM_Op) to
the weights' FP32 rounding. Then a full step in mixed precision must stay
within FP32-sized distances of the FP64 step.
"""
from __future__ import annotations

import numpy as np
import pytest

from common import earth_params, grid

pytestmark = pytest.mark.gpu


def _ring(g):
    """Stencil slots (0 = self, 1..6 = ring) per point; a pentagon's or pole's
    sixth slot repeats the fifth (its weight is zero)."""
    P = g.P
    ring = np.asarray(g.point_local).reshape(P, 6).copy()
    for p in g.pent_ind:
        ring[p, 5] = ring[p, 4]
    return np.concatenate([np.arange(P)[:, None], ring], axis=1)


@pytest.fixture(scope="module")
def setup(built):
    capi = built
    g = grid(capi, 4, 16)
    p = earth_params(capi, 4, dt=600.0, precision=1)
    d = capi.Dyn(g, p)
    w = {n: d.get_field(n) for n in ("wlap", "wdd", "udd")}
    d.close()
    return capi, g, p, w


def test_laplacian_weights_reproduce_oracle_diffusion(setup):
    from oracle.oracle import Oracle
    capi, g, p, w = setup
    P, nv = g.P, g.nv
    rng = np.random.default_rng(11)
    rho = 1.0 + 0.1 * rng.standard_normal((P, nv))
    o = Oracle(g, p)
    o.set("Rhok", rho)
    o.kernel("diffusion", 0, 0)
    want = o.get("diff").reshape(P, nv, 6)[:, :, 0]  # var 0: Laplacian of rho
    slots = _ring(g)
    wl = w["wlap"].reshape(P, 7)
    rs = p.A / (p.A + np.asarray(g.Altitude))
    got = np.einsum("ps,psl->pl", wl[:, 1:], rho[slots[:, 1:], :] - rho[:, None, :]) * rs[None, :] ** 2
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"Laplacian from weights vs oracle: max rel diff {err:.2e}")
    assert err < 2e-6


def test_divergence_damping_weights_reproduce_oracle_divm(setup):
    from oracle.oracle import Oracle
    capi, g, p, w = setup
    P, nv = g.P, g.nv
    rng = np.random.default_rng(12)
    M = rng.standard_normal((P, nv, 3))
    Wh = 0.1 * rng.standard_normal((P, nv + 1))
    o = Oracle(g, p)
    o.set("Mhk", M)
    o.set("Whk", Wh)
    o.kernel("divm", 0, 0)
    want = o.get("divg_Mh").reshape(P, nv, 3)
    slots = _ring(g)
    W = w["wdd"].reshape(P, 3, 7, 3)
    U = w["udd"].reshape(P, 3, 7)
    alt, alth = np.asarray(g.Altitude), np.asarray(g.Altitudeh)
    rs = p.A / (p.A + alt)
    dz = 1.0 / (alth[:-1] - alth[1:])
    Ms = M[slots]                                  # [P, 7, nv, 3]
    dw = (Wh[:, :-1] - Wh[:, 1:])[slots]           # [P, 7, nv]
    got = np.einsum("pcsk,pslk->plc", W, Ms) * (rs ** 2)[None, :, None] \
        + np.einsum("pcs,psl->plc", U, dw) * (rs * dz)[None, :, None]
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"divergence damping from weights vs oracle: max rel diff {err:.2e}")
    assert err < 2e-6


def test_mixed_step_stays_close_to_fp64(setup):
    capi, g, p64, _ = setup
    st = capi.ic_rest(g, p64, seed=1001, temperature_noise=1e-3, wind_noise=1.0)
    out = {}
    for prec in (0, 1):
        p = earth_params(capi, 4, dt=600.0, precision=prec)
        d = capi.Dyn(g, p)
        d.set_state(st)
        d.run(3)
        d.sync()
        out[prec] = d.get_state()
        d.close()
    # FP32 storage of the momentum deviation and FP32 damping operators move
    # the state by FP32-sized amounts; density and pressure stay FP64-exact
    # to that level too.
    tol = {"Mh": 1e-4, "Wh": 1e-3, "W": 1e-3, "Rho": 1e-6, "pressure": 1e-6}
    for k, t in tol.items():
        a, b = out[1][k], out[0][k]
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"3 steps mixed vs FP64 {k}: max rel diff {err:.2e}")
        assert np.isfinite(a).all()
        assert err < t, k


def test_mixed_order_six_stays_close_to_fp64(built):
    """HyDiffOrder 6: the intermediate pass runs on the FP32 weights in the
    mixed form; three steps stay within the order-4 mixed-vs-FP64 bounds."""
    capi = built
    g = grid(capi, 4, 16)
    print("case: g4 L16, HyDiffOrder 6, 3 steps mixed vs FP64", flush=True)
    p64 = earth_params(capi, 4, dt=600.0, HyDiffOrder=6)
    st = capi.ic_rest(g, p64, seed=1001, temperature_noise=1e-3, wind_noise=1.0)
    out = {}
    for prec in (0, 1):
        d = capi.Dyn(g, earth_params(capi, 4, dt=600.0, precision=prec, HyDiffOrder=6))
        d.set_state(st)
        d.run(3)
        d.sync()
        out[prec] = d.get_state()
        d.close()
    tol = {"Mh": 1e-4, "Wh": 1e-3, "W": 1e-3, "Rho": 1e-6, "pressure": 1e-6}
    for k, t in tol.items():
        a, b = out[1][k], out[0][k]
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"3 steps mixed vs FP64, HyDiffOrder 6, {k}: max rel diff {err:.2e}")
        assert np.isfinite(a).all()
        assert err < t, k


def test_mixed_vertical_hyperdiffusion_stays_close_to_fp64(built):
    """VertHyDiff: the vertical terms are formed in FP64 and added to the FP32
    horizontal tendencies in the mixed form; three steps stay within the
    mixed-vs-FP64 bounds, and the vertical terms do change the state."""
    capi = built
    g = grid(capi, 4, 16)
    print("case: g4 L16, VertHyDiff order 6, 3 steps mixed vs FP64", flush=True)
    kw = dict(VertHyDiff=1, VertHyDiffOrder=6, Diffc_v=2e-2)
    p64 = earth_params(capi, 4, dt=600.0, **kw)
    st = capi.ic_rest(g, p64, seed=1001, temperature_noise=1.0, wind_noise=1.0)
    out = {}
    for prec, vert in ((0, 1), (1, 1), (0, 0)):
        d = capi.Dyn(g, earth_params(capi, 4, dt=600.0, precision=prec, **dict(kw, VertHyDiff=vert)))
        d.set_state(st)
        d.run(3)
        d.sync()
        out[prec, vert] = d.get_state()
        d.close()
    tol = {"Mh": 1e-4, "Wh": 1e-3, "W": 1e-3, "Rho": 1e-6, "pressure": 1e-6}
    for k, t in tol.items():
        a, b = out[1, 1][k], out[0, 1][k]
        err = np.abs(a - b).max() / np.abs(b).max()
        eff = np.abs(out[0, 0][k] - b).max() / np.abs(b).max()
        print(f"3 steps mixed vs FP64, VertHyDiff, {k}: max rel diff {err:.2e} (vertical terms move it by {eff:.2e})")
        assert np.isfinite(a).all()
        assert err < t, k
    assert np.abs(out[0, 0]["Mh"] - out[0, 1]["Mh"]).max() > 0


@pytest.mark.parametrize("chunked", ["1", "0"])
def test_mixed_diffusive_sponge_stays_close_to_fp64(built, chunked):
    """The diffusive sponge in the mixed form's last hyperdiffusion pass (the
    level-chunked kernel and the one-level-per-block one): three steps within
    the mixed-vs-FP64 bounds, and the sponge does change the state."""
    import os
    capi = built
    g = grid(capi, 4, 16)
    print(f"case: g4 L16, DiffSponge order 2, THOR_DIFF2_CHUNKED={chunked}, 3 steps mixed vs FP64", flush=True)
    kw = dict(DiffSponge=1, order_diff_sponge=2, Dv_sponge=0.05, ns_diff_sponge=0.33)
    st = capi.ic_rest(g, earth_params(capi, 4, dt=600.0), seed=1001, temperature_noise=1.0, wind_noise=1.0)
    out = {}
    old = os.environ.get("THOR_DIFF2_CHUNKED")
    os.environ["THOR_DIFF2_CHUNKED"] = chunked
    try:
        for prec, on in ((0, 1), (1, 1), (0, 0)):
            d = capi.Dyn(g, earth_params(capi, 4, dt=600.0, precision=prec, **dict(kw, DiffSponge=on)))
            d.set_state(st)
            d.run(3)
            d.sync()
            out[prec, on] = d.get_state()
            d.close()
    finally:
        if old is None:
            os.environ.pop("THOR_DIFF2_CHUNKED", None)
        else:
            os.environ["THOR_DIFF2_CHUNKED"] = old
    tol = {"Mh": 1e-4, "Wh": 1e-3, "W": 1e-3, "Rho": 1e-6, "pressure": 1e-6}
    for k, t in tol.items():
        a, b = out[1, 1][k], out[0, 1][k]
        err = np.abs(a - b).max() / np.abs(b).max()
        eff = np.abs(out[0, 0][k] - b).max() / np.abs(b).max()
        print(f"3 steps mixed vs FP64, DiffSponge, {k}: max rel diff {err:.2e} (the sponge moves it by {eff:.2e})")
        assert np.isfinite(a).all()
        assert err < t, k
    assert np.abs(out[0, 0]["Mh"] - out[0, 1]["Mh"]).max() > 0


def _two_steps(capi, g, p, st, env):
    """Two HS steps of a Dyn created with the given tuning variables set."""
    import os
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        d = capi.Dyn(g, p)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    d.set_state(st)
    d.run(2, capi.FORCING_HELD_SUAREZ)
    d.sync()
    out = d.get_state()
    d.close()
    return out


@pytest.mark.parametrize("deep,nonhydro,ghv", [(1, 1, 0), (0, 0, 1), (1, 0, 1)])
def test_chunked_slow_modes_equal_per_level_kernel(built, deep, nonhydro, ghv):
    """The level-chunked slow-mode kernels (DivM pass, then the slow
    tendencies with the gradient weights serving the divergence) reproduce
    the one-level-per-block kernel bit for bit over two steps (all three RK
    stages, so both the zero start and the state deviations of the stage),
    when the substep kernel forms the whole pressure flux divergence itself."""
    capi = built
    g = grid(capi, 4, 20)
    p = earth_params(capi, 4, dt=600.0, precision=1, HyDiff=1, DivDampP=1, DeepModel=deep, NonHydro=nonhydro,
                     GravHeightVar=ghv)
    print(f"case: mixed g4 L20, DeepModel {deep} NonHydro {nonhydro} GravHeightVar {ghv}: chunked vs per-level "
          f"slow modes", flush=True)
    st = capi.ic_rest(g, p, seed=1011, temperature_noise=1.0, wind_noise=2.0)
    a = _two_steps(capi, g, p, st, {"THOR_SLOW_CHUNKED": "0"})
    # both chunked kernels: k_slow_lc and k_slow_lc2 (two blocks per CU)
    for v2 in ("0", "1"):
        b = _two_steps(capi, g, p, st, {"THOR_PREPARE_SPLIT": "0", "THOR_SLOW_V2": v2})
        for k in a:
            assert np.array_equal(a[k], b[k]), (v2, k)
    # Default: the substep pressure flux divergence adds the stage part
    # (formed by the slow-mode kernel) to the deviation part, whose h and pt
    # factors are read as FP32 copies (they multiply the FP32 momentum
    # deviation); the same sums grouped differently with one more FP32
    # rounding, so equal to ~1e-8 carried through two steps, far below the
    # mixed form's distance to FP64 (1e-4 in Mh).
    c = _two_steps(capi, g, p, st, {})
    for k in a:
        err = float(np.abs(c[k] - a[k]).max() / np.abs(a[k]).max())
        print(f"  split vs unsplit pressure flux, {k}: max rel diff {err:.2e}")
        assert err < 1e-6, k


def test_thomas_loop_unroll_is_bitwise_unchanged(built):
    """The fused vertical-solve kernel with its level loops unrolled by two
    (the default) gives the same states as without unrolling, in both
    precision forms."""
    capi = built
    g = grid(capi, 4, 20)
    for prec in (0, 1):
        p = earth_params(capi, 4, dt=600.0, precision=prec, HyDiff=1, DivDampP=1)
        print(f"case: precision {prec} g4 L20, Thomas level loops unrolled by 2 vs 1", flush=True)
        st = capi.ic_rest(g, p, seed=1012, temperature_noise=1.0, wind_noise=2.0)
        a = _two_steps(capi, g, p, st, {"THOR_THOMAS_UNROLL": "1"})
        b = _two_steps(capi, g, p, st, {"THOR_THOMAS_UNROLL": "2"})
        for k in a:
            assert np.array_equal(a[k], b[k]), (prec, k)


def test_chunked_last_diffusion_pass_stays_close(built):
    """The last hyperdiffusion pass and the stage's divergence-damping pass 0
    over level chunks (FP32 with fused multiply-adds) against the
    one-level-per-block kernel: two steps agree to FP32 rounding of those
    tendencies, far inside the mixed-vs-FP64 distances."""
    capi = built
    g = grid(capi, 4, 20)
    p = earth_params(capi, 4, dt=600.0, precision=1, HyDiff=1, DivDampP=1)
    print("case: mixed g4 L20, chunked vs per-level last diffusion pass + divergence-damping pass 0", flush=True)
    st = capi.ic_rest(g, p, seed=1013, temperature_noise=1.0, wind_noise=2.0)
    a = _two_steps(capi, g, p, st, {"THOR_DIFF2_CHUNKED": "0"})
    b = _two_steps(capi, g, p, st, {})
    tol = {"Mh": 1e-5, "Wh": 1e-5, "W": 1e-5, "Rho": 1e-7, "pressure": 1e-7}
    for k in a:
        err = float(np.abs(b[k] - a[k]).max() / np.abs(a[k]).max())
        print(f"  {k}: max rel diff {err:.2e}")
        assert err < tol[k], k


@pytest.mark.parametrize("deep,nonhydro,ghv", [(1, 1, 0), (0, 0, 1)])
def test_column_prep_level_window_is_bitwise_unchanged(built, deep, nonhydro, ghv):
    """The column-preparation kernel with 2 or 4 levels per thread (per-level
    values formed once into a window) gives the same states as with one
    level per thread, in both precision forms."""
    capi = built
    g = grid(capi, 4, 20)
    for prec in (0, 1):
        p = earth_params(capi, 4, dt=600.0, precision=prec, HyDiff=1, DivDampP=1, DeepModel=deep,
                         NonHydro=nonhydro, GravHeightVar=ghv)
        print(f"case: precision {prec} g4 L20, DeepModel {deep} NonHydro {nonhydro} GravHeightVar {ghv}: column "
              f"preparation with 1, 2, 4 levels per thread", flush=True)
        st = capi.ic_rest(g, p, seed=1014, temperature_noise=1.0, wind_noise=2.0)
        a = _two_steps(capi, g, p, st, {"THOR_COLPREP_LEVELS": "1"})
        for lv in ("2", "4"):
            b = _two_steps(capi, g, p, st, {"THOR_COLPREP_LEVELS": lv})
            for k in a:
                assert np.array_equal(a[k], b[k]), (prec, lv, k)


def test_row_parallel_vertical_sweeps_stay_close(built):
    """The mixed path's vertical solve with both sweeps on all rows of the
    block (segment maps chained through LDS) against the one-lane sweeps:
    two steps agree to FP64 rounding."""
    capi = built
    for nv in (20, 48):
        g = grid(capi, 4, nv)
        p = earth_params(capi, 4, dt=600.0, precision=1, HyDiff=1, DivDampP=1)
        print(f"case: mixed g4 L{nv}, row-parallel vs one-lane vertical sweeps", flush=True)
        st = capi.ic_rest(g, p, seed=1017, temperature_noise=1.0, wind_noise=2.0)
        a = _two_steps(capi, g, p, st, {"THOR_THOMAS_ROWS": "0"})
        b = _two_steps(capi, g, p, st, {})
        for k in a:
            err = float(np.abs(b[k] - a[k]).max() / np.abs(a[k]).max())
            print(f"  L{nv} {k}: max rel diff {err:.2e}")
            assert err < 1e-9, (nv, k)


@pytest.mark.parametrize("nv,deep,nonhydro,ghv", [(20, 1, 1, 0), (33, 0, 0, 1), (48, 1, 1, 1), (64, 1, 0, 0)])
def test_segment_vertical_solve_stays_close(built, nv, deep, nonhydro, ghv):
    """The segment form of the mixed vertical solve (each row of the block
    owns a contiguous run of interfaces in all three phases and keeps its
    cells' values in registers for the density/pressure update), with 16 and
    32 columns per block, and the segment form of the stage's factorisation
    (Moebius maps of the pivot recurrence chained over the rows), against the
    one-lane sweeps and recurrence: two steps agree to FP64 rounding,
    including ragged last segments and rows without interfaces."""
    capi = built
    g = grid(capi, 4, nv)
    p = earth_params(capi, 4, dt=600.0, precision=1, HyDiff=1, DivDampP=1, DeepModel=deep, NonHydro=nonhydro,
                     GravHeightVar=ghv)
    st = capi.ic_rest(g, p, seed=1018, temperature_noise=1.0, wind_noise=2.0)
    a = _two_steps(capi, g, p, st, {"THOR_THOMAS_ROWS": "0", "THOR_THOMAS_SEG": "0", "THOR_FACTOR_SEG": "0"})
    for cb in ("16", "32"):
        print(f"case: mixed g4 L{nv}, DeepModel {deep} NonHydro {nonhydro} GravHeightVar {ghv}: segment vertical "
              f"solve, {cb} columns per block, vs one-lane sweeps", flush=True)
        b = _two_steps(capi, g, p, st, {"THOR_THOMAS_SEG": cb, "THOR_FACTOR_SEG": "1"})
        for k in a:
            err = float(np.abs(b[k] - a[k]).max() / np.abs(a[k]).max())
            print(f"  L{nv} cb {cb} {k}: max rel diff {err:.2e}")
            assert err < 1e-9, (nv, cb, k)


@pytest.mark.parametrize("prec,extra", [(1, {}), (0, {}), (1, dict(VertHyDiff=1, VertHyDiffOrder=6, Diffc_v=2e-3)),
                                        (0, dict(thermo_equation=1))])
def test_stage_overlap_is_bitwise_unchanged(built, prec, extra):
    """The stage's divergence-damping passes (mixed form) and the stage
    factorisation on a second stream beside the hyperdiffusion chain give
    the same states bit for bit as everything on one stream."""
    capi = built
    g = grid(capi, 4, 20)
    p = earth_params(capi, 4, dt=600.0, precision=prec, HyDiff=1, DivDampP=1, **extra)
    print(f"case: precision {prec} g4 L20 {extra}: second stream vs one stream, 2 steps", flush=True)
    st = capi.ic_rest(g, p, seed=1019, temperature_noise=1.0, wind_noise=2.0)
    a = _two_steps(capi, g, p, st, {"THOR_STAGE_OVERLAP": "0"})
    b = _two_steps(capi, g, p, st, {"THOR_STAGE_OVERLAP": "1"})
    for k in a:
        assert np.array_equal(a[k], b[k]), (prec, k)
