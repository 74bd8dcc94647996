"""The FP64 form's multi-level horizontal stencils (one block per tile and
NL levels, each triangle's geometry loaded once for all NL levels) against
the per-level kernels (THOR_FP64_MULTILEVEL=0) and against the oracle.

Level counts that NL does not divide (23 = 5 x 4 + 3 = 11 x 2 + 1) exercise
the last chunk's clamped levels, which are read but never written.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from common import STATE, earth_params, grid

pytestmark = pytest.mark.gpu


def _run(capi, g, p, st, steps, variant="", multilevel=True):
    old = os.environ.get("THOR_FP64_MULTILEVEL")
    os.environ["THOR_FP64_MULTILEVEL"] = "1" if multilevel else "0"
    try:
        d = capi.Dyn(g, p, variant=variant)  # the switch is read when the handle is created
    finally:
        if old is None:
            del os.environ["THOR_FP64_MULTILEVEL"]
        else:
            os.environ["THOR_FP64_MULTILEVEL"] = old
    d.set_state(st)
    d.run(steps)
    d.sync()
    out = d.get_state()
    d.close()
    return out


@pytest.mark.parametrize("nv,deep,nonhydro,order", [(20, 1, 1, 4), (23, 1, 1, 4), (23, 0, 0, 4), (21, 1, 1, 6)])
def test_multilevel_equals_per_level_kernels(built, nv, deep, nonhydro, order):
    """Production build, Held-Suarez g4, hyperdiffusion and divergence
    damping on: two steps with the multi-level kernels equal two steps with
    the per-level kernels bit for bit."""
    capi = built
    print(f"case: FP64 g4 L{nv}, DeepModel {deep}, NonHydro {nonhydro}, HyDiffOrder {order}, multi-level vs "
          f"per-level, 2 steps", flush=True)
    g = grid(capi, 4, nv)
    p = earth_params(capi, 4, dt=600.0, ns=6, DeepModel=deep, NonHydro=nonhydro, HyDiffOrder=order)
    st = capi.ic_rest(g, p, seed=1040 + nv, temperature_noise=1e-3, wind_noise=1.0)
    a = _run(capi, g, p, st, 2, multilevel=True)
    b = _run(capi, g, p, st, 2, multilevel=False)
    for k in STATE:
        assert np.isfinite(a[k]).all()
        assert np.array_equal(a[k], b[k]), k


def test_multilevel_parity_build_odd_levels_bitwise(built):
    """Parity build at g4 L23 (both chunk sizes leave a partial last chunk):
    two ProfX + Thor steps equal the oracle bit for bit."""
    from oracle.oracle import Oracle
    capi = built
    print("case: parity build g4 L23, HS, 2 steps, multi-level kernels vs oracle", flush=True)
    g = grid(capi, 4, 23)
    p = earth_params(capi, 4, dt=600.0, ns=6)
    st = capi.ic_rest(g, p, seed=1063, temperature_noise=1e-3, wind_noise=1.0)
    gs = _run(capi, g, p, st, 2, variant="pm", multilevel=True)
    o = Oracle(g, p, variant="pm")
    o.set_state(st)
    o.steps(2)
    os_ = o.get_state()
    for k in STATE:
        assert np.array_equal(gs[k], os_[k]), k


def _run_env(capi, g, p, st, steps, env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        d = capi.Dyn(g, p)  # the switches are read when the handle is created
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    d.set_state(st)
    d.run(steps)
    d.sync()
    out = d.get_state()
    d.close()
    return out


@pytest.mark.parametrize("nv,extra", [
    (20, dict(DeepModel=1, NonHydro=1)),
    (23, dict(DeepModel=0, NonHydro=0, GravHeightVar=1)),
    (21, dict(DeepModel=1, NonHydro=1, GravHeightVar=1, VertHyDiff=1, VertHyDiffOrder=6, Diffc_v=2e-3)),
    (20, dict(DeepModel=1, NonHydro=1, thermo_equation=1)),
    (20, dict(DeepModel=1, NonHydro=1, DivDampP=0)),
])
def test_chunked_slow_modes_equal_per_level_kernel(built, nv, extra):
    """FP64 slow modes over level chunks (divergence-damping pass 1 over four
    levels per block, then the slow tendencies with each point's gradient
    weights held for the chunk) against the one-level-per-block kernel: two
    steps equal bit for bit, with the optional terms (vertical momentum
    hyperdiffusion, energy equation, no divergence damping) and a partial
    last chunk."""
    capi = built
    print(f"case: FP64 g4 L{nv} {extra}: chunked vs per-level slow modes, 2 steps", flush=True)
    g = grid(capi, 4, nv)
    p = earth_params(capi, 4, dt=600.0, ns=6, **extra)
    st = capi.ic_rest(g, p, seed=1070 + nv, temperature_noise=1e-3, wind_noise=1.0)
    a = _run_env(capi, g, p, st, 2, {"THOR_SLOW_CHUNKED": "1"})
    b = _run_env(capi, g, p, st, 2, {"THOR_SLOW_CHUNKED": "0"})
    for k in STATE:
        assert np.isfinite(a[k]).all()
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("nv,extra", [
    (20, dict(DeepModel=1, NonHydro=1)),
    (23, dict(DeepModel=0, NonHydro=0, GravHeightVar=1)),
    (48, dict(DeepModel=1, NonHydro=1, thermo_equation=1)),
])
def test_segment_vertical_solve_equals_fp64_kernel(built, nv, extra):
    """FP64 vertical solve in segment form (16 columns per block, each row a
    run of interfaces and the cells below them, values kept for the
    density/pressure update, one-lane sweeps in the reference's order)
    against the 64-column kernel: two steps equal bit for bit, with a hook
    splitting the solve from the update as well."""
    capi = built
    print(f"case: FP64 g4 L{nv} {extra}: segment vs 64-column vertical solve, 2 steps", flush=True)
    g = grid(capi, 4, nv)
    p = earth_params(capi, 4, dt=600.0, ns=6, **extra)
    st = capi.ic_rest(g, p, seed=1080 + nv, temperature_noise=1e-3, wind_noise=1.0)
    a = _run_env(capi, g, p, st, 2, {"THOR_THOMAS_SEG64": "16"})
    b = _run_env(capi, g, p, st, 2, {"THOR_THOMAS_SEG64": "0"})
    for k in STATE:
        assert np.isfinite(a[k]).all()
        assert np.array_equal(a[k], b[k]), k
    # with a fast-mode hook the solve runs without the update, then k_dp
    old = os.environ.get("THOR_THOMAS_SEG64")
    os.environ["THOR_THOMAS_SEG64"] = "16"
    try:
        d = capi.Dyn(g, p)
    finally:
        if old is None:
            del os.environ["THOR_THOMAS_SEG64"]
        else:
            os.environ["THOR_THOMAS_SEG64"] = old
    d.set_hooks(fast_modes=lambda stream, nstep, times, view: 0)
    d.set_state(st)
    d.run(2)
    d.sync()
    c = d.get_state()
    d.close()
    for k in STATE:
        assert np.array_equal(a[k], c[k]), ("hook split", k)


@pytest.mark.parametrize("nv,extra,env", [
    (20, dict(DeepModel=1, NonHydro=1), {"THOR_PREPARE_LC64": "1"}),
    (23, dict(DeepModel=0, NonHydro=0, GravHeightVar=1), {"THOR_PREPARE_LC64": "1"}),
    (23, dict(DeepModel=1, NonHydro=1, thermo_equation=1), {"THOR_PREPARE_LC64": "1", "THOR_STENCIL_LEVELS": "5"}),
    (23, dict(DeepModel=1, NonHydro=1), {"THOR_ML_LEVELS": "2"}),
])
def test_chunked_prepare_equals_multilevel_kernel(built, nv, extra, env):
    """FP64 implicit-solve sources and flux divergences over level chunks
    (each point's divergence weights held for the chunk, the next level
    loaded while the current one is summed) against the two-level kernel:
    two steps equal bit for bit, with partial last chunks (and with the
    divergence-damping pass 0 and momentum over two levels per block)."""
    capi = built
    print(f"case: FP64 g4 L{nv} {extra} {env}: chunked vs two-level prepare, 2 steps", flush=True)
    g = grid(capi, 4, nv)
    p = earth_params(capi, 4, dt=600.0, ns=6, **extra)
    st = capi.ic_rest(g, p, seed=1090 + nv, temperature_noise=1e-3, wind_noise=1.0)
    a = _run_env(capi, g, p, st, 2, env)
    b = _run_env(capi, g, p, st, 2, {"THOR_PREPARE_LC64": "0"})
    for k in STATE:
        assert np.isfinite(a[k]).all()
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("nv,extra,env", [
    (20, dict(DeepModel=1, NonHydro=1), {"THOR_DIFF2_SPLIT64": "1"}),
    (23, dict(DeepModel=0, NonHydro=0, HyDiffOrder=6), {"THOR_DIFF2_SPLIT64": "1"}),
    (21, dict(DeepModel=1, NonHydro=1, DivDampP=0), {"THOR_DIFF2_SPLIT64": "1"}),
])
def test_split_last_diffusion_pass_equals_fused_kernel(built, nv, extra, env):
    """FP64 last hyperdiffusion pass and stage divergence-damping pass 0 as
    two kernels against the fused kernel: two steps equal bit for bit."""
    capi = built
    print(f"case: FP64 g4 L{nv} {extra} {env}: split vs fused last diffusion pass, 2 steps", flush=True)
    g = grid(capi, 4, nv)
    p = earth_params(capi, 4, dt=600.0, ns=6, **extra)
    st = capi.ic_rest(g, p, seed=1100 + nv, temperature_noise=1e-3, wind_noise=1.0)
    a = _run_env(capi, g, p, st, 2, env)
    b = _run_env(capi, g, p, st, 2, {"THOR_DIFF2_SPLIT64": "0"})
    for k in STATE:
        assert np.isfinite(a[k]).all()
        assert np.array_equal(a[k], b[k]), k
