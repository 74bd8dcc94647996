"""Physics-module hooks and RK-advanced module arrays on the GPU.

THOR calls its physics modules at four points of the dynamics step
(phy_modules.h:76-86; thor_driver.cu:116, 641, 1290, 1465) and advances the
arrays they register like the state (UpdateRK / UpdateRK2,
thor_auxiliary.h:376-387, 482-494). Checked here:
  * the call sequence and the `times` argument of every call;
  * a hook between the vertical solve and the density/pressure update
    (which splits the fused column kernel in two) leaves the step bit for
    bit unchanged;
  * a module array with a constant tendency S added in every acoustic
    substep ends the step at x0 + S * dt exactly as the RK3 combination
    implies (stage weights 1/3, 1/2 - 1/3 ... sum to one step), for both
    precision modes. Closed form: the stage deviations start at 0, then
    x_i - x_k, so after stage 3 x_k = x0 + S * (ns * dt / ns) = x0 + S dt.
"""
from __future__ import annotations

import numpy as np
import pytest

from common import earth_params, grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(built):
    capi = built
    g = grid(capi, 4, 20)
    return capi, g


def _state(capi, g, p):
    return capi.ic_rest(g, p, seed=1004, temperature_noise=1e-3, wind_noise=1.0)


def test_hook_sequence_and_times(setup):
    capi, g = setup
    ns, dt = 6, 600.0
    print(f"case: FP64 g4 L20, ns={ns}, dt={dt}: hook call sequence over 2 steps", flush=True)
    p = earth_params(capi, 4, dt=dt, ns=ns)
    d = capi.Dyn(g, p)
    d.set_state(_state(capi, g, p))
    calls = []

    def rec(kind):
        if kind in ("init", "end"):
            def f(stream, nstep, view):
                assert stream != 0 and view.nv == g.nv and view.point_num == g.P and view.mhs_fp32 == 0
                calls.append((kind, nstep, None))
                return 0
        else:
            def f(stream, nstep, times, view):
                assert view.Whs and view.Rhok and view.Mhs
                calls.append((kind, nstep, times))
                return 0
        return f

    d.set_hooks(init=rec("init"), slow_modes=rec("slow"), fast_modes=rec("fast"), end=rec("end"))
    d.thor()
    d.thor()
    d.sync()
    want = []
    for step in (1, 2):
        want.append(("init", step, None))
        for rk, nit in enumerate((1, ns // 2, ns)):
            times = dt / 3.0 if rk == 0 else dt / ns
            want.append(("slow", step, times))
            want += [("fast", step, times)] * nit
        want.append(("end", step, None))
    assert calls == want
    d.close()


def test_hook_split_step_is_bitwise_unchanged(setup):
    capi, g = setup
    print("case: FP64 and mixed g4 L20, 2 steps with a no-op fast-mode hook vs none", flush=True)
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        p = earth_params(capi, 4, dt=600.0, ns=6, precision=prec)
        st = _state(capi, g, p)
        a, b = capi.Dyn(g, p), capi.Dyn(g, p)
        a.set_state(st)
        b.set_state(st)
        b.set_hooks(fast_modes=lambda stream, nstep, times, view: 0)
        for d in (a, b):
            d.run(2, capi.FORCING_HELD_SUAREZ)
            d.sync()
        sa, sb = a.get_state(), b.get_state()
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), (prec, k)
        a.close()
        b.close()


def _hip():
    """The HIP runtime the library itself links (device buffers for the test module)."""
    import ctypes as C
    import os
    hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipMemsetAsync.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    return hip


def test_rk_array_constant_tendency(setup):
    """A module in the style of THOR's chemistry module (chemistry.cu:335-345,
    494-502): init copies tracer -> tracerk and clears the deviation, the
    fast-mode hook adds S * times to the deviation, end copies tracerk back."""
    import ctypes as C
    capi, g = setup
    hip = _hip()
    H2D, D2H, D2D = 1, 2, 3
    ns, dt, S, dims = 6, 600.0, 2.5e-3, 2
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        print(f"case: precision {prec}, g4 L20: module array with constant tendency S = {S}", flush=True)
        p = earth_params(capi, 4, dt=dt, ns=ns, precision=prec)
        d = capi.Dyn(g, p)
        d.set_state(_state(capi, g, p))
        n = g.P * g.nv * dims
        nb = n * 8
        x0 = np.linspace(-1.0, 1.0, n)
        bufs = {}
        for k in ("tracer", "tracerk", "tracers"):
            ptr = C.c_void_p()
            assert hip.hipMalloc(C.byref(ptr), nb) == 0
            bufs[k] = ptr.value
        assert hip.hipMemcpy(bufs["tracer"], x0.ctypes.data, nb, H2D) == 0
        d.register_rk_array(bufs["tracers"], bufs["tracerk"], bufs["tracer"], dims)
        host = np.empty(n)

        def init(stream, nstep, view):
            hip.hipMemcpyAsync(bufs["tracerk"], bufs["tracer"], nb, D2D, stream)
            return hip.hipMemsetAsync(bufs["tracers"], 0, nb, stream)

        def fast(stream, nstep, times, view):
            hip.hipStreamSynchronize(stream)
            hip.hipMemcpy(host.ctypes.data, bufs["tracers"], nb, D2H)
            host[:] += S * times
            return hip.hipMemcpy(bufs["tracers"], host.ctypes.data, nb, H2D)

        def end(stream, nstep, view):
            return hip.hipMemcpyAsync(bufs["tracer"], bufs["tracerk"], nb, D2D, stream)

        d.set_hooks(init=init, fast_modes=fast, end=end)
        out = np.empty(n)
        for step in (1, 2):
            d.thor()
            d.sync()
            assert hip.hipMemcpy(out.ctypes.data, bufs["tracer"], nb, D2H) == 0
            err = np.abs(out - (x0 + step * S * dt)).max()
            print(f"  after {step} step(s): max |x - (x0 + n S dt)| = {err:.2e}")
            assert err < 1e-12
        d.clear_hooks()
        d.clear_rk_arrays()
        d.close()
        for v in bufs.values():
            hip.hipFree(v)


def test_hook_error_stops_the_step(setup):
    capi, g = setup
    print("case: FP64 g4 L20: a slow-mode hook that fails", flush=True)
    p = earth_params(capi, 4)
    d = capi.Dyn(g, p)
    d.set_state(_state(capi, g, p))
    d.set_hooks(slow_modes=lambda stream, nstep, times, view: 1)
    with pytest.raises(capi.ThorError) as e:
        d.thor()
    assert e.value.code == capi.THOR_ERR_HOOK
    d.sync()
    d.close()
