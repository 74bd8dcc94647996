"""The oracle's restatement against THOR's own kernels, kernel by kernel.

THOR's kernels for advection/Coriolis, hyperdiffusion, divergence damping,
slow modes and Held-Suarez compile for gfx950 straight from the reference
tree (oracle/ref_kernels.hip, `make -C oracle ref` into oracle/_ref/). Both
sides get the same arrays in THOR's layout, taken from an oracle state two
steps into a run with winds, and run one kernel; the outputs must agree
bit for bit (both are FP64 without FMA contraction), except where THOR's
kernel calls pow/log/sin/cos, which come from the device libm on one side
and glibc on the other (1e-14 relative).

Where oracle/_ref/ has not been built (no reference tree when the tree was
built), these tests are skipped.
"""
from __future__ import annotations

import numpy as np
import pytest

from common import earth_params, grid

pytestmark = pytest.mark.gpu

# kernel name, a0, a1, outputs, relative tolerance
# Kernels that call pow (deep-atmosphere r^2, r^3, |M|^2) get the device
# libm's pow on THOR's side and glibc's on the oracle's: those differ by an
# ulp, hence 1e-14; the others must agree exactly. Compute_Slow_Modes forms
# dr2dz = (r_top^3 - r_bot^3) / 3 (thor_slowmodes.h:237), a cancellation that
# magnifies a 1-ulp difference in pow(r, 3) by about r / (3 dz) ~ 1e3 at this
# grid's 2.25 km layers, hence 1e-12 there; a dropped term or a wrong sign
# would still show up as an O(1) difference. vertical_diff calls pow(r, 2)
# (1e-14; measured 8e-16).
CASES = [
    ("advec_cori1", 0, 0, ["Adv", "v"], 0.0),
    ("advec_cori2", 0, 0, ["Adv"], 1e-14),
    ("diffusion", 0, 0, ["diff"], 0.0),
    ("diffusion", 1, 1, ["diffmh", "diffw", "diffrh", "diffpr", "diff_sponge"], 0.0),
    ("vertical_diff", 0, 0, ["diffmv", "diffwv", "diffprv"], 1e-14),
    ("correct_horizontal", 0, 0, ["diffmh"], 0.0),
    ("divm", 0, 0, ["divg_Mh"], 0.0),
    ("divm", 1, 0, ["DivM"], 0.0),
    ("slow_modes", 0, 0, ["SlowMh", "SlowWh", "SlowRho", "Slowpressure"], 1e-12),
    ("held_suarez", 0, 0, ["Mh", "temperature"], 1e-14),
]

NAMES = ["Mh", "Wh", "W", "Rho", "pressure", "temperature", "Rd", "Cp", "profx_Qheat", "profx_dMh", "profx_dWh",
         "Mhk", "Whk", "Wk", "Rhok", "pressurek", "h", "pt", "gtil", "hh", "pth", "gtilh", "Adv", "v", "SlowMh",
         "SlowWh", "SlowRho", "Slowpressure", "diff", "diffmh", "diffw", "diffrh", "diffpr", "DivM", "divg_Mh",
         "Kdh4", "Kdhz", "Kdv6", "diffmv", "diffwv", "diffrv", "diffprv", "Kdh2", "diff_sponge"]


@pytest.fixture(scope="module")
def setup(built):
    from oracle import oracle as om
    if not om.ref_available():
        pytest.skip("THOR's kernels are not built (oracle/_ref/libthor_ref.so missing)")
    capi = built
    g = grid(capi, 4, 16)
    # vertical hyperdiffusion on, so the slow modes also see its terms; the
    # second-order diffusive sponge on from half the model top, so the last
    # hyperdiffusion pass also forms and adds THOR's diff_sponge
    p = earth_params(capi, 4, dt=600.0, VertHyDiff=1, VertHyDiffOrder=6, Diffc_v=2e-3, DiffSponge=1,
                     order_diff_sponge=2, Dv_sponge=0.01, ns_diff_sponge=0.5)
    st = capi.ic_rest(g, p, seed=1000, temperature_noise=1e-3, wind_noise=1.0)
    o = om.Oracle(g, p)
    o.set_state(st)
    o.steps(2)  # every intermediate array now holds realistic values
    ref = om.RefKernels(g, p)
    return o, ref, {n: o.get(n) for n in NAMES}


@pytest.mark.parametrize("name,a0,a1,outputs,tol", CASES,
                         ids=[f"{c[0]}-{c[1]}{c[2]}" for c in CASES])
def test_oracle_matches_thor_kernel(setup, name, a0, a1, outputs, tol):
    o, ref, base = setup
    rng = np.random.default_rng(7)
    pre = dict(base)
    for n in outputs:  # perturbed outputs: overwrite vs accumulate must agree too
        a = base[n]
        pre[n] = a * (1.0 + 0.01 * rng.standard_normal(a.size))
    for n, a in pre.items():
        o.set(n, a)
        ref.set(n, a)
    print(f"case {name} a0={a0} a1={a1}", flush=True)
    o.kernel(name, a0, a1)
    ref.kernel(name, a0, a1)
    for n in outputs:
        want, got = ref.get(n), o.get(n)
        scale = max(np.abs(want).max(), 1e-300)
        err = np.abs(got - want).max() / scale
        changed = np.abs(want - pre[n]).max() / scale
        print(f"  {n}: max rel diff {err:.3e} (output changed by {changed:.3e})", flush=True)
        assert changed > 0, f"{name}: {n} not written by THOR's kernel"
        assert err <= tol, f"{name}: {n} differs from THOR's kernel by {err:.3e}"


@pytest.fixture(scope="module")
def setup_energy(built):
    """The same state advanced by the total-energy form (thermo_equation =
    energy), which switches the fifth hyperdiffused field to (Cp - Rd) T
    (thor_diff.h:196-202) and scales its slow-mode pressure tendency by
    Rd / Cv (thor_slowmodes.h:375-381)."""
    from oracle import oracle as om
    if not om.ref_available():
        pytest.skip("THOR's kernels are not built (oracle/_ref/libthor_ref.so missing)")
    capi = built
    g = grid(capi, 4, 16)
    p = earth_params(capi, 4, dt=600.0, VertHyDiff=1, VertHyDiffOrder=6, Diffc_v=2e-3, DiffSponge=1,
                     order_diff_sponge=2, Dv_sponge=0.01, ns_diff_sponge=0.5, thermo_equation=capi.THERMO_ENERGY)
    st = capi.ic_rest(g, p, seed=1001, temperature_noise=1e-3, wind_noise=1.0)
    o = om.Oracle(g, p)
    o.set_state(st)
    o.steps(2)
    ref = om.RefKernels(g, p)
    return o, ref, {n: o.get(n) for n in NAMES}


ENERGY_CASES = [c for c in CASES if c[0] in ("diffusion", "slow_modes")]


@pytest.mark.parametrize("name,a0,a1,outputs,tol", ENERGY_CASES,
                         ids=[f"{c[0]}-{c[1]}{c[2]}" for c in ENERGY_CASES])
def test_oracle_matches_thor_kernel_energy(setup_energy, name, a0, a1, outputs, tol):
    test_oracle_matches_thor_kernel(setup_energy, name, a0, a1, outputs, tol)


@pytest.mark.parametrize("deep", [1, 0])
def test_tracer_kernels_match_thor(built, deep):
    """The chemistry module's tracer kernels (thor_chemistry.h), two tracers:
    the first diffusion pass (THOR writes it into diff_d), the last pass
    (difftr; the pole kernel writes nothing there, which the oracle keeps) and
    Tracer_Eq (tracers), from the same random tracers on an oracle state with
    winds, bit for bit (deep: r^2 by pow on both sides, glibc's in the oracle
    and the device's in THOR's kernel; their 1-ulp differences pass through
    the cancelling flux sums, a few 1e-14 of the largest tracer tendency
    depending on the grid's last bits, so 1e-13)."""
    from oracle import oracle as om
    if not om.ref_available():
        pytest.skip("THOR's kernels are not built (oracle/_ref/libthor_ref.so missing)")
    capi = built
    ntr = 2
    g = grid(capi, 4, 16)
    p = earth_params(capi, 4, dt=600.0, DeepModel=deep)
    st = capi.ic_rest(g, p, seed=1008, temperature_noise=1e-3, wind_noise=1.0)
    o = om.Oracle(g, p)
    o.set_state(st)
    o.steps(1)
    P, nv = g.P, g.nv
    rng = np.random.default_rng(9)
    rho = o.get("Rhok").reshape(P, nv)
    trk = (rho[:, :, None] * (1.0 + 0.2 * rng.uniform(-1, 1, (P, nv, ntr)))).ravel()
    o.set_tracers(ntr)
    ref = om.RefKernels(g, p)
    shared = ["Rhok", "Rhos", "Mhs", "Mhk", "Whs", "Whk", "Kdh4"]
    vals = {n: o.get(n) for n in shared}
    vals["Rhos"] = rng.uniform(-1e-4, 1e-4, P * nv)
    vals["Mhs"] = rng.uniform(-0.1, 0.1, 3 * P * nv)
    vals["Whs"] = rng.uniform(-1e-3, 1e-3, P * (nv + 1))
    for n, a in vals.items():
        o.set(n, a)
        ref.set(n, a)
    o.set("tracerk", trk)
    ref.set("tracerk", trk)
    tol = 1e-13 if deep else 0.0

    def cmp(label, got, want):
        err = np.abs(got - want).max() / max(np.abs(want).max(), 1e-300)
        print(f"  {label}: max rel diff {err:.3e}", flush=True)
        assert np.abs(want).max() > 0 and err <= tol, label

    print(f"case tracers, DeepModel {deep}", flush=True)
    o.kernel("tracer_diffusion", 0, 0)
    ref.kernel("tracer_diffusion", 0, 0)
    first = ref.get("diff").reshape(P, nv, 6)[:, :, :ntr].ravel()
    cmp("first pass", o.get("trdiff"), first)
    o.kernel("tracer_diffusion", 1, 0)
    ref.kernel("tracer_diffusion", 1, 0)
    cmp("difftr", o.get("difftr"), ref.get("difftr"))
    o.kernel("tracer_eq", 100000, 0)
    ref.kernel("tracer_eq", 100000, 0)
    cmp("tracers", o.get("tracers"), ref.get("tracers"))
