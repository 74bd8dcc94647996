"""Pins the CPU oracle (serial restatement of THOR's kernels) against closed
forms, invariants and textbook special cases (CPU only).

THOR ships no numeric golden vectors for the dynamical core (its BENCH_POINT
and test_data references are written at run time); the checks below are the
ones a plausible restatement mistake (dropped term, wrong sign or index,
transposed operand) would fail.
"""
import numpy as np
import pytest

from tests.common import earth_params, grid

A = 6.371e6


@pytest.fixture(scope="module")
def setup(built):
    capi = built
    g = grid(capi, 4, 16)
    p = earth_params(capi, 4)
    return capi, g, p


def test_tree_sum_matches_model_and_exact_integers(built):
    from oracle import oracle
    rng = np.random.default_rng(7)
    ints = rng.integers(-1000, 1000, size=5001).astype(np.float64)
    assert oracle.tree_sum(ints) == float(ints.sum())
    x = rng.standard_normal(7777)
    # model of reduction_add.h:59-111: 2048-element blocks, stride halving, host sum
    out = 0.0
    for b in range(0, x.size, 2048):
        buf = np.zeros(2048)
        chunk = x[b:b + 2048]
        buf[:chunk.size] = chunk
        s = 1024
        while s > 0:
            buf[:s] = buf[:s] + buf[s:2 * s]
            s //= 2
        out += buf[0]
    assert oracle.tree_sum(x) == out


@pytest.mark.parametrize("variant", ["", "pm"])
def test_elementary_functions(built, variant):
    from oracle import oracle
    rng = np.random.default_rng(3)
    for _ in range(2000):
        x = float(np.exp(rng.uniform(-8, 8)))
        y = float(rng.uniform(0.1, 3.0))
        assert abs(oracle.math("pow", x, y, variant) / x ** y - 1) < 2e-14
        assert abs(oracle.math("log", x, 0, variant) - np.log(x)) <= 4e-16 * max(1, abs(np.log(x)))
        t = float(rng.uniform(-3.2, 3.2))
        assert abs(oracle.math("sin", t, 0, variant) - np.sin(t)) < 2e-15
        assert abs(oracle.math("cos", t, 0, variant) - np.cos(t)) < 2e-15
    for x in (1.0, 3.5, 6.371e6 + 36000.0):
        assert oracle.math("pow", x, 2.0, variant) == x * x
    # exp over its whole range, including subnormal results, overflow and underflow
    for x in np.concatenate([rng.uniform(-745, 709.7, 3000), [-1e6, -746.0, -745.0, -708.4, 709.78, 710.0]]):
        want = np.exp(x)
        got = oracle.math("exp", float(x), 0, variant)
        if want == 0.0 or np.isinf(want):
            assert got == want or (want == 0.0 and abs(got) < 1e-320), x
        elif want < 2.2250738585072014e-308:
            assert abs(got - want) <= 2 * 4.9e-324, x  # subnormal: within two of the smallest step
        else:
            assert abs(got / want - 1) < 4e-16, x
    # sin / cos beyond [-pi, pi] and log at its edges
    for t in rng.uniform(-10.0, 10.0, 500):
        assert abs(oracle.math("sin", float(t), 0, variant) - np.sin(t)) < 4e-15
        assert abs(oracle.math("cos", float(t), 0, variant) - np.cos(t)) < 4e-15
    assert oracle.math("log", 0.0, 0, variant) == -np.inf
    assert np.isnan(oracle.math("log", -1.0, 0, variant))
    assert abs(oracle.math("log", 1e-310, 0, variant) - np.log(1e-310)) < 1e-13


@pytest.mark.parametrize("Rd,Cp", [(287.0, 1005.0), (3779.0, 13226.0)])
def test_portable_pow_over_the_step_ranges(built, Rd, Cp):
    """pm::pow, which the production mixed form calls on the GPU for the
    pressure update, against glibc over the arguments the step uses:
    p / P_Ref from 1e-8 to 10 with the exponents Cv/Cp and Cp/Cv, and
    1 + x for the binomial-series cut-over |x| ~ 2^-7 and beyond. Bound:
    2.5 (1 + |y log x|) ulp as the header states, i.e. below 2e-14 relative."""
    from oracle import oracle
    Cv = Cp - Rd
    rng = np.random.default_rng(17)
    xs = np.concatenate([10.0 ** rng.uniform(-8, 1, 4000), 1.0 + rng.uniform(-0.05, 0.05, 1000),
                         [1e-8, 1.0, 10.0, 1.0 + 2.0 ** -7, 1.0 - 2.0 ** -7]])
    for y in (Cv / Cp, Cp / Cv):
        worst = 0.0
        for x in xs:
            got = oracle.math("pow", float(x), y, "pm")
            want = float(x) ** y
            ulp = np.spacing(want)
            bound = 2.5 * (1.0 + abs(y * np.log(x))) * ulp
            assert abs(got - want) <= bound, (x, y, got, want)
            worst = max(worst, abs(got / want - 1))
        assert worst < 2e-14, worst


def test_rest_state_is_discrete_hydrostatic(setup):
    """The IC solves (p_k - p_{k-1})/dz + g rho_interp = 0 level by level
    (esp_initial.cu:520-587); isothermal: close to P0 exp(-z/H)."""
    capi, g, p = setup
    st = capi.ic_rest(g, p)
    pr, rho = st["pressure"][0], st["Rho"][0]
    z = g.Altitude
    for k in range(1, g.nv):
        dz = z[k] - z[k - 1]
        l_int = (z[k] - g.Altitudeh[k]) / dz
        r_int = (g.Altitudeh[k] - z[k - 1]) / dz
        res = (pr[k] - pr[k - 1]) / dz + p.Gravit * (rho[k] * r_int + rho[k - 1] * l_int)
        assert abs(res) < 1e-9 * p.Gravit * rho[k]
    # isothermal closed form of the discrete recursion (uniform layers, l = r = 1/2):
    # p_k = p_{k-1} (1 - b) / (1 + b), b = g dz / (2 R T), with p_{-1} = P_Ref at z = 0
    RT = p.Rd * p.Tmean
    expect = np.empty(g.nv)
    prev, zprev = p.P_Ref, 0.0
    for k in range(g.nv):
        b = 0.5 * p.Gravit * (z[k] - zprev) / RT
        expect[k] = prev * (1 - b) / (1 + b)
        prev, zprev = expect[k], z[k]
    assert np.allclose(pr, expect, rtol=1e-12)
    H = RT / p.Gravit
    assert np.allclose(pr, p.P_Ref * np.exp(-z / H), rtol=3e-2)
    assert np.allclose(pr / (p.Rd * rho), p.Tmean, rtol=1e-13)
    assert (st["pressure"] == st["pressure"][0]).all()  # horizontally uniform


def test_held_suarez_closed_form(setup):
    """ProfX with Held & Suarez (1994) forcing: Teq, kT, kv of the paper; the
    surface pressure is THOR's linear extrapolation (profx_held_suarez.h:89-91),
    exact for a column whose pressure is linear in height and hydrostatic."""
    from oracle.oracle import Oracle
    capi, g, p = setup
    P, nv = g.P, g.nv
    z = g.Altitude
    a0, b = 1.0e5, -2.0                    # p = a0 + b z  (Pa), rho0 = -b / g
    pr = np.tile(a0 + b * z, (P, 1))
    rho = np.full((P, nv), 1.1)
    rho[:, 0] = -b / p.Gravit
    rng = np.random.default_rng(11)
    Mh = rng.uniform(-1, 1, (P, nv, 3))
    o = Oracle(g, p)
    o.set_state({"Mh": Mh, "Wh": np.zeros((P, nv + 1)), "W": np.zeros((P, nv)), "Rho": rho, "pressure": pr})
    o.profx()
    s = o.get_state()
    lat = g.lonlat[:, 1][:, None]
    ps = a0
    sigma = pr / ps
    kappa = p.Rd / p.Cp
    teq = np.maximum(200.0, (315.0 - 60.0 * np.sin(lat) ** 2 - 10.0 * np.log(pr / 1e5) * np.cos(lat) ** 2)
                     * (pr / 1e5) ** kappa)
    day = 86400.0
    kt = 1 / (40 * day) + (1 / (4 * day) - 1 / (40 * day)) * np.maximum(0, (sigma - 0.7) / 0.3) * np.cos(lat) ** 4
    kv = (1 / day) * np.maximum(0, (sigma - 0.7) / 0.3)
    T0 = pr / (p.Rd * rho)
    T1 = T0 - kt * p.timestep * (T0 - teq)
    assert np.allclose(s["pressure"] / (p.Rd * rho), T1, rtol=1e-12)
    assert np.allclose(s["Mh"], Mh / (1 + kv * p.timestep)[:, :, None], rtol=1e-13)
    # only the layers with sigma > 0.7 are damped; the others keep Mh bit for bit
    free = sigma < 0.7 - 1e-12
    damped = sigma > 0.7 + 1e-12
    assert free.any() and damped.any()
    assert (kv[free] == 0).all()
    assert np.array_equal(s["Mh"][free], Mh[free])
    assert (np.abs(s["Mh"][damped]) < np.abs(Mh[damped])).all()


def test_mass_conserved_and_rest_preserved(setup):
    """Flux-form continuity: sum(rho Vol) is conserved by the step (CalcMass);
    a horizontally uniform hydrostatic rest state stays at rest."""
    from oracle.oracle import Oracle
    capi, g, p = setup
    st = capi.ic_rest(g, p)
    o = Oracle(g, p)
    o.set_state(st)
    m0 = o.globdiag()[0]
    for _ in range(2):
        o.thor()
    s = o.get_state()
    m1 = o.globdiag()[0]
    assert abs(m1 / m0 - 1) < 1e-14
    cs = np.sqrt(p.Cp / (p.Cp - p.Rd) * p.Rd * p.Tmean)
    assert np.abs(s["Mh"]).max() < 1e-10 * st["Rho"].max() * cs
    assert np.abs(s["Wh"]).max() < 1e-8 * st["Rho"].max() * cs
    assert np.allclose(s["pressure"], st["pressure"], rtol=1e-10)


def test_noise_breaks_symmetry_but_conserves_mass(setup):
    from oracle.oracle import Oracle
    capi, g, p = setup
    st = capi.ic_rest(g, p, seed=1000, temperature_noise=1e-3)
    o = Oracle(g, p)
    o.set_state(st)
    m0 = o.globdiag()[0]
    o.steps(2)
    s = o.get_state()
    assert abs(o.globdiag()[0] / m0 - 1) < 1e-14
    assert np.abs(s["Mh"]).max() > 0
    assert np.isfinite(s["pressure"]).all()
