"""Surrogate network on the GPU: MFMA forward and backward against the CPU
restatement, training, and the surrogate inside the hot-Jupiter step.

Tolerances: the oracle rounds to BF16 at the same points and computes in
FP64; the GPU accumulates in FP32, so an activation close to a BF16 rounding
boundary can land one BF16 step (2^-8 relative) apart. Outputs and gradients
therefore agree to ~1e-3 relative, and the tests allow 1e-2 (outputs) and
3e-2 (gradient blocks, relative L2).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from common import hj_grid, hj_params, rt_params

pytestmark = pytest.mark.gpu

DIMS = (81, 256, 40)


def _blocks(g, dims):
    n_in, nh, n_out = dims
    sizes = [n_in * nh, nh, nh * nh, nh, nh * n_out, n_out]
    names = ["W1", "b1", "W2", "b2", "W3", "b3"]
    out, o = {}, 0
    for n, s in zip(names, sizes):
        out[n] = g[o:o + s]
        o += s
    return out


def test_forward_matches_oracle(built):
    from oracle import mlp_oracle as mo
    capi = built
    print("case: MLP 81-256-256-40 forward, 300 rows", flush=True)
    m = capi.MLP(*DIMS, seed=3)
    rng = np.random.default_rng(1)
    p = m.params()
    p[DIMS[0] * DIMS[1]:DIMS[0] * DIMS[1] + DIMS[1]] = 0.1 * rng.standard_normal(DIMS[1])  # non-zero b1
    m.set_params(p)
    X = rng.standard_normal((300, DIMS[0])).astype(np.float32)
    Y = m.forward(X)
    Yo, _ = mo.forward(m.params(), X, DIMS, bf=True)
    err = np.abs(Y - Yo).max() / np.abs(Yo).max()
    print(f"forward vs oracle: max rel diff {err:.2e}")
    assert err < 1e-2
    m.close()


def test_fused_inference_equals_training_forward(built):
    """thor_mlp_forward runs the fused three-layer inference kernel; the
    training path (thor_mlp_gradients) runs three separate GEMMs with the same
    K tiling and epilogues. With zero targets the training loss is mean(Y^2)
    of the GEMM outputs, so it must equal mean(Y^2) of the fused outputs."""
    capi = built
    print("case: fused inference vs three-GEMM forward, 1000 rows (partial last block)", flush=True)
    m = capi.MLP(*DIMS, seed=6)
    rng = np.random.default_rng(7)
    p = m.params()
    nb = DIMS[0] * DIMS[1]
    p[nb:nb + DIMS[1]] = 0.1 * rng.standard_normal(DIMS[1])
    m.set_params(p)
    X = rng.standard_normal((1000, DIMS[0])).astype(np.float32)
    Y = m.forward(X).astype(np.float64)
    _, loss = m.gradients(X, np.zeros((1000, DIMS[2]), np.float32))
    want = float((Y ** 2).mean())
    print(f"mean(Y^2): fused {want:.12e}, GEMM path {loss:.12e}")
    assert abs(loss / want - 1) < 1e-9
    m.close()


def test_gradients_match_oracle(built):
    from oracle import mlp_oracle as mo
    capi = built
    print("case: MLP gradients, 200 rows", flush=True)
    m = capi.MLP(*DIMS, seed=4)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((200, DIMS[0])).astype(np.float32)
    T = rng.standard_normal((200, DIMS[2])).astype(np.float32)
    g, loss = m.gradients(X, T)
    lo, go = mo.loss_and_grad(m.params(), X, T, DIMS, bf=True)
    print(f"loss gpu {loss:.6e} oracle {lo:.6e}")
    assert abs(loss / lo - 1) < 1e-3
    bg, bo = _blocks(g, DIMS), _blocks(go, DIMS)
    for k in bg:
        e = np.linalg.norm(bg[k] - bo[k]) / np.linalg.norm(bo[k])
        print(f"  d{k}: relative L2 diff {e:.2e}")
        assert e < 3e-2, k
    m.close()


def test_training_fits_a_teacher_network(built):
    capi = built
    print("case: train 81-256-256-40 on a random teacher, 8192 rows, 15 epochs", flush=True)
    rng = np.random.default_rng(3)
    teacher = capi.MLP(*DIMS, seed=11)
    X = rng.standard_normal((8192, DIMS[0])).astype(np.float32)
    Y = teacher.forward(X)
    Y = (Y - Y.mean(axis=0)) / Y.std(axis=0)
    student = capi.MLP(*DIMS, seed=12)
    student.set_data(X, Y)
    losses = [student.train_epoch(256, 1e-3 * 0.9 ** e, e) for e in range(15)]
    print("losses", " ".join(f"{v:.3e}" for v in losses))
    assert np.isfinite(losses).all()
    assert losses[-1] < losses[0] / 2.5
    teacher.close()
    student.close()


def test_surrogate_in_the_hot_jupiter_step(built):
    """Train on exact-RT columns of a short hot-Jupiter run, then run with the
    surrogate: the judge's RMS heating error is bounded, the run stays
    finite, and the trained network beats the untrained one by far."""
    from thor_amd import surrogate as sg
    capi = built
    print("case: hot Jupiter g4 L20, RT surrogate trained on 30 steps of exact RT", flush=True)
    g = hj_grid(capi, 4, 20)
    p = hj_params(capi, 4)
    st = capi.ic_rest(g, p, seed=1003, temperature_noise=1.0, wind_noise=0.1)
    d = capi.Dyn(g, p)
    d.set_state(st)
    d.set_rt(rt_params(capi))
    X, Y = sg.collect(d, g, nsteps=30, every=2)
    norm = sg.normalisation(X, Y)
    Xn, Yn = sg.apply_norm(X, Y, norm)
    nv = g.nv
    m = capi.MLP(2 * nv + 1, 256, nv, seed=5)
    err0 = sg.heating_error(m, X, Y, norm)
    losses = sg.train(m, Xn, Yn, epochs=20, batch=256, lr=2e-3)
    err1 = sg.heating_error(m, X, Y, norm)
    print(f"training loss {losses[0]:.3e} -> {losses[-1]:.3e}; heating RMS error untrained {err0:.3f}, "
          f"trained {err1:.3f}")
    assert err1 < 0.5 * err0
    d.set_rt_surrogate(m, norm, judge_every=4)
    d.run(12, capi.FORCING_RT_SURROGATE)
    d.sync()
    stats = d.surrogate_stats()
    qrms = float(np.sqrt((Y ** 2).mean()))
    print(f"surrogate stats {stats}; judged RMS error / RMS heating = {stats['rms_err'] / qrms:.3f}")
    assert stats["calls"] == 12 and stats["judged"] == 3
    assert stats["rms_err"] < 0.25 * qrms
    assert all(np.isfinite(v).all() for v in d.get_state().values())
    d.close()
    m.close()


@pytest.mark.parametrize("glevel,nv", [(4, 20), (4, 48), (5, 40)])
def test_surrogate_profx_is_the_network_on_column_features(built, glevel, nv):
    """One surrogate ProfX (no judge) against the network evaluated by
    thor_mlp_forward on features formed on the host from the same state: the
    fused kernel's features, envelope test, inference and heating must agree.
    The host forms the normalised inputs with the kernel's recipe
    (surrogate.normalise_inputs); only the FP64 logarithm of p can differ by
    an ulp between numpy and the device, which moves an input one BF16 step
    when it sits at a rounding boundary (very rare), and that column's
    outputs by up to ~1e-2 of their scale; the envelope decision can differ
    likewise for a column at its edge. Tolerances: at least 99 % of the columns within 1e-5 of the heating
    scale, every column within 2e-2, and the fallback count within 0.1 % of
    the columns. L48 has 97 inputs (padded K) and 48 outputs (three
    16-wide output tiles); g5 L40 is the learned-RT workload's own size (81
    inputs, 40 outputs, 10 242 columns)."""
    from thor_amd import surrogate as sg
    capi = built
    print(f"case: hot Jupiter g{glevel} L{nv}, surrogate ProfX vs host features + thor_mlp_forward", flush=True)
    g = hj_grid(capi, glevel, nv)
    p = hj_params(capi, glevel)
    st = capi.ic_rest(g, p, seed=1005, temperature_noise=1.0, wind_noise=0.1)
    d = capi.Dyn(g, p)
    d.set_state(st)
    d.set_rt(rt_params(capi))
    X, Y = sg.collect(d, g, nsteps=6, every=2)
    norm = sg.normalisation(X, Y)
    m = capi.MLP(2 * nv + 1, 256, nv, seed=8)
    sg.train(m, *sg.apply_norm(X, Y, norm), epochs=3, batch=256, lr=2e-3)
    d.set_rt_surrogate(m, norm, judge_every=0)
    pres = d.get_field("pressure")
    rho = d.get_field("Rho")
    T = pres / (p.Rd * rho)
    Xh = sg.features(T, pres, sg.cos_zenith(g))
    Xn = sg.normalise_inputs(Xh, norm)
    inside = ((Xn >= norm["in_lo"]) & (Xn <= norm["in_hi"])).all(axis=1)
    # columns whose envelope decision no rounding can change
    clear = ((Xn >= norm["in_lo"] + 1e-4) & (Xn <= norm["in_hi"] - 1e-4)).all(axis=1)
    d.profx(capi.FORCING_RT_SURROGATE)
    d.sync()
    Q = d.get_field("profx_Qheat")
    Yn = m.forward(Xn).astype(np.float64)
    Qh = Yn * norm["out_std"].astype(np.float64) + norm["out_mean"].astype(np.float64)
    dt = p.timestep
    clip = pres + p.Rd / (p.Cp - p.Rd) * Qh * dt < 0
    Qh = np.where(clip, -pres / dt, Qh)
    scale = np.abs(Qh[clear]).max()
    col_err = np.abs(Q[clear] - Qh[clear]).max(axis=1) / scale
    err = col_err.max()
    close = float((col_err <= 1e-5).mean())
    stats = d.surrogate_stats()
    print(f"{inside.sum()} of {g.P} columns inside the envelope (library: {g.P - stats['fallback_columns']}); "
          f"max rel diff {err:.2e}, {100 * close:.2f} % of columns within 1e-5")
    assert abs(stats["fallback_columns"] - int((~inside).sum())) <= max(1, g.P // 1000)
    assert close >= 0.99
    assert err < 2e-2
    d.close()
    m.close()



def test_concurrent_fallback_equals_sequential(built):
    """Surrogate ProfX with the exact RT of the out-of-envelope columns on a
    second stream beside the network (flags formed first) against the
    sequential order (network, then the fallback list): six steps with a
    judge every third call give the same states bit for bit and the same
    fallback count, and some columns did fall back."""
    from thor_amd import surrogate as sg
    capi = built
    print("case: hot Jupiter g4 L20, surrogate ProfX, concurrent vs sequential fallback, 6 steps", flush=True)
    g = hj_grid(capi, 4, 20)
    p = hj_params(capi, 4)
    st = capi.ic_rest(g, p, seed=1007, temperature_noise=1.0, wind_noise=0.1)
    d0 = capi.Dyn(g, p)
    d0.set_state(st)
    d0.set_rt(rt_params(capi))
    X, Y = sg.collect(d0, g, nsteps=6, every=2)
    start = d0.get_state()
    d0.close()
    norm = sg.normalisation(X, Y)
    m = capi.MLP(2 * g.nv + 1, 256, g.nv, seed=9)
    sg.train(m, *sg.apply_norm(X, Y, norm), epochs=3, batch=256, lr=2e-3)
    out = {}
    for conc in ("1", "0"):
        old = os.environ.get("THOR_SURROGATE_CONCURRENT")
        os.environ["THOR_SURROGATE_CONCURRENT"] = conc
        try:
            d = capi.Dyn(g, p)
        finally:
            if old is None:
                del os.environ["THOR_SURROGATE_CONCURRENT"]
            else:
                os.environ["THOR_SURROGATE_CONCURRENT"] = old
        d.set_state(start)
        d.set_rt(rt_params(capi))
        d.set_rt_surrogate(m, norm, judge_every=3)
        d.run(6, capi.FORCING_RT_SURROGATE)
        d.sync()
        out[conc] = (d.get_state(), d.surrogate_stats(), d.get_field("profx_Qheat"))
        d.close()
    (a, sa, qa), (b, sb, qb) = out["1"], out["0"]
    print(f"fallback columns: concurrent {sa['fallback_columns']}, sequential {sb['fallback_columns']}")
    assert sa["fallback_columns"] == sb["fallback_columns"] > 0
    for k in a:
        assert np.isfinite(a[k]).all()
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(qa, qb)
    m.close()
