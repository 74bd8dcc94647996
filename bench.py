"""Benchmark: simulated days per wall-day of the HE-VI dynamical-core step.

Workload (the configuration the headline metric is quoted on): Held-Suarez
Earth, icosahedral glevel 6 (40 962 columns), 48 levels, 36 km top, dt = 300 s,
8 acoustic substeps, 4th-order hyperdiffusion and 3D divergence damping on.
One "step" = ProfX (Held-Suarez forcing) + Thor (one large RK3 step), the same
sequence as THOR's main loop (esp.cu:1393-1398). Synthetic input: the rest
state plus seeded +-1e-3 K temperature noise; all state resident in HBM.

Both forms are timed: the plain FP64 path (THOR's arithmetic) and the mixed-
precision path (FP32 hyperdiffusion / divergence-damping operators and
momentum deviation, FP64 elsewhere). `value` is the mixed-precision form's
throughput; `fp64` and `speedup_vs_fp64` give the FP64 form beside it.
Held-Suarez has no radiative transfer, so the RT surrogate is not on this
workload's path; the learned_rt object below times it on its own workload.

Prints ONE JSON line. Extra objects:
  roofline       - dominant kernel of the headline form: algorithmic bytes per
                   launch (per-cell arrays of the reference kernels it fuses,
                   each at its storage width, x P x nv) / average launch time
                   from HIP events on the library's stream, against 8 TB/s;
                   "traffic" = HBM bytes per launch from the committed
                   rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/hbm_traffic_*.json);
                   "step" = the whole large step's algorithmic bytes / wall time;
  error_vs_fp64  - FP64, mixed and a one-ulp-perturbed FP64 twin run
                   --error-days simulated days from the same state; each day
                   the RMS / max differences of T, u, v, w and surface
                   pressure (point-wise, and of the zonal means of T and u)
                   of the mixed form from FP64, beside the perturbed twin's
                   (the chaotic spread precision cannot be told apart from);
  cpu_baseline   - the FP64 oracle (THOR's kernels as a host loop) timed on
                   the host's threads (and on one) for steps of the same
                   workload;
  learned_rt     - the learned-RT workload (hot Jupiter g5 L40, double-gray
                   RT): 50-day exact-RT spin-up, GPU training of the MLP,
                   ProfX and step timing of FP64 + exact RT, mixed + exact RT
                   and mixed + MLP (judge every 10 calls), judge error.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Sim-days/wall-day at g6 L48, mixed+MLP vs FP64 at stated err; HBM GB/s % peak"
HBM_PEAK_GBS = 8000.0
# launches whose bytes stay out of the step's algorithmic total (no counterpart in the reference's kernels)
NOT_REFERENCE = ("thomas_factor", "divm1")
CACHE = os.path.join(ROOT, ".thor_cache")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_probe(mib: int) -> None:
    """Plain copy kernel of known size: reads and writes mib MiB per launch."""
    import torch
    a = torch.ones((mib << 20) // 8, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy probe: 3 x {mib} MiB device copy")
    del a, b


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter runs
    (profiles/hbm_traffic_<workload>.json: FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by
    what the same counter reads for a copy of known size), or None when the file
    does not cover this kernel and workload."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if t.get("workload") != workload or k is None:
        return None
    return k["hbm_bytes_per_launch"]


# kernel symbol (demangled, rocprofv3) -> launch name of the library's per-kernel timing
TRAFFIC_NAMES = [
    (r"k_profx_relax<1>", "profx_held_suarez"),
    (r"k_column_prep", "column_prep"),
    (r"k_adv_diff1", "adv_diff1"),
    (r"k_diff2_divm0", "diff2_divm0"),
    (r"k_diff2_lc", "diff2"),
    (r"k_divm0_lc<true>", "divm0_stage"),
    (r"k_divm0", "divm0"),
    (r"k_divm1", "divm1"),
    (r"k_slow_levels", None),  # the slow modes' level table, once per handle
    (r"k_slow", "slow_modes"),
    (r"k_thomas_factor", "thomas_factor"),
    (r"k_momentum", "momentum"),
    (r"k_prepare", "prepare"),
    (r"k_thomas_dp", "thomas_dp"),
    (r"k_update_rk2", "update_rk2"),
]


def traffic_summary(fetch_dir: str, write_dir: str, workload: str, mib: int, commands: list) -> dict:
    """profiles/hbm_traffic_<workload>.json from two rocprofv3 counter runs of
    the same program (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch, csv
    output): each kernel's mean per dispatch, each counter divided by what it
    read for the program's plain copy of `mib` MiB (the largest copy
    dispatch, --copy-probe-mib)."""
    import csv
    import glob
    import re

    def load(d, counter):
        rows = []
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(path)):
                if r["Counter_Name"] == counter:
                    rows.append((r["Kernel_Name"], float(r["Counter_Value"])))
        return rows

    true_kib = mib * 1024.0
    out = {"workload": workload, "copy_probe_MiB": mib, "commands": commands, "kernels": {}}
    per = {}
    for counter, d, key in (("FETCH_SIZE", fetch_dir, "fetch"), ("WRITE_SIZE", write_dir, "write")):
        rows = load(d, counter)
        copies = [v for n, v in rows if "copyBuffer" in n or "elementwise" in n or "copy" in n.lower()]
        if not copies:
            raise SystemExit(f"no copy dispatch in {d}: run the program with --copy-probe-mib")
        factor = max(copies) / true_kib
        out[f"{key}_counter_per_true_byte"] = round(factor, 5)
        for n, v in rows:
            name = next((t for rx, t in TRAFFIC_NAMES if re.search(rx, n.replace(" ", ""))), "")
            if not name:
                continue
            e = per.setdefault(name, {"fetch": [], "write": []})
            e[key].append(v / factor)
    for name, e in per.items():
        if not e["fetch"] or not e["write"]:
            continue
        rd = sum(e["fetch"]) / len(e["fetch"]) * 1024.0
        wr = sum(e["write"]) / len(e["write"]) * 1024.0
        out["kernels"][name] = {"hbm_read_bytes": round(rd), "hbm_write_bytes": round(wr),
                                "hbm_bytes_per_launch": round(rd + wr), "dispatches": len(e["fetch"])}
    return out


def timed_run(capi, g, p, st, args):
    """K timed steps after W warm-up steps, then per-kernel HIP-event timing."""
    d = capi.Dyn(g, p, variant=args.variant)
    d.set_state(st)
    d.run(args.warmup)
    d.sync()
    t1 = time.perf_counter()
    d.run(args.steps)
    d.sync()
    wall = time.perf_counter() - t1
    nprof = max(3, min(20, args.steps // 5))
    d.kernel_timing(True)
    d.run(nprof)
    d.sync()
    stats = d.kernel_stats()
    d.kernel_timing(False)
    d.close()
    tot = sum(s["total_ms"] for s in stats)
    for s in sorted(stats, key=lambda s: -s["total_ms"]):
        avg = s["total_ms"] / s["launches"]
        gbs = s["bytes_per_launch"] / (avg * 1e-3) / 1e9
        log(f"  {s['name']:20s} {s['launches'] / nprof:6.1f}/step {avg * 1e3:9.1f} us "
            f"{100 * s['total_ms'] / tot:5.1f}%  {gbs:7.0f} GB/s alg")
    return {"ms_step": 1000.0 * wall / args.steps, "sdpd": args.dt * args.steps / wall, "stats": stats,
            "nprof": nprof}


def roofline(run, workload):
    stats, nprof = run["stats"], run["nprof"]
    dom = max(stats, key=lambda s: s["total_ms"])
    dom_avg_ms = dom["total_ms"] / dom["launches"]
    achieved = dom["bytes_per_launch"] / (dom_avg_ms * 1e-3) / 1e9
    # whole step: algorithmic bytes of every launch except those with no
    # reference counterpart (the per-stage Thomas factorisation, the separate
    # divergence-damping pass of the chunked slow modes); in FP64 this is the
    # reference decomposition's N * 8 B * (470 + 70 n_sub + 24)
    step_bytes = sum(s["bytes_per_launch"] * s["launches"] for s in stats if s["name"] not in NOT_REFERENCE) / nprof
    step_gbs = step_bytes / (run["ms_step"] * 1e-3) / 1e9
    traffic = measured_traffic(dom["name"], workload)
    # the same kernel's rate on the HBM bytes the counters saw (its fused form
    # moves fewer bytes than the reference decomposition books)
    t_gbs = traffic / (dom_avg_ms * 1e-3) / 1e9 if traffic else None
    return {
        "bound": "hbm", "kernel": dom["name"], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_achieved": round(t_gbs, 1) if t_gbs else None,
        "traffic_frac": round(t_gbs / HBM_PEAK_GBS, 4) if t_gbs else None,
        "alg_bytes_per_launch": round(dom["bytes_per_launch"]),
        "avg_launch_us": round(dom_avg_ms * 1e3, 2),
        "step": {"alg_bytes": round(step_bytes), "achieved": round(step_gbs, 1),
                 "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
    }


def _host_cpu() -> str:
    """Model name and logical CPU count of the host (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as fh:
            names = [ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")]
        return f"{names[0]} x {len(names)} logical CPUs (affinity: {len(os.sched_getaffinity(0))})"
    except (OSError, IndexError):
        return "unknown"


def diagnostics(g, p, st):
    """T, zonal / meridional wind, w and surface pressure of a state."""
    P = g.P
    rho, pr = st["Rho"], st["pressure"]
    T = pr / (rho * p.Rd)
    v = st["Mh"] / rho[:, :, None]
    lon, lat = np.asarray(g.lonlat).reshape(P, 2).T
    el = np.stack([-np.sin(lon), np.cos(lon), np.zeros(P)], axis=1)
    ep = np.stack([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)], axis=1)
    u = np.einsum("plc,pc->pl", v, el)
    vm = np.einsum("plc,pc->pl", v, ep)
    w = st["W"] / rho
    alt = np.asarray(g.Altitude)
    # surface pressure as the Held-Suarez forcing forms it (profx_held_suarez.h:89-91)
    psm1 = pr[:, 1] - rho[:, 0] * p.Gravit * (-alt[0] - alt[1])
    ps = 0.5 * (pr[:, 0] + psm1)
    return {"T": T, "u": u, "v": vm, "w": w, "ps": ps}


def _profx_us(stats, names):
    """Average device microseconds per ProfX call of the listed kernels."""
    calls = max((s["launches"] for s in stats
                 if s["name"] in ("profx_double_gray", "surrogate_features", "surrogate_profx")), default=1)
    return 1e3 * sum(s["total_ms"] for s in stats if s["name"] in names) / calls


def rt_surrogate_run(capi, args):
    """Hot-Jupiter g5 L40 with double-gray RT (mixed-precision dynamics):
    spin-up with the exact RT while sampling (T, ln p, cos z) -> Qheat columns,
    GPU training of the 81-256-256-40 surrogate, then ProfX and whole-step
    timing with the exact RT and with the surrogate (judge every 10 calls,
    exact RT for out-of-envelope columns)."""
    from thor_amd import surrogate as sg
    g, p = sg.hot_jupiter(5, 40, dt=300.0, ns=6, precision=1, cache_dir=CACHE)
    st = capi.ic_rest(g, p, seed=1003, temperature_noise=1.0, wind_noise=0.1)
    d = capi.Dyn(g, p)
    d.set_state(st)
    d.set_rt(sg.double_gray())
    d.set_sponge(sg.hot_jupiter_sponge())
    t0 = time.perf_counter()
    X, Y = sg.collect(d, g, nsteps=args.rt_spinup_steps, every=args.rt_sample_every)
    t_spin = time.perf_counter() - t0
    log(f"spin-up: {args.rt_spinup_steps} steps ({args.rt_spinup_steps * 300.0 / 86400:.1f} days) with exact RT, "
        f"{len(X)} columns sampled, {t_spin:.1f} s")
    norm = sg.normalisation(X, Y)
    Xn, Yn = sg.apply_norm(X, Y, norm)
    m = capi.MLP(2 * g.nv + 1, 256, g.nv, seed=5)
    err0 = sg.heating_error(m, X, Y, norm)
    t0 = time.perf_counter()
    losses = sg.train(m, Xn, Yn, epochs=args.rt_epochs, batch=512, lr=2e-3, log=log)
    t_train = time.perf_counter() - t0
    err1 = sg.heating_error(m, X, Y, norm)
    log(f"training: {args.rt_epochs} epochs in {t_train:.1f} s; heating RMS error untrained {err0:.3f}, "
        f"trained {err1:.3f}")
    n = args.rt_steps
    out = {}
    # the plain FP64 form from the same state: exact RT, FP64 dynamics
    _, p64 = sg.hot_jupiter(5, 40, dt=300.0, ns=6, precision=0, cache_dir=CACHE)
    d64 = capi.Dyn(g, p64)
    d64.set_state(d.get_state())
    d64.set_rt(sg.double_gray())
    d64.set_sponge(sg.hot_jupiter_sponge())
    d64.run(3, capi.FORCING_DOUBLE_GRAY)
    d64.sync()
    t1 = time.perf_counter()
    d64.run(n, capi.FORCING_DOUBLE_GRAY)
    d64.sync()
    wall = time.perf_counter() - t1
    out["fp64_exact_rt"] = {"ms_per_step": round(1e3 * wall / n, 4), "sdpd": round(300.0 * n / wall, 1)}
    d64.close()
    for form, forcing in (("exact_rt", capi.FORCING_DOUBLE_GRAY), ("surrogate", capi.FORCING_RT_SURROGATE)):
        if forcing == capi.FORCING_RT_SURROGATE:
            d.set_rt_surrogate(m, norm, judge_every=10)
            d.surrogate_stats()
        d.run(3, forcing)
        d.sync()
        if forcing == capi.FORCING_RT_SURROGATE:
            d.surrogate_stats()
        t1 = time.perf_counter()
        d.run(n, forcing)
        d.sync()
        wall = time.perf_counter() - t1
        sstats = d.surrogate_stats() if forcing == capi.FORCING_RT_SURROGATE else None
        d.kernel_timing(True)
        d.run(10, forcing)
        d.sync()
        ks = d.kernel_stats()
        d.kernel_timing(False)
        if forcing == capi.FORCING_RT_SURROGATE:
            d.surrogate_stats()
        rt_names = ("profx_double_gray", "profx_double_gray_fallback", "surrogate_features", "surrogate_mlp",
                    "surrogate_apply", "surrogate_profx")
        out[form] = {"ms_per_step": round(1e3 * wall / n, 4), "sdpd": round(300.0 * n / wall, 1),
                     "profx_us": round(_profx_us(ks, rt_names), 1),
                     "kernels_us": {s["name"]: round(1e3 * s["total_ms"] / s["launches"], 1) for s in ks
                                    if s["name"] in rt_names}}
        if sstats is not None:
            qrms = float(np.sqrt((Y ** 2).mean()))
            out[form]["judge"] = {"calls": sstats["calls"], "judged": sstats["judged"],
                                  "fallback_columns_per_call": round(sstats["fallback_columns"] / max(1, sstats["calls"]), 1),
                                  "rms_err_W_m3": sstats["rms_err"], "max_err_W_m3": sstats["max_err"],
                                  "rms_err_rel_to_rms_heating": round(sstats["rms_err"] / qrms, 4)}
    d.close()
    m.close()
    return {"metric": "double-gray RT vs MLP surrogate, hot Jupiter g5 L40 (10242 columns), mixed-precision "
                      "dynamics", "workload": "hot-Jupiter g5 L40, dt 300 s, 6 substeps, double-gray RT, Rayleigh "
                                              "sponge (top quarter, 1e-4 1/s, 20 bands)",
            "spinup_steps": args.rt_spinup_steps, "train_columns": int(len(X)), "epochs": args.rt_epochs,
            "train_seconds": round(t_train, 1), "train_loss_first_last": [losses[0], losses[-1]],
            "heating_rel_rms_err_untrained": round(err0, 4), "heating_rel_rms_err_trained": round(err1, 4),
            "timed_steps": n, **out,
            "profx_speedup": round(out["exact_rt"]["profx_us"] / out["surrogate"]["profx_us"], 2),
            "step_speedup": round(out["exact_rt"]["ms_per_step"] / out["surrogate"]["ms_per_step"], 3),
            "mixed_mlp_vs_fp64_plain": round(out["fp64_exact_rt"]["ms_per_step"] / out["surrogate"]["ms_per_step"], 3)}


def _spinup_and_train(capi, args, nv):
    """Trains the surrogate for nv levels on columns of a hot-Jupiter g5 spin-up
    with the exact double-gray RT (mixed dynamics, dt 300 s, Rayleigh sponge)."""
    from thor_amd import surrogate as sg
    g, p = sg.hot_jupiter(5, nv, dt=300.0, ns=6, precision=1, cache_dir=CACHE)
    d = capi.Dyn(g, p, variant=args.variant)
    d.set_state(capi.ic_rest(g, p, seed=1003, temperature_noise=1.0, wind_noise=0.1))
    d.set_rt(sg.double_gray())
    d.set_sponge(sg.hot_jupiter_sponge())
    t0 = time.perf_counter()
    X, Y = sg.collect(d, g, nsteps=args.rt_spinup_steps, every=args.rt_sample_every)
    d.close()
    log(f"g5 L{nv} spin-up: {args.rt_spinup_steps} steps ({args.rt_spinup_steps * 300.0 / 86400:.1f} days) with "
        f"exact RT, {len(X)} columns sampled, {time.perf_counter() - t0:.1f} s")
    norm = sg.normalisation(X, Y)
    m = capi.MLP(2 * nv + 1, 256, nv, seed=5, variant=args.variant)
    t0 = time.perf_counter()
    losses = sg.train(m, *sg.apply_norm(X, Y, norm), epochs=args.rt_epochs, batch=512, lr=2e-3, log=log)
    err = sg.heating_error(m, X, Y, norm)
    log(f"training: {args.rt_epochs} epochs in {time.perf_counter() - t0:.1f} s, heating RMS error {err:.3f}")
    return m, norm, {"train_columns": int(len(X)), "epochs": args.rt_epochs,
                     "train_loss_first_last": [losses[0], losses[-1]], "heating_rel_rms_err_trained": round(err, 4),
                     "rms_heating_W_m3": float(np.sqrt((Y ** 2).mean()))}


def at_scale_run(capi, args):
    """Hot Jupiter at glevel 7 (163 842 columns), 48 levels, dt 120 s, 6 acoustic
    substeps, Rayleigh sponge, seeded +-0.1 m/s wind noise: mixed-precision
    dynamics with the MLP surrogate (exact RT for out-of-envelope columns, judge
    every 10 calls) against the FP64 form with the exact RT, from the same state
    over the same simulated days, with the mixed form with the exact RT beside
    them to separate the two sources of difference, and an FP64 twin started
    from the state with every pressure one ulp up: the flow's own sensitivity to
    a last-bit change, which the other differences are set against. The network
    is trained on a g5 L48 exact-RT spin-up: its inputs (T, ln p, cos z per
    column) and outputs do not depend on the horizontal grid. The four runs
    advance a simulated day at a time (each timed on its own); each day the
    differences from the FP64 run are reported point by point and of the zonal
    means (5-degree bands, per level). Reports SDPD and the step's algorithmic
    HBM rate per form."""
    from thor_amd import surrogate as sg
    nv, gl, dt = 48, args.scale_glevel, 120.0
    m, norm, train = _spinup_and_train(capi, args, nv)
    t0 = time.perf_counter()
    g, p64 = sg.hot_jupiter(gl, nv, dt=dt, ns=6, precision=0, cache_dir=CACHE)
    pmx = sg.hot_jupiter_params(gl, dt=dt, ns=6, precision=1)
    log(f"grid g{gl} L{nv}: P={g.P} in {time.perf_counter() - t0:.1f} s")
    st = capi.ic_rest(g, p64, seed=1010, temperature_noise=0.0, wind_noise=0.1)
    st_p = dict(st)
    st_p["pressure"] = np.nextafter(st["pressure"], np.inf)
    nsteps = int(round(args.scale_days * 86400.0 / dt))
    per_day = int(round(86400.0 / dt))
    forms = {"fp64_exact_rt": (p64, capi.FORCING_DOUBLE_GRAY, st),
             "fp64_ulp_exact_rt": (p64, capi.FORCING_DOUBLE_GRAY, st_p),
             "mixed_exact_rt": (pmx, capi.FORCING_DOUBLE_GRAY, st),
             "mixed_surrogate": (pmx, capi.FORCING_RT_SURROGATE, st)}
    ds, walls = {}, {}
    for name, (prm, forcing, s0) in forms.items():
        d = capi.Dyn(g, prm, variant=args.variant)
        d.set_state(s0)
        d.set_rt(sg.double_gray())
        d.set_sponge(sg.hot_jupiter_sponge())
        if forcing == capi.FORCING_RT_SURROGATE:
            d.set_rt_surrogate(m, norm, judge_every=10)
        ds[name], walls[name] = d, 0.0
    lat = np.asarray(g.lonlat).reshape(g.P, 2)[:, 1]
    band = np.minimum(((lat + np.pi / 2) / np.pi * 36).astype(int), 35)
    cnt = np.bincount(band, minlength=36)

    def zonal(x):  # zonal mean per (5-degree latitude band, level)
        return np.stack([np.bincount(band, weights=x[:, l], minlength=36) / cnt for l in range(x.shape[1])], 1)

    def diffs(a, b):
        err = {}
        for k in ("T", "u", "v", "w"):
            diff = b[k] - a[k]
            zd = zonal(b[k]) - zonal(a[k])
            err[k] = {"rms": float(f"{np.sqrt((diff ** 2).mean()):.4e}"), "max": float(f"{np.abs(diff).max():.4e}"),
                      "zonal_mean_rms": float(f"{np.sqrt((zd ** 2).mean()):.4e}")}
        return err

    done, days = 0, []
    while done < nsteps:
        n = min(per_day, nsteps - done)
        for name, (prm, forcing, _) in forms.items():
            t1 = time.perf_counter()
            ds[name].run(n, forcing)
            ds[name].sync()
            walls[name] += time.perf_counter() - t1
        done += n
        diag = {name: diagnostics(g, p64, d.get_state()) for name, d in ds.items()}
        a = diag["fp64_exact_rt"]
        days.append({"day": round(done * dt / 86400.0, 3),
                     **{f"{o}_vs_fp64": diffs(a, diag[o]) for o in ("fp64_ulp_exact_rt", "mixed_exact_rt",
                                                                    "mixed_surrogate")}})
        last = days[-1]
        log(f"day {last['day']:g}: T rms vs FP64: 1-ulp FP64 {last['fp64_ulp_exact_rt_vs_fp64']['T']['rms']:.3e}, "
            f"mixed {last['mixed_exact_rt_vs_fp64']['T']['rms']:.3e}, mixed + surrogate "
            f"{last['mixed_surrogate_vs_fp64']['T']['rms']:.3e} K")
    a = diagnostics(g, p64, ds["fp64_exact_rt"].get_state())
    fp64_scale = {k: {"rms": float(f"{np.sqrt((a[k] ** 2).mean()):.4e}"), "max_abs": float(f"{np.abs(a[k]).max():.4e}")}
                  for k in ("T", "u", "v", "w")}
    out = {}
    for name, (prm, forcing, _) in forms.items():
        d = ds[name]
        sstats = d.surrogate_stats() if forcing == capi.FORCING_RT_SURROGATE else None
        d.kernel_timing(True)
        d.run(10, forcing)
        d.sync()
        ks = d.kernel_stats()
        d.kernel_timing(False)
        d.close()
        step_bytes = sum(k["bytes_per_launch"] * k["launches"] for k in ks if k["name"] not in NOT_REFERENCE) / 10
        ms = 1e3 * walls[name] / nsteps
        gbs = step_bytes / (ms * 1e-3) / 1e9
        tot = sum(k["total_ms"] for k in ks)
        top = sorted(ks, key=lambda k: -k["total_ms"])[:6]
        rt_names = ("profx_double_gray", "profx_double_gray_fallback", "surrogate_profx", "sponge_zonal",
                    "sponge_mean", "sponge_apply")
        out[name] = {"ms_per_step": round(ms, 3), "sdpd": round(dt * nsteps / walls[name], 1),
                     "step_alg_bytes": round(step_bytes), "hbm_gbs_alg": round(gbs, 1),
                     "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                     "top_kernels_us": {k["name"]: round(1e3 * k["total_ms"] / 10, 1) for k in top},
                     "profx_kernels_us_per_step": {k["name"]: round(1e3 * k["total_ms"] / 10, 1) for k in ks
                                                   if k["name"] in rt_names},
                     "kernel_sum_ms_per_step": round(tot / 10, 3)}
        if sstats is not None:
            out[name]["judge"] = {"calls": sstats["calls"], "judged": sstats["judged"],
                                  "fallback_columns_per_call": round(sstats["fallback_columns"] / max(1, sstats["calls"]), 1),
                                  "rms_err_rel_to_rms_heating": round(sstats["rms_err"] / train["rms_heating_W_m3"], 4)}
        log(f"{name}: {out[name]}")
    m.close()
    last = days[-1]
    within = {o: {k: {"point": bool(last[f"{o}_vs_fp64"][k]["rms"] <= last["fp64_ulp_exact_rt_vs_fp64"][k]["rms"]),
                      "zonal": bool(last[f"{o}_vs_fp64"][k]["zonal_mean_rms"]
                                    <= last["fp64_ulp_exact_rt_vs_fp64"][k]["zonal_mean_rms"])}
                  for k in ("T", "u", "v", "w")} for o in ("mixed_exact_rt", "mixed_surrogate")}
    return {"metric": "SDPD and algorithmic HBM rate, hot Jupiter g7 L48: mixed-precision dynamics + MLP RT "
                      "surrogate vs FP64 dynamics + exact double-gray RT",
            "workload": f"hot-Jupiter g{gl} L{nv} ({g.P} columns), dt {dt:g} s, 6 substeps, double-gray RT, Rayleigh "
                        f"sponge (top quarter, 1e-4 1/s, 20 bands), +-0.1 m/s wind noise, {args.scale_days:g} days",
            "steps": nsteps, "surrogate_training": {"spinup": f"g5 L{nv}, {args.rt_spinup_steps} steps of 300 s",
                                                    **train},
            **out, "speedup": round(out["fp64_exact_rt"]["ms_per_step"] / out["mixed_surrogate"]["ms_per_step"], 3),
            "fp64_scale": fp64_scale, "within_ulp_spread_at_end": within, "per_day": days,
            "units": {"T": "K", "u": "m/s", "v": "m/s", "w": "m/s"}}


def _zonal_means(g, x, nbands=36):
    """Zonal means of a [P, nv] (or [P]) field in latitude bands of equal width."""
    lat = np.asarray(g.lonlat).reshape(g.P, 2)[:, 1]
    band = np.minimum(((lat + np.pi / 2) / np.pi * nbands).astype(int), nbands - 1)
    cnt = np.bincount(band, minlength=nbands).astype(np.float64)
    if x.ndim == 1:
        return np.bincount(band, weights=x, minlength=nbands) / np.maximum(cnt, 1)
    return np.stack([np.bincount(band, weights=x[:, k], minlength=nbands) for k in range(x.shape[1])], axis=1) \
        / np.maximum(cnt, 1)[:, None]


def _diffs(g, a, b):
    """RMS and max differences of T, u, v, w and ps between two diagnostic
    sets, point-wise and of the zonal means of T and u."""
    out = {}
    for k in ("T", "u", "v", "w", "ps"):
        d = b[k] - a[k]
        out[k] = {"rms": float(f"{np.sqrt((d ** 2).mean()):.4e}"), "max": float(f"{np.abs(d).max():.4e}")}
    for k in ("T", "u"):
        d = _zonal_means(g, b[k]) - _zonal_means(g, a[k])
        out[k + "_zonal"] = {"rms": float(f"{np.sqrt((d ** 2).mean()):.4e}"), "max": float(f"{np.abs(d).max():.4e}")}
    return out


def error_run(capi, g, p64, pmx, st, args):
    """Three runs of --error-days from the same state: FP64, the mixed form,
    and FP64 from the state with every pressure one ulp up. Each simulated
    day, the mixed form's distance from FP64 is set beside the perturbed FP64
    run's (the flow's own sensitivity to a last-bit change: what precision
    alone can be told apart from)."""
    nsteps = int(round(args.error_days * 86400.0 / args.dt))
    per_day = int(round(86400.0 / args.dt))
    st_p = dict(st)
    st_p["pressure"] = np.nextafter(st["pressure"], np.inf)
    ds = [capi.Dyn(g, q, variant=args.variant) for q in (p64, pmx, p64)]
    for d, s0 in zip(ds, (st, st, st_p)):
        d.set_state(s0)
    t0 = time.perf_counter()
    done, days = 0, []
    while done < nsteps:
        n = min(per_day, nsteps - done)
        for d in ds:
            d.run(n)
        for d in ds:
            d.sync()
        done += n
        a, b, c = (diagnostics(g, q, d.get_state()) for q, d in zip((p64, pmx, p64), ds))
        days.append({"day": round(done * args.dt / 86400.0, 3), "mixed": _diffs(g, a, b), "fp64_ulp": _diffs(g, a, c)})
        log(f"error run: day {days[-1]['day']:g}: T rms mixed {days[-1]['mixed']['T']['rms']:.3e}, "
            f"1-ulp FP64 {days[-1]['fp64_ulp']['T']['rms']:.3e} K ({time.perf_counter() - t0:.1f} s)")
    a = diagnostics(g, p64, ds[0].get_state())
    gd = [d.globdiag() for d in ds]
    for d in ds:
        d.close()
    last = days[-1]
    out = {"days": args.error_days, "steps": nsteps, "final": last,
           "fp64_rms": {k: float(f"{np.sqrt((a[k] ** 2).mean()):.4e}") for k in ("T", "u", "v", "w", "ps")},
           "mixed_within_ulp_spread": {k: bool(last["mixed"][k]["rms"] <= last["fp64_ulp"][k]["rms"])
                                       for k in ("T", "u", "v", "w", "ps", "T_zonal", "u_zonal")},
           "per_day": days,
           "units": {"T": "K", "u": "m/s", "v": "m/s", "w": "m/s", "ps": "Pa"},
           "mass_rel_diff": float(f"{(gd[1][0] - gd[0][0]) / gd[0][0]:.3e}")}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--glevel", type=int, default=6)
    ap.add_argument("--nv", type=int, default=48)
    ap.add_argument("--dt", type=float, default=300.0)
    ap.add_argument("--substeps", type=int, default=8)
    ap.add_argument("--variant", default="", help='"" production build, "pm" parity build')
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both",
                    help="forms to time; the headline is the mixed form when it is timed")
    ap.add_argument("--error-days", type=float, default=10.0,
                    help="simulated days of the FP64-vs-mixed error run (0: skip)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-rt", action="store_true",
                    help="skip the learned-RT leg (hot Jupiter g5 L40: exact RT vs the trained surrogate)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain device-to-device copy of this many MiB (three times); "
                         "calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE against known bytes")
    ap.add_argument("--rt-surrogate", action="store_true",
                    help="instead of the headline: exact double-gray RT vs the trained MLP surrogate at hot-Jupiter "
                         "g5 L40 (spin-up, GPU training, timing, judge statistics)")
    ap.add_argument("--rt-spinup-steps", type=int, default=14400, help="spin-up steps with exact RT (50 days)")
    ap.add_argument("--rt-sample-every", type=int, default=720, help="sample all columns every this many steps")
    ap.add_argument("--rt-epochs", type=int, default=40)
    ap.add_argument("--rt-steps", type=int, default=100, help="timed steps per form")
    ap.add_argument("--at-scale", action="store_true",
                    help="instead of the headline: hot Jupiter g7 L48, dt 120 s, mixed dynamics + MLP surrogate vs "
                         "FP64 + exact RT over --scale-days (surrogate trained on a g5 L48 spin-up)")
    ap.add_argument("--scale-days", type=float, default=5.0)
    ap.add_argument("--scale-glevel", type=int, default=7)
    ap.add_argument("--traffic-summary", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "WORKLOAD"),
                    help="instead of running: print the profiles/hbm_traffic_WORKLOAD.json table from two "
                         "rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE) of a run with --copy-probe-mib")
    args = ap.parse_args()
    if args.traffic_summary:
        fd, wd, wl = args.traffic_summary
        print(json.dumps(traffic_summary(fd, wd, wl, args.copy_probe_mib or 512, []), indent=1))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU (rhombus-sharded) "
                                   "path is a design only, so --gpus must be 1"}))
        sys.exit(2)

    import thor_amd.capi as capi

    if args.rt_surrogate:
        os.makedirs(CACHE, exist_ok=True)
        print(json.dumps(rt_surrogate_run(capi, args)), flush=True)
        return
    if args.at_scale:
        os.makedirs(CACHE, exist_ok=True)
        print(json.dumps(at_scale_run(capi, args)), flush=True)
        return
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
    os.makedirs(CACHE, exist_ok=True)
    t0 = time.perf_counter()
    g = capi.Grid(args.glevel, args.nv, 6.371e6, 36000.0, spring=True, beta=1.15, cache_dir=CACHE)
    log(f"grid g{args.glevel} L{args.nv}: P={g.P} built in {time.perf_counter() - t0:.1f}s")
    base = dict(glevel=args.glevel, timestep=args.dt, n_substeps=args.substeps)
    p64 = capi.default_params(precision=0, **base)
    pmx = capi.default_params(precision=1, **base)
    st = capi.ic_rest(g, p64, seed=1002, temperature_noise=1e-3)
    wl = f"g{args.glevel}L{args.nv}"

    runs = {}
    if args.precision in ("both", "fp64"):
        log("FP64 form:")
        runs["fp64"] = timed_run(capi, g, p64, st, args)
    if args.precision in ("both", "mixed"):
        log("mixed-precision form:")
        runs["mixed"] = timed_run(capi, g, pmx, st, args)
    head = "mixed" if "mixed" in runs else "fp64"
    hr = runs[head]
    roof = roofline(hr, wl + ("_mixed" if head == "mixed" else ""))

    err = None
    if args.precision == "both" and args.error_days > 0:
        err = error_run(capi, g, p64, pmx, st, args)

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        from oracle.oracle import Oracle
        # the oracle's block and column loops on every host thread (same bits
        # as the serial loop), then one thread for the serial figure
        nthr = orc.set_threads(0, args.variant)
        o = Oracle(g, p64, variant=args.variant)
        o.set_state(st)
        tc = time.perf_counter()
        o.steps(2)
        cpu_s = (time.perf_counter() - tc) / 2
        orc.set_threads(1, args.variant)
        o1 = Oracle(g, p64, variant=args.variant)
        o1.set_state(st)
        tc = time.perf_counter()
        o1.steps(1)
        cpu1_s = time.perf_counter() - tc
        orc.set_threads(nthr, args.variant)
        cpu = {"value": round(args.dt / cpu_s, 3), "unit": "sim-days/wall-day", "cores": nthr, "kind": "port",
               "sample": f"2 steps (ProfX + Thor) of the same g{args.glevel} L{args.nv} workload, FP64 oracle "
                         f"(THOR's kernels as a host loop) on {nthr} host threads: {cpu_s:.2f} s per step; "
                         f"1 thread: {cpu1_s:.2f} s per step",
               "value_1_thread": round(args.dt / cpu1_s, 3), "host_cpu": _host_cpu()}
        d2 = capi.Dyn(g, p64, variant=args.variant)
        d2.set_state(st)
        d2.run(1)
        d2.sync()
        gs, os_ = d2.get_state(), o1.get_state()
        # density and pressure relative to their maxima; the momenta relative
        # to the momentum scale rho c_s (near rest their own maxima are
        # round-off residuals of hydrostatic balance)
        scale = np.abs(os_["Rho"]).max() * np.sqrt(p64.Cp / (p64.Cp - p64.Rd) * p64.Rd * p64.Tmean)
        errs = {k: float(np.abs(gs[k] - os_[k]).max() / (scale if k in ("Mh", "Wh", "W") else np.abs(os_[k]).max()))
                for k in ("Mh", "Wh", "W", "Rho", "pressure")}
        cpu["fp64_first_step_err"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
        d2.close()

    mixed = head == "mixed"
    out = {
        "metric": METRIC, "value": round(hr["sdpd"], 1), "unit": "sim-days/wall-day", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(hr["ms_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic (rest state + seeded 1e-3 K noise)",
        "config": {"workload": f"Held-Suarez Earth g{args.glevel} L{args.nv} ({g.P} columns), dt={args.dt:g} s, "
                               f"{args.substeps} acoustic substeps, HyDiff 4 + DivDamp, "
                               + ("mixed precision (FP32 damping operators and momentum deviation, FP64 rest)"
                                  if mixed else "FP64"),
                   "glevel": args.glevel, "nv": args.nv, "dt": args.dt, "n_substeps": args.substeps,
                   "precision": head, "build": args.variant or "production"},
        "roofline": roof,
    }
    if "fp64" in runs:
        out["fp64"] = {"value": round(runs["fp64"]["sdpd"], 1), "ms_per_step": round(runs["fp64"]["ms_step"], 4),
                       "roofline": roofline(runs["fp64"], wl)}
        if mixed:
            out["speedup_vs_fp64"] = round(runs["fp64"]["ms_step"] / hr["ms_step"], 3)
    out["error_vs_fp64"] = err
    out["cpu_baseline"] = cpu
    if not args.no_rt:
        log("learned-RT leg (hot Jupiter g5 L40):")
        out["learned_rt"] = rt_surrogate_run(capi, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
