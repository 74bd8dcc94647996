// Host side of the dynamical core: handle life cycle, layout conversion,
// launch sequence of one large step (ESP::Thor, thor_driver.cu:70-1466),
// ProfX, global diagnostics and per-kernel timing.
#include "../../include/thor_dyn.h"
#include "../../include/thor_mlp.h"
#include "common.h"
#include "dyn_kernels.h"
#include "dyn_kernels_ml.h"
#include "mlp_internal.h"
#include "tracer_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <type_traits>
#include <vector>

namespace thor {

#define HIP_TRY(expr)                                                                                \
    do {                                                                                             \
        hipError_t _e = (expr);                                                                      \
        if (_e != hipSuccess)                                                                        \
            return fail(THOR_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e));             \
    } while (0)

constexpr int kLC = 8;  // levels per thread in the stencil kernels

struct KernelStat {
    long launches = 0;
    double total_ms = 0.0;
    double bytes = 0.0;  // sum over launches of algorithmic bytes
};

// Device scratch of thor_dyn_globdiag_noalloc, one copy per GPU, held until
// the process ends: the per-block partial sums of the five diagnostics
// (kGdMaxBlocks blocks of 2048 cells, so at most 8 388 608 cells per handle)
// followed by the five totals. Host code uses the address that
// hipGetSymbolAddress returns for the current device (gd_scratch_address),
// never the array's name.
constexpr int kGdMaxBlocks = 4096;
constexpr size_t kGdTotals = (size_t)5 * kGdMaxBlocks;
static __device__ __align__(256) double gd_scratch[kGdTotals + 8];

// The address of gd_scratch on the current device, resolved once per device
// (called by thor_dyn_create with the handle's device current).
static hipError_t gd_scratch_address(int dev, double **out) {
    static std::mutex mu;
    static std::map<int, double *> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    double *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(gd_scratch));
    if (e != hipSuccess) return e;
    addr[dev] = ptr;
    *out = ptr;
    return hipSuccess;
}

struct Pending {
    std::string name;
    hipEvent_t a, b;
};

struct Dyn {
    Geo g{};
    thor_params prm{};
    int P = 0, S = 0, nv = 0;
    std::vector<double> alt_host, alth_host;  // Altitude [nv], Altitudeh [nv + 1]
    hipEvent_t ev_in = nullptr, ev_out = nullptr;  // ordering against a caller's stream (thor_dyn_step)
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    // state y(n) and step buffers (see Fields)
    double *Mh, *Wh, *W, *Rho, *p;
    double *Mhk, *WhkA, *WhkB, *Wk, *Rhok, *pk;
    Fields f{};
    double *partials = nullptr;
    int npart = 0;
    int device = 0;                  // the HIP device the handle was created on
    double *gd_scratch = nullptr;    // that device's address of gd_scratch (thor_dyn_globdiag_noalloc)
    bool timing = false;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, KernelStat> stats;
    std::vector<std::string> stat_order;
    double diff_sign = 1.0;
    bool mixed = false;  // THOR_PRECISION_MIXED
    int lc = 16;         // levels per block of the mixed-precision momentum and divergence stencils
    int lc_prep = 16;    // levels per block of the mixed-precision prepare stencil
    int lc_slow = 16;    // levels per block of the chunked slow modes
    int tcb = 64;        // columns per block of the substep Thomas kernel (16, 32 or 64)
    bool slow_lc = false;  // mixed path: level-chunked slow modes (k_divm1_lc + k_slow_lc)
    bool slow_v2 = true;   // mixed path: k_slow_lc2 (two blocks per CU) in place of k_slow_lc
    // Stage kernels that do not depend on the hyperdiffusion chain can run on
    // a second stream beside it (no physics hooks installed): the stage
    // factorisation after the column preparation, and in the mixed form the
    // stage's divergence-damping passes (into dmbuf) from the stage start.
    // Bit for bit the same states, but slower on MI355X (3.53 -> 3.68 ms mixed,
    // 8.46 -> 8.52 ms FP64 at g6 L48: the concurrent kernels slow each other
    // more than they fill the gaps), so off unless THOR_STAGE_OVERLAP=1.
    bool overlap = false;
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_s[4] = {};
    float *dmbuf = nullptr;  // mixed form: DivM of the slow modes, [3][nv][S], when overlapped
    bool slow_lc64 = false;  // FP64 path: k_divm1_ml + k_slow_lc2<false> in place of k_slow<false>
    bool fp64_ml = true;   // FP64 path: horizontal stencils over several levels per block (*_ml kernels)
    int ml_nl = 4;         // FP64 path: levels per block of the divergence-damping pass 0 and momentum (2 or 4)
    bool diff2_split64 = true;  // FP64 path: the last hyperdiffusion pass and the stage divergence-damping pass 0
                                // as two kernels (false: k_diff2_divm0_ml, both in one; 7.86 -> 7.82 ms)
    bool prep64 = true;    // FP64 path: level-chunked prepare k_prepare_lc64 (false: k_prepare_ml; 107.7 -> 63.6 us
                           // at g6 L48; at three blocks per CU it spills, 95 us)
    int tur = 2;           // unroll of the level loops of the fused Thomas kernel (1 or 2; 64 columns)
    int tseg = 16;         // mixed path: columns per block of the segment Thomas kernel (16 or 32; 0: off)
    bool fseg = true;      // mixed path: segment form of the stage's Thomas factorisation
    int tseg64 = 0;        // FP64 path: segment Thomas kernel, 16 columns per block (0: the 64-column kernel;
                           // the segment form is bit for bit the same but slower: its one-lane sweeps run
                           // on 16 lanes per block, 118 against 97 us at g6 L48)
    bool prep_split = false;  // mixed path: k_prepare_lc takes the stage part of hflx_pt from k_slow_lc
    bool diff2_lc = true;     // mixed path: k_diff2_lc + k_divm0_lc<true> instead of k_diff2_divm0<true>
    int cp_lv = 2;            // levels per thread of the column-preparation kernel (1, 2 or 4)
    int sponge_pass = -1;     // hyperdiffusion pass whose output the diffusive sponge reads (-1: none)
    double mid_scale = 1.0;   // mixed form: the factor each intermediate hyperdiffusion pass applies
    // Rayleigh sponge
    bool sponge_on = false;
    SpongePar spg{};
    int *ztab = nullptr;                                // [3][S]
    int *zcnt = nullptr;                                // [nlat] points per latitude bin
    double *ztrig = nullptr, *zkf = nullptr;            // [4][S] sin/cos of lon and lat, [3][nv] rates
    double *zdw = nullptr;                              // [nv][S] explicit sponge: drag on W
    double *ztmp = nullptr, *zout = nullptr, *vbar = nullptr;
    // physics-module hooks and RK-advanced module arrays
    thor_hooks hooks{};
    RKArrays rka{};
    long nstep = 0;  // Thor steps completed by this handle
    int shrink_sponge = 0, t_shrink = 0;  // the sponge's shrinking (thor_sponge_params)
    // forcings
    bool nr_set = false, rt_set = false;
    NRPar nr{};
    BLPar bl{};  // boundary-layer module (RayleighHS), off by default
    RTPar rt{};
    double *insol = nullptr, *asr = nullptr, *olr = nullptr;  // [P] RT column diagnostics
    // RT surrogate
    thor_mlp *sur = nullptr;
    SurPar sp{};
    int judge_every = 0;
    int *sur_mask = nullptr, *sur_nflag = nullptr;
    // out-of-envelope columns of a surrogate call, and the two per-call
    // counters of that list (call k appends with counter k % 2 and clears the other)
    int *sur_list = nullptr, *sur_cnt = nullptr;
    int ncu = 0;  // compute units of the device (persistent surrogate kernel)
    int rt_list_cb = 4;  // columns per block of the exact RT over the surrogate's fallback list
    int rt_cb = 32;      // columns per block of the exact RT over all columns (nv <= 63)
    bool sur_fused = true;  // surrogate ProfX as one kernel (else features, network, heating apart)
    bool sur_conc = false;  // fused surrogate ProfX: the exact RT of the flagged columns on the second stream,
                            // beside the network (flags formed first by k_sur_flags); bit for bit the same, but
                            // slower on MI355X (hot Jupiter g5 L40: 1.408 -> 1.491 ms per step)
    double *qsur = nullptr, *err_part = nullptr;  // [nv][S] surrogate Qheat on judge steps; 3 x nblk partials
    int err_nblk = 0;
    long sur_calls = 0, sur_judged = 0, sur_fallback = 0;
    double err_sq = 0.0, err_max = 0.0, err_cnt = 0.0;
};

// Zero-filled device allocations. hipMemset runs on the null stream, which
// the handle's non-blocking stream does not wait for: the fill is waited for
// here, so a kernel the handle queues next cannot run before it (and have its
// output cleared by it).
static hipError_t dalloc(Dyn &d, double **ptr, size_t n) {
    hipError_t e = hipMalloc((void **)ptr, n * sizeof(double));
    if (e == hipSuccess) {
        d.allocs.push_back(*ptr);
        e = hipMemset(*ptr, 0, n * sizeof(double));
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    }
    return e;
}

static hipError_t falloc(Dyn &d, float **ptr, size_t n) {
    hipError_t e = hipMalloc((void **)ptr, n * sizeof(float));
    if (e == hipSuccess) {
        d.allocs.push_back(*ptr);
        e = hipMemset(*ptr, 0, n * sizeof(float));
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    }
    return e;
}

static hipEvent_t get_event(Dyn &d) {
    if (!d.free_events.empty()) {
        hipEvent_t e = d.free_events.back();
        d.free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
}

static void drain_timing(Dyn &d) {
    if (d.pending.empty()) return;
    hipStreamSynchronize(d.stream);
    for (auto &pd : d.pending) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, pd.a, pd.b);
        d.stats[pd.name].total_ms += ms;
        d.free_events.push_back(pd.a);
        d.free_events.push_back(pd.b);
    }
    d.pending.clear();
}

// Launches a kernel through fn(); when timing is on, brackets it with events
// on the handle's stream and books bytes_per_launch of algorithmic traffic.
// fn enqueues its kernels on `on` (the handle's stream when null); with
// per-kernel timing the events bracket them on that stream.
template <class Fn> static void launch(Dyn &d, const char *name, double bytes, Fn fn, hipStream_t on = nullptr) {
    if (d.timing) {
        auto &st = d.stats[name];
        if (st.launches == 0 && std::find(d.stat_order.begin(), d.stat_order.end(), name) == d.stat_order.end())
            d.stat_order.push_back(name);
        st.launches++;
        st.bytes += bytes;
        Pending pd{name, get_event(d), get_event(d)};
        hipEventRecord(pd.a, on ? on : d.stream);
        fn();
        hipEventRecord(pd.b, on ? on : d.stream);
        d.pending.push_back(pd);
        if (d.pending.size() > 2048) drain_timing(d);
    } else {
        fn();
    }
}

// Rv_fac of the shrinking sponge at step `step` (profx_driver.cu:153-160,
// thor_driver.cu:822-830): 1 before t_shrink, then exp(-(step - t_shrink) dt / (1000 dt)).
static double sponge_shrink(const Dyn &d, long step) {
    double rv = 1;
    if (d.shrink_sponge) {
        const double timestep = d.prm.timestep;
        if (step * timestep >= d.t_shrink * timestep) {
            const double shrink_scale = timestep * 1000;
            rv = std::exp(-(step * timestep - d.t_shrink * timestep) / shrink_scale);
        }
    }
    return rv;
}

static Fields stage_fields(Dyn &d, int rk) {
    Fields f = d.f;
    f.Mh = d.Mh;
    f.Wh = d.Wh;
    f.W = d.W;
    f.Rho = d.Rho;
    f.p = d.p;
    if (rk == 0) {
        f.Mhk = d.Mh;
        f.Whk = d.Wh;
        f.Wk = d.W;
        f.Rhok = d.Rho;
        f.pk = d.p;
        f.Whk_out = d.WhkA;
    } else {
        f.Mhk = d.Mhk;
        f.Whk = rk == 1 ? d.WhkA : d.WhkB;
        f.Wk = d.Wk;
        f.Rhok = d.Rhok;
        f.pk = d.pk;
        f.Whk_out = rk == 1 ? d.WhkB : d.WhkA;
    }
    f.Mhk_out = d.Mhk;
    f.Wk_out = d.Wk;
    f.Rhok_out = d.Rhok;
    f.pk_out = d.pk;
    return f;
}

// Algorithmic bytes booked per launch: the per-cell doubles of the reference
// kernels that each fused kernel replaces (each field array read once and
// written once per reference kernel; geometry excluded), times P * nv cells
// * 8 B. Reference kernels (doubles per cell): Compute_Advec_Cori1 11,
// Compute_Advec_Cori2 11, Compute_Temperature_H_Pt_Geff 12, Diffusion_Op 13
// per pass, Correct_Horizontal 12, DivM_Op 7 (pass 0) + 6 (pass 1),
// Compute_Slow_Modes 40, UpdateRK 19, UpdateRK2 19, Momentum_Eq 13,
// Prepare_Implicit_Vertical 10, Vertical_Eq 12, Density_Pressure_Eqs 22,
// ProfX (Held-Suarez) 24. The per-stage Thomas factorisation has no reference
// counterpart and books its own traffic (hh, gtilh in; three planes out).
// In mixed precision each array counts with its storage width: FP32 for Mhs,
// Adv, diff, diffmh, diffw/rh/pr, divg, DivM, SlowMh (and THOR's vertical-
// diffusion arrays, zero here), FP64 for the rest; below, (n64, n32) pairs.
template <bool MX> static int thor_step_t(Dyn &d) {
    const double N = (double)d.P * d.nv, B = 8.0 * N;
    // bytes per cell for n64 FP64 arrays and n32 mixed-precision arrays
    auto W = [&](double n64, double n32) { return MX ? 8.0 * N * n64 + 4.0 * N * n32 : 8.0 * N * (n64 + n32); };
    const int ns_total = d.prm.n_substeps;
    const int hydiff = d.prm.HyDiff, divdamp = d.prm.DivDampP;
    const dim3 blk(kTile), grid_s((d.P + kTile - 1) / kTile, (d.nv + kLC - 1) / kLC);
    const dim3 grid_t2(8 * d.g.tiles_per_xcd * d.nv);  // LDS-tiled stencils: one tile x one level per block
    // mixed-precision substep stencils: one tile x lc levels per block
    const dim3 grid_lc(8 * d.g.tiles_per_xcd * ((d.nv + d.lc - 1) / d.lc));
    const dim3 grid_lp(8 * d.g.tiles_per_xcd * ((d.nv + d.lc_prep - 1) / d.lc_prep));
    const dim3 grid_ls(8 * d.g.tiles_per_xcd * ((d.nv + d.lc_slow - 1) / d.lc_slow));
    const int lc = d.lc, lcp = d.lc_prep, lcs = d.lc_slow;
    // FP64 multi-level stencils: one tile x NL levels per block
    const bool ml = !MX && d.fp64_ml;
    auto grid_ml = [&](int nl) { return dim3(8 * d.g.tiles_per_xcd * ((d.nv + nl - 1) / nl)); };
    // column_prep: (256 columns, level 0..nv) per block, XCD-major
    const dim3 grid_c(8 * (((d.P + kTile - 1) / kTile + 7) / 8) * (d.nv + 1));
    const dim3 grid_c2(8 * (((d.P + kTile - 1) / kTile + 7) / 8) * ((d.nv + 2) / 2));
    const dim3 grid_c4(8 * (((d.P + kTile - 1) / kTile + 7) / 8) * ((d.nv + 4) / 4));
    const int cb = kThomasCB(d.nv);
    const dim3 grid_t((d.P + cb - 1) / cb);
    const size_t lds_t = 4 * (size_t)(d.nv + 1) * cb * sizeof(double);
    hipStream_t st = d.stream, s2 = d.stream2;
    const int nstep = (int)(d.nstep + 1);
    // physics hooks may read or write any array of the step: no second stream then
    const bool overlap_ok = d.overlap && s2 && !d.hooks.dyn_core_loop_init && !d.hooks.dyn_core_loop_slow_modes
                            && !d.hooks.dyn_core_loop_fast_modes && !d.hooks.dyn_core_loop_end;
    auto view = [&](const Fields &f) {
        thor_step_view v{};
        v.plane = d.S;
        v.point_num = d.P;
        v.nv = d.nv;
        v.mhs_fp32 = MX ? 1 : 0;
        v.Mhk = f.Mhk;
        v.Whk = f.Whk;
        v.Wk = f.Wk;
        v.Rhok = f.Rhok;
        v.pressurek = f.pk;
        v.Mhs = MX ? (const void *)f.mMhs : (const void *)f.Mhs;
        v.Whs = f.Whs;
        v.Rhos = f.Rhos;
        v.pressures = f.ps;
        v.temperature = f.T;
        v.h = f.h;
        v.pt = f.pt;
        v.pth = f.pth;
        return v;
    };
    auto hook_fail = [&](const char *which) {
        return fail(THOR_ERR_HOOK, std::string("thor step: physics hook ") + which + " returned an error");
    };
    const long rk_count = d.rka.n ? *std::max_element(d.rka.count, d.rka.count + d.rka.n) : 0;
    auto rk_arrays = [&](int mode) {
        if (!d.rka.n) return;
        double bytes = 0.0;
        for (int j = 0; j < d.rka.n; j++) bytes += 8.0 * d.rka.count[j] * 3;
        launch(d, mode == 0 ? "rk_arrays_start" : "rk_arrays_end", bytes, [&] {
            k_rk_arrays<<<(unsigned)((rk_count + 255) / 256), blk, 0, st>>>(d.rka, mode);
        });
    };
    if (d.hooks.dyn_core_loop_init) {
        const thor_step_view v = view(stage_fields(d, 0));
        if (d.hooks.dyn_core_loop_init(d.hooks.user, (void *)st, nstep, &v)) return hook_fail("dyn_core_loop_init");
    }
    for (int rk = 0; rk < 3; rk++) {
        const int ns_it = rk == 0 ? 1 : (rk == 1 ? ns_total / 2 : ns_total);
        const double times = rk == 0 ? d.prm.timestep / 3.0 : d.prm.timestep / (double)ns_total;
        Fields f = stage_fields(d, rk);
        const Geo g = d.g;
        // second stream: the divergence-damping passes of the slow modes need
        // only the stage values, the factorisation only the column preparation
        const bool ovl = overlap_ok;
        const bool ovl_divm = ovl && MX && d.slow_lc && d.diff2_lc && divdamp;
        if (ovl) {
            hipEventRecord(d.ev_s[0], st);  // the stage values are ready
            hipStreamWaitEvent(s2, d.ev_s[0], 0);
        }
        if (ovl_divm) {
            launch(d, "divm0_stage", W(4, 3), [&] { k_divm0_lc<true><<<grid_lc, blk, 0, s2>>>(g, f, lc); }, s2);
            launch(d, "divm1", W(0, 6), [&] { k_divm1_lc<<<grid_lc, blk, 0, s2>>>(g, f, d.dmbuf, lc); }, s2);
            hipEventRecord(d.ev_s[1], s2);  // DivM ready
        }
        launch(d, "column_prep", W(12 + 5, 6), [&] {
            if (d.cp_lv == 4) k_column_prep_lv<MX, 4><<<grid_c4, blk, 0, st>>>(g, f);
            else if (d.cp_lv == 2) k_column_prep_lv<MX, 2><<<grid_c2, blk, 0, st>>>(g, f);
            else k_column_prep<MX><<<grid_c, blk, 0, st>>>(g, f);
        });
        auto factor = [&](hipStream_t on) {
            launch(d, "thomas_factor", 5 * B, [&] {
                const int fseg = (d.nv - 1 + 7) / 8;  // interfaces per row of the segment form (32 columns)
                if (MX && d.fseg && d.nv > 1 && fseg <= 8) {
                    const dim3 gs((d.P + 31) / 32);
                    if (fseg <= 6) k_thomas_factor_seg<32, 6><<<gs, 256, 0, on>>>(g, f, times);
                    else k_thomas_factor_seg<32, 8><<<gs, 256, 0, on>>>(g, f, times);
                } else if (cb == 32) k_thomas_factor<32, MX><<<grid_t, 256, lds_t, on>>>(g, f, times);
                else k_thomas_factor<16, MX><<<grid_t, 256, lds_t, on>>>(g, f, times);
            }, on);
        };
        if (ovl) {
            hipEventRecord(d.ev_s[2], st);  // hh and gtilh ready
            hipStreamWaitEvent(s2, d.ev_s[2], 0);
            factor(s2);
            hipEventRecord(d.ev_s[3], s2);  // factorisation ready
        }
        if (g.energy)
            launch(d, "energy_prep", W(10, 3), [&] {
                k_energy_prep<<<dim3((d.P + 255) / 256, d.nv + 1), blk, 0, st>>>(g, f);
            });
        launch(d, "adv_diff1", W(8 + (hydiff ? 7 : 0), 3 + (hydiff ? 6 : 0)), [&] {
            if (MX) k_adv_diff1_lc<<<grid_lc, blk, 0, st>>>(g, f, hydiff, lc);
            else if (ml) k_adv_diff1_ml<2><<<grid_ml(2), blk, 0, st>>>(g, f, hydiff);
            else k_adv_diff1<MX><<<grid_t2, blk, 0, st>>>(g, f, hydiff);
        });
        // HyDiffOrder / 2 - 2 intermediate passes (HyDiffOrder >= 6) between
        // the first pass (above) and the last (below); fd points the last pass
        // at the output of the final one
        Fields fd = f;
        // the last pass's sign, divided by the intermediate passes' scaling (mixed form)
        const double last_sign =
            MX && hydiff ? d.diff_sign / std::pow(d.mid_scale, d.prm.HyDiffOrder / 2 - 2) : d.diff_sign;
        if (hydiff) {
            for (int ih = 0; ih < d.prm.HyDiffOrder / 2 - 2; ih++) {
                auto *src = FS<MX>::diff(fd);
                auto *dst = ih % 2 ? FS<MX>::diff(f) : FS<MX>::diffx(f);
                launch(d, "diff_mid", W(0, 12),
                       [&] { k_diff_mid<MX><<<grid_t2, blk, 0, st>>>(g, src, dst, (float)d.mid_scale); });
                FS<MX>::set_diff(fd, dst);
            }
            // pass s writes diff (s even) or the second array (s odd); the
            // order limits keep the sponge pass's output there until the last pass
            if (d.sponge_pass >= 0)
                FS<MX>::set_dsp(fd, d.sponge_pass % 2 ? FS<MX>::diffx(f) : FS<MX>::diff(f));
        }
        if (MX && d.diff2_lc) {
            // the last hyperdiffusion pass and the stage's divergence-damping
            // pass 0, each over level chunks
            if (hydiff)
                launch(d, "diff2", W(1, 12 + 12), [&] { k_diff2_lc<<<grid_lc, blk, 0, st>>>(g, fd, last_sign, lc); });
            if (divdamp && !ovl_divm)
                launch(d, "divm0_stage", W(4, 3), [&] { k_divm0_lc<true><<<grid_lc, blk, 0, st>>>(g, f, lc); });
        } else if (ml && d.diff2_split64) {
            // FP64: the last hyperdiffusion pass, then divergence-damping pass 0
            // on the stage values, each kernel with its own register budget
            if (hydiff)
                launch(d, "diff2", W(1, 12 + 12),
                       [&] { k_diff2_divm0_ml<2, false><<<grid_ml(2), blk, 0, st>>>(g, fd, hydiff, 0, last_sign); });
            if (divdamp) {
                Fields fk = f;
                fk.Mhs = f.Mhk;
                fk.Whs = f.Whk;
                launch(d, "divm0_stage", W(4, 3), [&] { k_divm0_ml<4><<<grid_ml(4), blk, 0, st>>>(g, fk); });
            }
        } else if (hydiff || divdamp) {
            launch(d, "diff2_divm0", W((hydiff ? 1 : 0) + (divdamp ? 4 : 0), (hydiff ? 12 + 12 : 0) + (divdamp ? 3 : 0)),
                   [&] {
                if (ml) k_diff2_divm0_ml<2><<<grid_ml(2), blk, 0, st>>>(g, fd, hydiff, divdamp, last_sign);
                else k_diff2_divm0<MX><<<grid_t2, blk, 0, st>>>(g, fd, hydiff, divdamp, last_sign);
            });
        }
        if (hydiff && d.prm.VertHyDiff) {
            // vertical hyperdiffusion, after the horizontal passes as in THOR
            const int cbv = vert_diff_columns(d.nv);
            const size_t lds = vert_diff_lds(d.nv, cbv);
            launch(d, "vert_diff", MX ? W(6, 10) : W(6, 7), [&] {
                k_vert_diff<MX><<<dim3((d.P + cbv - 1) / cbv), dim3(cbv, 6), lds, st>>>(g, f, d.prm.VertHyDiffOrder);
            });
        }
        if (d.hooks.dyn_core_loop_slow_modes) {
            const thor_step_view v = view(f);
            if (d.hooks.dyn_core_loop_slow_modes(d.hooks.user, (void *)st, nstep, times, &v))
                return hook_fail("dyn_core_loop_slow_modes");
        }
        if (d.sponge_on && d.spg.mode != 0) {
            // explicit Rayleigh sponge on the stage values, before the slow modes
            Fields fz = f;
            fz.Mh = f.Mhk;
            fz.Rho = f.Rhok;
            fz.W = f.Wk;
            fz.p = f.pk;
            const SpongePar sp = d.spg;
            const dim3 gpl((d.P + kTile - 1) / kTile, d.nv);
            if (rk == 0 || sp.mode == 3) {
                const int nc = sp.ncomp();
                launch(d, "sponge_zonal", 6 * B, [&] {
                    k_sponge_zonal<<<gpl, blk, 0, st>>>(g, fz, sp, d.ztab, d.ztmp);
                });
                launch(d, "sponge_mean", 8.0 * nc * sp.nlat * d.nv * sp.max_count, [&] {
                    k_sponge_reduce<<<d.nv * nc * sp.nlat, 1024, 0, st>>>(sp, d.nv, d.ztmp, d.zout);
                    k_sponge_vbar<<<(nc * sp.nlat * d.nv + 255) / 256, 256, 0, st>>>(sp, d.nv, d.zout, d.vbar);
                });
            }
            launch(d, "sponge_drag", 12 * B, [&] {
                k_sponge_drag<<<gpl, blk, 0, st>>>(g, fz, sp, d.ztab, d.vbar, f.dMh, d.zdw, f.Qheat,
                                                   sponge_shrink(d, nstep));
                k_sponge_dwh<<<dim3((d.P + kTile - 1) / kTile, d.nv + 1), blk, 0, st>>>(g, d.zdw, f.dWh);
            });
        }
        if (MX && d.slow_lc) {
            // DivM from the second stream, or into the idle hyperdiffusion
            // scratch (its first three planes)
            float *dm = ovl_divm ? d.dmbuf : f.mdiff;
            if (ovl_divm) hipStreamWaitEvent(st, d.ev_s[1], 0);
            else if (divdamp)
                launch(d, "divm1", W(0, 6), [&] { k_divm1_lc<<<grid_lc, blk, 0, st>>>(g, f, dm, lc); });
            launch(d, "slow_modes", W(19 + (rk ? 16 : 0), (divdamp ? 6 : 0) + 21 + (rk ? 3 : 0)),
                   [&] {
                       if (d.slow_v2) k_slow_lc2<true><<<grid_ls, blk, 0, st>>>(g, f, dm, divdamp, rk, lcs);
                       else k_slow_lc<<<grid_ls, blk, 0, st>>>(g, f, dm, divdamp, rk, lcs);
                   });
        } else if (!MX && d.slow_lc64) {
            // FP64: DivM over several levels per block into the idle
            // hyperdiffusion scratch, then the slow modes over level chunks
            if (divdamp)
                launch(d, "divm1", W(0, 6), [&] { k_divm1_ml<4><<<grid_ml(4), blk, 0, st>>>(g, f, f.diff); });
            // booked as the reference's slow modes with DivM pass 1 (as the per-level kernel)
            launch(d, "slow_modes", W(19 + (rk ? 16 : 0), (divdamp ? 6 : 0) + 21 + (rk ? 3 : 0)),
                   [&] { k_slow_lc2<false><<<grid_ls, blk, 0, st>>>(g, f, f.diff, divdamp, rk, lcs); });
        } else {
            launch(d, "slow_modes", W(19 + (rk ? 16 : 0), (divdamp ? 6 : 0) + 21 + (rk ? 3 : 0)),
                   [&] { k_slow<MX><<<grid_t2, blk, 0, st>>>(g, f, divdamp, rk); });
        }
        if (rk > 0) rk_arrays(0);
        auto divm0 = [&] {
            launch(d, "divm0", W(1, 6), [&] {
                if (MX) k_divm0_lc<false><<<grid_lc, blk, 0, st>>>(g, f, lc);
                else if (ml && d.ml_nl == 2) k_divm0_ml<2><<<grid_ml(2), blk, 0, st>>>(g, f);
                else if (ml) k_divm0_ml<4><<<grid_ml(4), blk, 0, st>>>(g, f);
                else k_divm0<MX><<<grid_t2, blk, 0, st>>>(g, f);
            });
        };
        if (divdamp) divm0();
        if (ovl) hipStreamWaitEvent(st, d.ev_s[3], 0);
        else factor(st);
        for (int ns = 0; ns < ns_it; ns++) {
            const int next = divdamp && ns < ns_it - 1;
            launch(d, "momentum", W(1, (divdamp ? 6 : 0) + 12), [&] {
                if (MX) k_momentum_lc<<<grid_lc, blk, 0, st>>>(g, f, divdamp, times, lc);
                else if (ml && d.ml_nl == 2) k_momentum_ml<2><<<grid_ml(2), blk, 0, st>>>(g, f, divdamp, times);
                else if (ml) k_momentum_ml<4><<<grid_ml(4), blk, 0, st>>>(g, f, divdamp, times);
                else k_momentum<MX><<<grid_t2, blk, 0, st>>>(g, f, divdamp, times);
            });
            // Prepare_Implicit_Vertical, plus the horizontal flux divergences of
            // Density_Pressure_Eqs (booked with the column kernel below); the
            // split form reads h and pt as FP32 copies and writes no Sd (the
            // mixed vertical solve forms it from hflx_r and SlowRho)
            launch(d, "prepare", MX && d.prep_split ? W(4, 5) : W(7, 3), [&] {
                if (MX && d.prep_split) k_prepare_lc<true><<<grid_lp, blk, 0, st>>>(g, f, lcp);
                else if (MX) k_prepare_lc<false><<<grid_lp, blk, 0, st>>>(g, f, lcp);
                else if (ml && d.prep64) k_prepare_lc64<<<grid_lp, blk, 0, st>>>(g, f, lcp);
                else if (ml) k_prepare_ml<2><<<grid_ml(2), blk, 0, st>>>(g, f);
                else k_prepare<MX><<<grid_t2, blk, 0, st>>>(g, f);
            });
            // Vertical_Eq (12) + Density_Pressure_Eqs (22, of which Mhs and diffpr FP32 in mixed)
            // the solve's plane plus the row maps of the mixed path's sweeps
            const size_t lds = ((size_t)(d.nv + 1) * d.tcb + 2 * 256) * sizeof(double);
            const dim3 grid_th((d.P + d.tcb - 1) / d.tcb);
            // segment kernel: rows of (nv - 1 + NROW - 1) / NROW interfaces
            const int seg16 = (d.nv - 1 + 15) / 16, seg32 = (d.nv - 1 + 7) / 8;
            // mixed: 16 or 32 columns (rows chained); FP64: 16 columns (one-lane sweeps)
            const int cbs = MX ? d.tseg : d.tseg64;
            const bool tseg = d.nv > 1 && (MX ? d.f.thomas_rows && (cbs == 16 ? seg16 <= 4 : cbs == 32 && seg32 <= 8)
                                              : cbs == 16 && seg16 <= 4);
            const size_t lds_seg = ((size_t)(d.nv + 1) * cbs + 2 * 256) * sizeof(double);
            auto seg_solve = [&](auto dp) {
                constexpr bool DP = decltype(dp)::value;
                const dim3 gs((d.P + cbs - 1) / cbs);
                if (cbs == 16 && seg16 <= 3) k_thomas_dp_seg<16, 3, DP, MX><<<gs, 256, lds_seg, st>>>(g, f, times);
                else if (cbs == 16) k_thomas_dp_seg<16, 4, DP, MX><<<gs, 256, lds_seg, st>>>(g, f, times);
                else if (seg32 <= 6) k_thomas_dp_seg<32, 6, DP, MX><<<gs, 256, lds_seg, st>>>(g, f, times);
                else k_thomas_dp_seg<32, 8, DP, MX><<<gs, 256, lds_seg, st>>>(g, f, times);
            };
            if (!d.hooks.dyn_core_loop_fast_modes && tseg) {
                launch(d, "thomas_dp", W(12 + 18, 4), [&] { seg_solve(std::true_type{}); });
            } else if (!d.hooks.dyn_core_loop_fast_modes) {
                launch(d, "thomas_dp", W(12 + 18, 4), [&] {
                    if (d.tcb == 64 && d.tur == 2) k_thomas_dp<64, MX, true, 2><<<grid_th, 256, lds, st>>>(g, f, times);
                    else if (d.tcb == 64) k_thomas_dp<64, MX><<<grid_th, 256, lds, st>>>(g, f, times);
                    else if (d.tcb == 32) k_thomas_dp<32, MX><<<grid_th, 256, lds, st>>>(g, f, times);
                    else k_thomas_dp<16, MX><<<grid_th, 256, lds, st>>>(g, f, times);
                });
            } else {
                // the module runs between the vertical solve and the density/pressure update
                launch(d, "thomas", W(12, 0), [&] {
                    if (tseg) seg_solve(std::false_type{});
                    else if (d.tcb == 64) k_thomas_dp<64, MX, false><<<grid_th, 256, lds, st>>>(g, f, times);
                    else if (d.tcb == 32) k_thomas_dp<32, MX, false><<<grid_th, 256, lds, st>>>(g, f, times);
                    else k_thomas_dp<16, MX, false><<<grid_th, 256, lds, st>>>(g, f, times);
                });
                const thor_step_view v = view(f);
                if (d.hooks.dyn_core_loop_fast_modes(d.hooks.user, (void *)st, nstep, times, &v))
                    return hook_fail("dyn_core_loop_fast_modes");
                launch(d, "density_pressure", W(18, 4),
                       [&] { k_dp<MX><<<dim3((d.P + kTile - 1) / kTile, d.nv), blk, 0, st>>>(g, f, times); });
            }
            // next substep's divergence-damping pass 0
            if (next) divm0();
        }
        launch(d, "update_rk2", W(16, 3), [&] { k_update_rk2<kLC, MX><<<grid_s, blk, 0, st>>>(g, f); });
        rk_arrays(1);
    }
    std::swap(d.Mh, d.Mhk);
    std::swap(d.Wh, d.WhkA);
    std::swap(d.W, d.Wk);
    std::swap(d.Rho, d.Rhok);
    std::swap(d.p, d.pk);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(THOR_ERR_HIP, std::string("thor step launch: ") + hipGetErrorString(e));
    d.nstep++;
    if (d.hooks.dyn_core_loop_end) {
        const thor_step_view v = view(stage_fields(d, 0));
        if (d.hooks.dyn_core_loop_end(d.hooks.user, (void *)st, nstep, &v)) return hook_fail("dyn_core_loop_end");
    }
    return THOR_OK;
}

static int thor_step(Dyn &d) { return d.mixed ? thor_step_t<true>(d) : thor_step_t<false>(d); }


// Double-gray ProfX over all columns (list = nullptr) or over the *count
// columns of list: 32 columns per block up to 63 levels, 16 above (4 LDS
// planes of (nv + 1) x columns doubles within 64 KiB). The list's length is
// known only on the device, so its grid covers every column and the blocks
// past the list's end return at once.
static void launch_rt(Dyn &d, const Geo &g, const Fields &f, double dt, const int *list = nullptr,
                      const int *count = nullptr, hipStream_t on = nullptr) {
    const int cb = list ? d.rt_list_cb : (d.nv <= 63 ? d.rt_cb : 16);
    const size_t lds = 4 * (size_t)(d.nv + 1) * cb * sizeof(double);
    const dim3 grid((d.P + cb - 1) / cb);
    hipStream_t st = on ? on : d.stream;
    auto go = [&](auto kern) { kern<<<grid, 256, lds, st>>>(g, f, d.rt, dt, d.insol, d.asr, d.olr, list, count); };
    if (cb == 32) go(k_profx_rt<32>);
    else if (cb == 16) go(k_profx_rt<16>);
    else if (cb == 8) go(k_profx_rt<8>);
    else go(k_profx_rt<4>);
}

static int profx(Dyn &d, int forcing) {
    if (forcing == THOR_FORCING_NONE) return THOR_OK;
    Fields f = stage_fields(d, 0);
    const Geo g = d.g;
    const dim3 blk(kTile), grid_c((d.P + 63) / 64);  // 64 columns x 4 level rows per block
    const double B = 8.0 * d.P * d.nv;
    const double dt = d.prm.timestep;
    if (d.sponge_on && d.spg.mode == 0) {
        const SpongePar sp = d.spg;
        const dim3 gpl((d.P + kTile - 1) / kTile, d.nv);
        // Wh 2, Mh 3, rho in; W and the three zonal slots out
        launch(d, "sponge_zonal", 10 * B, [&] {
            k_sponge_zonal<<<gpl, blk, 0, d.stream>>>(g, f, sp, d.ztab, d.ztmp);
        });
        if (sp.dmu || sp.dmw || sp.dmt) {
            const int nc = sp.ncomp();
            launch(d, "sponge_mean", 8.0 * nc * sp.nlat * d.nv * sp.max_count, [&] {
                k_sponge_reduce<<<d.nv * nc * sp.nlat, 1024, 0, d.stream>>>(sp, d.nv, d.ztmp, d.zout);
                k_sponge_vbar<<<(nc * sp.nlat * d.nv + 255) / 256, 256, 0, d.stream>>>(sp, d.nv, d.zout, d.vbar);
            });
        }
        // Mh 3, W, rho in; Mh 3, W out; then W 2 in, Wh out
        launch(d, "sponge_apply", 12 * B, [&] {
            k_sponge_apply<<<gpl, blk, 0, d.stream>>>(g, f, sp, d.ztab, d.vbar, dt, sponge_shrink(d, d.nstep + 1));
            k_sponge_wh<<<dim3((d.P + kTile - 1) / kTile, d.nv - 1), blk, 0, d.stream>>>(g, f);
        });
        f.skip_w = 1;
    }
    // The boundary-layer module runs after the forcing, on the pressure ProfX
    // started from: fused into the relaxation forcings; before the RT forcings,
    // which leave Mh alone (the same numbers as after them).
    if (d.bl.on && forcing != THOR_FORCING_HELD_SUAREZ && forcing != THOR_FORCING_NEWTONIAN)
        launch(d, "boundary_layer", 8 * B, [&] { k_bl_rayleigh<<<grid_c, blk, 0, d.stream>>>(g, f, d.bl, dt); });
    if (forcing == THOR_FORCING_HELD_SUAREZ) {
        launch(d, "profx_held_suarez", 24 * B, [&] {
            k_profx_relax<THOR_FORCING_HELD_SUAREZ><<<grid_c, blk, 0, d.stream>>>(g, f, dt, 1, d.nr, d.bl);
        });
    } else if (forcing == THOR_FORCING_NEWTONIAN) {
        if (!d.nr_set) return fail(THOR_ERR_BAD_ARG, "profx: Newtonian forcing parameters not set");
        // W, T, relaxation and pressure per cell: Wh 2, p, rho, Mh untouched -> 4 in, W, p, Qheat out
        launch(d, "profx_newtonian", 7 * B, [&] {
            k_profx_relax<THOR_FORCING_NEWTONIAN><<<grid_c, blk, 0, d.stream>>>(g, f, dt, 1, d.nr, d.bl);
        });
    } else if (forcing == THOR_FORCING_DOUBLE_GRAY) {
        if (!d.rt_set) return fail(THOR_ERR_BAD_ARG, "profx: RT parameters not set");
        // Wh 2, p, rho in; W, T, Qheat, p out; two flux planes written and read twice
        launch(d, "profx_double_gray", 14 * B, [&] { launch_rt(d, g, f, dt); });
    } else if (forcing == THOR_FORCING_RT_SURROGATE) {
        if (!d.rt_set || !d.sur) return fail(THOR_ERR_BAD_ARG, "profx: RT surrogate not set");
        const bool judge = d.judge_every > 0 && d.sur_calls % d.judge_every == 0;
        int *cnt = d.sur_cnt + (d.sur_calls & 1), *cnt_next = d.sur_cnt + ((d.sur_calls & 1) ^ 1);
        d.sur_calls++;
        hipError_t e = mlp::prepare(d.sur, d.P);
        // the network's queued work (training, new parameters) comes first
        if (e == hipSuccess) e = mlp::order_on(d.sur, d.stream);
        if (e != hipSuccess) return fail(THOR_ERR_HIP, std::string("profx surrogate: ") + hipGetErrorString(e));
        const dim3 gp((d.P + kTile - 1) / kTile), gpl((d.P + kTile - 1) / kTile, d.nv);
        uint16_t *X = mlp::input(d.sur);
        const float *Y = mlp::output(d.sur);
        mlp::InferWeights iw;
        const bool fused = d.sur_fused && d.nv <= 63 && mlp::fused_weights(d.sur, &iw);
        // concurrent fallback: the flags first, then the exact RT for the flagged
        // columns on the second stream while the network takes the others
        const bool conc = fused && d.sur_conc && !judge;
        if (conc) {
            launch(d, "surrogate_flags", 2 * B, [&] {
                k_sur_flags<<<(d.P + 255) / 256, 256, 0, d.stream>>>(g, f, d.sp, d.sur_mask, d.sur_list, cnt, cnt_next,
                                                                    d.sur_nflag);
            });
            hipEventRecord(d.ev_s[0], d.stream);
            hipStreamWaitEvent(d.stream2, d.ev_s[0], 0);
            launch(
                d, "profx_double_gray_fallback", 14 * B, [&] { launch_rt(d, g, f, dt, d.sur_list, cnt, d.stream2); },
                d.stream2);
            hipEventRecord(d.ev_s[1], d.stream2);
        }
        if (fused) {
            // features, network and heating in one kernel, one 8-wave block per CU
            launch(d, "surrogate_profx", 10 * B, [&] {
                const unsigned grid = (unsigned)std::max(1, std::min((d.P + 31) / 32, d.ncu));
                const int set_p = judge ? 0 : 1;
                // (n_in tiles, levels per thread) = (kpad / 32, ceil(nv / 16))
                auto go = [&](auto kern) {
                    kern<<<grid, mlp::kInferThreads, 0, d.stream>>>(g, f, d.sp, iw, d.sur_mask, d.sur_list, cnt,
                                                                    cnt_next, d.sur_nflag, dt, set_p,
                                                                    conc ? d.sur_mask : nullptr);
                };
                const int nl = (d.nv + 15) / 16;
                switch (iw.kpad / 32) {
                case 1: go(k_sur_persist<1, 1>); break;
                case 2: go(k_sur_persist<2, 2>); break;
                case 3: go(k_sur_persist<3, 3>); break;
                default:
                    if (nl <= 3) go(k_sur_persist<4, 3>);
                    else go(k_sur_persist<4, 4>);
                    break;
                }
            });
        } else {
        launch(d, "surrogate_features", 6 * B, [&] {
            const size_t lds = (size_t)64 * (2 * d.nv + 1) * sizeof(uint16_t);
            k_sur_features<<<(d.P + 63) / 64, kTile, lds, d.stream>>>(g, f, d.sp, X, d.sur_mask, d.sur_list, cnt,
                                                                       cnt_next, d.sur_nflag);
        });
        // BF16 inputs and weights, 193 kFLOP per column for 81 -> 256 -> 256 -> 40
        launch(d, "surrogate_mlp", 2.0 * d.P * (2 * d.nv + 1 + d.nv) * 2, [&] { mlp::forward_on(d.sur, d.P, d.stream); });
        launch(d, "surrogate_apply", 4 * B, [&] {
            const size_t lds = (size_t)64 * (d.nv + 1) * sizeof(float);
            k_sur_apply<<<(d.P + 63) / 64, kTile, lds, d.stream>>>(g, f, d.sp, Y, d.sur_mask, dt, judge ? 0 : 1);
        });
        }
        if (judge) {
            const size_t n1 = (size_t)d.nv * d.S;
            hipMemcpyAsync(d.qsur, f.Qheat, n1 * sizeof(double), hipMemcpyDeviceToDevice, d.stream);
            launch(d, "profx_double_gray", 14 * B, [&] { launch_rt(d, g, f, dt); });
            k_sur_error<<<d.err_nblk, 256, 0, d.stream>>>(g, f.Qheat, d.qsur, d.sur_mask, (long)n1, d.err_part,
                                                          d.err_part + d.err_nblk, d.err_part + 2 * d.err_nblk);
            std::vector<double> part((size_t)3 * d.err_nblk);
            hipMemcpyAsync(part.data(), d.err_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, d.stream);
            e = hipStreamSynchronize(d.stream);
            if (e != hipSuccess) return fail(THOR_ERR_HIP, std::string("profx surrogate judge: ") + hipGetErrorString(e));
            for (int b = 0; b < d.err_nblk; b++) {
                d.err_sq += part[b];
                d.err_max = std::max(d.err_max, part[d.err_nblk + b]);
                d.err_cnt += part[2 * d.err_nblk + b];
            }
            d.sur_judged++;
        } else if (conc) {
            hipStreamWaitEvent(d.stream, d.ev_s[1], 0);  // the fallback columns are done
        } else {
            // the exact RT for the columns outside the envelope only
            launch(d, "profx_double_gray_fallback", 14 * B, [&] { launch_rt(d, g, f, dt, d.sur_list, cnt); });
        }
    } else {
        return fail(THOR_ERR_BAD_ARG, "unknown forcing");
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(THOR_ERR_HIP, std::string("profx launch: ") + hipGetErrorString(e));
    return THOR_OK;
}

// Resolves a field name to (device pointer, levels, components).
static bool field_ptr(Dyn &d, const std::string &n, double *&ptr, int &nlev, int &ncomp) {
    const int nv = d.nv;
    ncomp = 1;
    nlev = nv;
    if (n == "Mh") { ptr = d.Mh; ncomp = 3; }
    else if (n == "Wh") { ptr = d.Wh; nlev = nv + 1; }
    else if (n == "W") ptr = d.W;
    else if (n == "Rho") ptr = d.Rho;
    else if (n == "pressure") ptr = d.p;
    else if (n == "temperature") ptr = d.f.T;
    else if (n == "pt") ptr = d.f.pt;
    else if (n == "h") ptr = d.f.h;
    else if (n == "gtil") ptr = d.f.gtil;
    else if (n == "hh") { ptr = d.f.hh; nlev = nv + 1; }
    else if (n == "pth") { ptr = d.f.pth; nlev = nv + 1; }
    else if (n == "gtilh") { ptr = d.f.gtilh; nlev = nv + 1; }
    else if (n == "Etotal" && d.f.Etot) ptr = d.f.Etot;
    else if (n == "profx_Qheat") ptr = d.f.Qheat;
    else if (n == "profx_dMh") { ptr = d.f.dMh; ncomp = 3; }
    else if (n == "profx_dWh") { ptr = d.f.dWh; nlev = nv + 1; }
    else return false;
    return true;
}

static int copy_field(Dyn &d, double *dev, int nlev, int ncomp, double *user, bool to_device, bool user_on_device) {
    const long n = (long)d.P * nlev * ncomp;
    const size_t plane = (size_t)ncomp * nlev * d.S;
    if (user_on_device) {
        const int tb = 256;
        const long nb = (n + tb - 1) / tb;
        if (to_device) k_aos_to_soa<<<nb, tb, 0, d.stream>>>(user, dev, d.P, d.S, nlev, ncomp);
        else k_soa_to_aos<<<nb, tb, 0, d.stream>>>(dev, user, d.P, d.S, nlev, ncomp);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(d.stream));
        return THOR_OK;
    }
    std::vector<double> tmp(plane, 0.0);
    if (to_device) {
        for (int p = 0; p < d.P; p++)
            for (int lev = 0; lev < nlev; lev++)
                for (int c = 0; c < ncomp; c++)
                    tmp[((size_t)c * nlev + lev) * d.S + p] = user[((size_t)p * nlev + lev) * ncomp + c];
        HIP_TRY(hipMemcpyAsync(dev, tmp.data(), plane * sizeof(double), hipMemcpyHostToDevice, d.stream));
        HIP_TRY(hipStreamSynchronize(d.stream));
    } else {
        HIP_TRY(hipMemcpyAsync(tmp.data(), dev, plane * sizeof(double), hipMemcpyDeviceToHost, d.stream));
        HIP_TRY(hipStreamSynchronize(d.stream));
        for (int p = 0; p < d.P; p++)
            for (int lev = 0; lev < nlev; lev++)
                for (int c = 0; c < ncomp; c++)
                    user[((size_t)p * nlev + lev) * ncomp + c] = tmp[((size_t)c * nlev + lev) * d.S + p];
    }
    return THOR_OK;
}

// Kd profiles of ESP::initial_values (esp_initial.cu:1292-1335): horizontal
// hyperdiffusion and divergence damping, and the diffusive sponge, whose
// strength adds to Kdh4 when its order is HyDiffOrder and fills Kdh2
// otherwise; the vertical hyperdiffusion strength.
struct KdProfiles {
    std::vector<double> Kdh4, Kdhz, Kdh2, Kdv6;
};
static KdProfiles kd_profiles(const thor_params &prm, const double *Alt, const double *Alth, int nv) {
    KdProfiles k;
    k.Kdh4.assign(nv, 0.0);
    k.Kdhz.assign(nv, 0.0);
    k.Kdh2.assign(nv, 0.0);
    k.Kdv6.assign(nv, 0.0);
    for (int lev = 0; lev < nv; lev++) {
        double dbar = std::sqrt(2 * M_PI / 5) * prm.A / (std::pow(2, prm.glevel));
        k.Kdh4[lev] = (prm.Diffc) * std::pow(dbar, 1.0 * prm.HyDiffOrder) / prm.timestep;
        k.Kdhz[lev] = (prm.DivDampc) * std::pow(dbar, 4.) / prm.timestep;
        if (prm.DiffSponge) {
            const double n = Alt[lev] / Alth[nv], ns = prm.ns_diff_sponge;
            const double ks = n > ns ? prm.Dv_sponge * std::pow(std::sin(0.5 * M_PI * (n - ns) / (1.0 - ns)), 2) : 0.0;
            if (prm.order_diff_sponge == prm.HyDiffOrder)
                k.Kdh4[lev] += ks * std::pow(dbar, 1.0 * prm.order_diff_sponge) / prm.timestep;
            else
                k.Kdh2[lev] = ks * std::pow(dbar, 1.0 * prm.order_diff_sponge) / prm.timestep;
        }
        const double dz = Alth[lev + 1] - Alth[lev];
        k.Kdv6[lev] = prm.Diffc_v * std::pow(dz, 1.0 * prm.VertHyDiffOrder) / prm.timestep;
    }
    return k;
}

// A new large time step: the Kd profiles scale with 1 / timestep. Waits for
// the queued work that reads the old profiles.
static int set_timestep(Dyn &d, double dt) {
    d.prm.timestep = dt;
    const KdProfiles k = kd_profiles(d.prm, d.alt_host.data(), d.alth_host.data(), d.nv);
    const size_t nb = (size_t)d.nv * sizeof(double);
    HIP_TRY(hipStreamSynchronize(d.stream));
    HIP_TRY(hipMemcpy((void *)d.g.Kdh4, k.Kdh4.data(), nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy((void *)d.g.Kdhz, k.Kdhz.data(), nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy((void *)d.g.Kdh2, k.Kdh2.data(), nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy((void *)d.g.Kdv6, k.Kdv6.data(), nb, hipMemcpyHostToDevice));
    if (d.g.lvc) {
        std::vector<float> kz(d.nv);
        for (int lev = 0; lev < d.nv; lev++) kz[lev] = (float)k.Kdhz[lev];
        HIP_TRY(hipMemcpy((void *)(d.g.lvc + 3 * d.nv), kz.data(), d.nv * sizeof(float), hipMemcpyHostToDevice));
    }
    return THOR_OK;
}

// Caller's array in THOR's layout <-> the library's level-planar copy, in
// stream order on the handle's stream (no host synchronisation).
static void transpose_async(Dyn &d, const double *src, double *dst, int nlev, int ncomp, bool to_planar) {
    const long n = (long)d.P * nlev * ncomp;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (to_planar) k_aos_to_soa<<<nb, 256, 0, d.stream>>>(src, dst, d.P, d.S, nlev, ncomp);
    else k_soa_to_aos<<<nb, 256, 0, d.stream>>>(src, dst, d.P, d.S, nlev, ncomp);
}

} // namespace thor

using thor::Dyn;
struct thor_dyn : Dyn {};

extern "C" thor_status thor_dyn_create(const thor_grid_desc *gd, const thor_params *prm, thor_dyn **out) {
    using namespace thor;
    if (!gd || !prm || !out) return fail(THOR_ERR_BAD_ARG, "thor_dyn_create: null argument");
    if (prm->HyDiff && (prm->HyDiffOrder < 4 || prm->HyDiffOrder % 2))
        return fail(THOR_ERR_UNSUPPORTED, "thor_dyn_create: HyDiffOrder must be even and >= 4");
    if (prm->HyDiff && prm->VertHyDiff && (prm->VertHyDiffOrder < 4 || prm->VertHyDiffOrder % 2))
        return fail(THOR_ERR_UNSUPPORTED, "thor_dyn_create: VertHyDiffOrder must be even and >= 4");
    if (prm->HyDiff && prm->DiffSponge
        && (prm->order_diff_sponge < 2 || prm->order_diff_sponge % 2 || prm->order_diff_sponge > prm->HyDiffOrder
            || prm->order_diff_sponge < prm->HyDiffOrder - 4))
        return fail(THOR_ERR_UNSUPPORTED, "thor_dyn_create: order_diff_sponge must be even, >= 2 and within "
                                          "HyDiffOrder - 4 .. HyDiffOrder");
    if (prm->n_substeps < 2 || prm->n_substeps % 2)
        return fail(THOR_ERR_BAD_ARG, "thor_dyn_create: n_substeps must be even and >= 2");
    if (prm->precision != THOR_PRECISION_FP64 && prm->precision != THOR_PRECISION_MIXED)
        return fail(THOR_ERR_BAD_ARG, "thor_dyn_create: unknown precision mode");
    if (prm->thermo_equation != THOR_THERMO_ENTROPY && prm->thermo_equation != THOR_THERMO_ENERGY)
        return fail(THOR_ERR_BAD_ARG, "thor_dyn_create: thermo_equation must be 0 (entropy) or 1 (energy)");
    if (prm->thermo_equation == THOR_THERMO_ENERGY && prm->precision != THOR_PRECISION_FP64)
        return fail(THOR_ERR_UNSUPPORTED, "thor_dyn_create: the energy equation has the FP64 form only");
    if (gd->nv < 3 || gd->point_num < 12) return fail(THOR_ERR_BAD_ARG, "thor_dyn_create: grid too small");
    if (4 * (gd->nv + 1) * kThomasCB(gd->nv) * 8 > 65536)
        return fail(THOR_ERR_UNSUPPORTED, "thor_dyn_create: at most 127 levels (the vertical solve stages 16 "
                                          "columns x (nv + 1) levels x 4 planes in 64 KiB of LDS)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(THOR_ERR_NO_DEVICE, "thor_dyn_create: no HIP device");
    thor_dyn *h = new thor_dyn();
    Dyn &d = *h;
    d.prm = *prm;
    d.mixed = prm->precision == THOR_PRECISION_MIXED;
    // Levels per block of the mixed-precision substep stencils (tuning knob).
    if (const char *e = std::getenv("THOR_STENCIL_LEVELS")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= kLcMax) d.lc = d.lc_prep = d.lc_slow = v;
    }
    if (const char *e = std::getenv("THOR_SLOW_LEVELS")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= kLcMax) d.lc_slow = v;
    }
    // Columns per block of the substep Thomas kernel (tuning knob).
    if (const char *e = std::getenv("THOR_THOMAS_COLUMNS")) {
        const int v = std::atoi(e);
        if (v == 16 || v == 32 || v == 64) d.tcb = v;
    }
    d.P = gd->point_num;
    d.nv = gd->nv;
    d.S = (d.P + kAlign - 1) / kAlign * kAlign;
    const int P = d.P, S = d.S, nv = d.nv;
    auto cleanup_fail = [&](int code) {
        thor_dyn_destroy(h);
        return code;
    };
    if (hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) != hipSuccess
        || hipEventCreateWithFlags(&d.ev_in, hipEventDisableTiming) != hipSuccess
        || hipEventCreateWithFlags(&d.ev_out, hipEventDisableTiming) != hipSuccess)
        return cleanup_fail(fail(THOR_ERR_HIP, "thor_dyn_create: stream"));
    // the second stream, default priority like the first
    if (hipStreamCreateWithFlags(&d.stream2, hipStreamNonBlocking) != hipSuccess)
        return cleanup_fail(fail(THOR_ERR_HIP, "thor_dyn_create: second stream"));
    for (auto &ev : d.ev_s)
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
            return cleanup_fail(fail(THOR_ERR_HIP, "thor_dyn_create: stream events"));
    if (const char *e = std::getenv("THOR_STAGE_OVERLAP")) d.overlap = std::atoi(e) != 0;
    if (hipGetDevice(&d.device) != hipSuccess
        || hipDeviceGetAttribute(&d.ncu, hipDeviceAttributeMultiprocessorCount, d.device) != hipSuccess)
        return cleanup_fail(fail(THOR_ERR_HIP, "thor_dyn_create: device attributes"));
    if (const hipError_t e = gd_scratch_address(d.device, &d.gd_scratch); e != hipSuccess)
        return cleanup_fail(fail(THOR_ERR_HIP, std::string("thor_dyn_create: hipGetSymbolAddress(gd_scratch): ")
                                                   + hipGetErrorString(e)));

    // Geometry in level-planar layout.
    std::vector<int> ring((size_t)6 * S, 0), nfc(S, 0);
    std::vector<char> pent(P, 0);
    for (int k = 0; k < 12; k++)
        if (gd->pent_ind[k] >= 0 && gd->pent_ind[k] < P) pent[gd->pent_ind[k]] = 1;
    for (int p = 0; p < P; p++) {
        nfc[p] = pent[p] ? 5 : 6;
        for (int m = 0; m < 6; m++) {
            int q = gd->point_local[p * 6 + m];
            if (pent[p] && m == 5) q = gd->point_local[p * 6 + 4];
            if (q < 0 || q >= P) return cleanup_fail(fail(THOR_ERR_BAD_ARG, "thor_dyn_create: bad point_local"));
            ring[(size_t)m * S + p] = q;
        }
    }
    auto planar = [&](const double *src, int ncoef) {
        std::vector<double> v((size_t)ncoef * S, 0.0);
        for (int p = 0; p < P; p++)
            for (int c = 0; c < ncoef; c++) v[(size_t)c * S + p] = src[(size_t)p * ncoef + c];
        return v;
    };
    // Kd profiles (esp_initial.cu:1292-1335)
    d.alt_host.assign(gd->Altitude, gd->Altitude + nv);
    d.alth_host.assign(gd->Altitudeh, gd->Altitudeh + nv + 1);
    const KdProfiles kd = kd_profiles(*prm, gd->Altitude, gd->Altitudeh, nv);
    const std::vector<double> &Kdh4 = kd.Kdh4, &Kdhz = kd.Kdhz, &Kdh2 = kd.Kdh2, &Kdv6 = kd.Kdv6;
    d.sponge_pass = prm->HyDiff && prm->DiffSponge && prm->order_diff_sponge != prm->HyDiffOrder
                        ? prm->order_diff_sponge / 2 - 1 : -1;
    d.g.ssp = (prm->order_diff_sponge / 2 + 1) % 2 ? -1.0 : 1.0;
    // mixed form, HyDiffOrder >= 6: each intermediate pass scales its FP32
    // output by dbar^2 (k_diff_mid); the last pass and the sponge undo it
    d.g.dsp_scale = 1.0;
    if (prm->precision == THOR_PRECISION_MIXED && prm->HyDiff && prm->HyDiffOrder >= 6) {
        const double dbar = std::sqrt(2 * M_PI / 5) * prm->A / (std::pow(2, prm->glevel));
        // the FP32 value the passes multiply by, so the undo divides by the same number
        d.mid_scale = (double)(float)(dbar * dbar);
        if (d.sponge_pass >= 1) d.g.dsp_scale = 1.0 / std::pow(d.mid_scale, d.sponge_pass);
    }
    d.diff_sign = std::pow(-1.0, prm->HyDiffOrder / 2);
    auto up_d = [&](const std::vector<double> &v, const double **dst) -> bool {
        double *ptr;
        if (hipMalloc((void **)&ptr, v.size() * sizeof(double)) != hipSuccess) return false;
        d.allocs.push_back(ptr);
        if (hipMemcpy(ptr, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return false;
        *dst = ptr;
        return true;
    };
    auto up_i = [&](const std::vector<int> &v, const int **dst) -> bool {
        int *ptr;
        if (hipMalloc((void **)&ptr, v.size() * sizeof(int)) != hipSuccess) return false;
        d.allocs.push_back(ptr);
        if (hipMemcpy(ptr, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return false;
        *dst = ptr;
        return true;
    };
    Geo &g = d.g;
    g.P = P;
    g.S = S;
    g.nv = nv;
    g.nr = gd->nr;
    g.tiles_per_xcd = (gd->nr + 1 + 7) / 8;
    g.A = prm->A;
    g.Omega = prm->Omega;
    g.Gravit = prm->Gravit;
    g.P_Ref = prm->P_Ref;
    g.Rd = prm->Rd;
    g.Cp = prm->Cp;
    g.deep = prm->DeepModel;
    g.nonhydro = prm->NonHydro;
    g.ghv = prm->GravHeightVar;
    g.energy = prm->thermo_equation == THOR_THERMO_ENERGY;
    std::vector<int> halo((size_t)gd->nr * NHALO);
    for (int r = 0; r < gd->nr; r++)
        for (int t = 0; t < NHALO; t++) halo[(size_t)r * NHALO + t] = gd->maps[(size_t)r * NHL2 + halo_pos(t)];
    if (gd->nr * kTile + 2 != P) return cleanup_fail(fail(THOR_ERR_BAD_ARG, "thor_dyn_create: P != 256 nr + 2"));
    // The level-chunked slow-mode kernel uses the gradient weights for the
    // neighbours of the divergence too; THOR's operators share them, checked
    // here bit for bit.
    bool shared_weights = true;
    for (int p = 0; p < P && shared_weights; p++)
        if (std::memcmp(gd->div + (size_t)p * 21 + 3, gd->grad + (size_t)p * 21 + 3, 18 * sizeof(double)) != 0)
            shared_weights = false;
    d.slow_lc = d.mixed && shared_weights;
    d.slow_lc64 = !d.mixed && shared_weights;
    if (const char *e = std::getenv("THOR_THOMAS_UNROLL")) {
        const int v = std::atoi(e);
        if (v == 1 || v == 2) d.tur = v;
    }
    if (const char *e = std::getenv("THOR_SLOW_CHUNKED")) {
        d.slow_lc = d.slow_lc && std::atoi(e) != 0;
        d.slow_lc64 = d.slow_lc64 && std::atoi(e) != 0;
    }
    if (const char *e = std::getenv("THOR_SLOW_V2")) d.slow_v2 = std::atoi(e) != 0;
    d.prep_split = d.slow_lc;
    if (const char *e = std::getenv("THOR_DIFF2_CHUNKED")) d.diff2_lc = std::atoi(e) != 0;
    d.f.thomas_rows = 1;
    if (const char *e = std::getenv("THOR_THOMAS_ROWS")) d.f.thomas_rows = std::atoi(e) != 0;
    if (const char *e = std::getenv("THOR_FACTOR_SEG")) d.fseg = std::atoi(e) != 0;
    if (const char *e = std::getenv("THOR_THOMAS_SEG64")) {
        const int v = std::atoi(e);
        if (v == 0 || v == 16) d.tseg64 = v;
    }
    if (const char *e = std::getenv("THOR_THOMAS_SEG")) {
        const int v = std::atoi(e);
        if (v == 0 || v == 16 || v == 32) d.tseg = v;
    }
    if (const char *e = std::getenv("THOR_COLPREP_LEVELS")) {
        const int v = std::atoi(e);
        if (v == 1 || v == 2 || v == 4) d.cp_lv = v;
    }
    if (const char *e = std::getenv("THOR_PREPARE_SPLIT")) d.prep_split = d.prep_split && std::atoi(e) != 0;
    if (const char *e = std::getenv("THOR_FP64_MULTILEVEL")) d.fp64_ml = std::atoi(e) != 0;
    if (const char *e = std::getenv("THOR_ML_LEVELS")) {
        const int v = std::atoi(e);
        if (v == 2 || v == 4) d.ml_nl = v;
    }
    if (const char *e = std::getenv("THOR_PREPARE_LC64")) d.prep64 = std::atoi(e) != 0;
    if (const char *e = std::getenv("THOR_DIFF2_SPLIT64")) d.diff2_split64 = std::atoi(e) != 0;
    if (const char *e = std::getenv("THOR_RT_LIST_COLUMNS")) {
        const int v = std::atoi(e);
        if (v == 4 || v == 8 || v == 16 || v == 32) d.rt_list_cb = v;
    }
    if (const char *e = std::getenv("THOR_RT_COLUMNS")) {
        const int v = std::atoi(e);
        if (v == 4 || v == 8 || v == 16 || v == 32) d.rt_cb = v;
    }
    if (const char *e = std::getenv("THOR_SURROGATE_FUSED")) d.sur_fused = std::atoi(e) != 0;
    if (const char *e = std::getenv("THOR_SURROGATE_CONCURRENT")) d.sur_conc = std::atoi(e) != 0;
    bool ok = up_i(ring, &g.ring) && up_i(nfc, &g.nfc) && up_i(halo, &g.halo) && up_d(planar(gd->div, 21), &g.div)
              && up_d(planar(gd->grad, 21), &g.grad) && up_d(planar(gd->areas, 18), &g.areas)
              && up_d(planar(gd->nvecoa, 18), &g.noa) && up_d(planar(gd->nvecti, 18), &g.nti)
              && up_d(planar(gd->nvecte, 18), &g.nte) && up_d(planar(gd->areasTr, 6), &g.areasTr)
              && up_d(planar(gd->func_r, 3), &g.fr) && up_d(planar(gd->lonlat, 2), &g.lonlat)
              && up_d(planar(gd->areasT, 1), &g.areasT)
              && up_d(std::vector<double>(gd->Altitude, gd->Altitude + nv), &g.Alt)
              && up_d(std::vector<double>(gd->Altitudeh, gd->Altitudeh + nv + 1), &g.Alth)
              && up_d(Kdh4, &g.Kdh4) && up_d(Kdhz, &g.Kdhz) && up_d(Kdv6, &g.Kdv6) && up_d(Kdh2, &g.Kdh2);
    if (!ok) return cleanup_fail(fail(THOR_ERR_HIP, "thor_dyn_create: geometry upload"));

    const size_t n1 = (size_t)nv * S, ni = (size_t)(nv + 1) * S, n3 = 3 * n1;
    Fields &f = d.f;
    hipError_t e = hipSuccess;
    auto A = [&](double **ptr, size_t n) {
        if (e == hipSuccess) e = dalloc(d, ptr, n);
    };
    A(&d.Mh, n3); A(&d.Wh, ni); A(&d.W, n1); A(&d.Rho, n1); A(&d.p, n1);
    A(&d.Mhk, n3); A(&d.WhkA, ni); A(&d.WhkB, ni); A(&d.Wk, n1); A(&d.Rhok, n1); A(&d.pk, n1);
    A(&f.Mhs, n3); A(&f.Whs, ni); A(&f.Rhos, n1); A(&f.ps, n1);
    A(&f.T, n1); A(&f.h, n1); A(&f.pt, n1); A(&f.gtil, n1);
    A(&f.hh, ni); A(&f.pth, ni); A(&f.gtilh, ni);
    A(&f.Adv, n3); A(&f.diff, 6 * n1); A(&f.diffmh, n3);
    A(&f.diffw, n1); A(&f.diffrh, n1); A(&f.diffpr, n1); A(&f.divg, n3);
    A(&f.SlowMh, n3); A(&f.SlowWh, ni); A(&f.SlowRho, n1); A(&f.Slowp, n1);
    A(&f.Sp, n1); A(&f.Sd, n1); A(&f.hflx_r, n1); A(&f.hflx_pt, n1); A(&f.tAA, n1); A(&f.tCC, n1); A(&f.tTT, n1);
    A(&f.Qheat, n1); A(&f.dMh, n3); A(&f.dWh, ni);
    if (d.mixed) A(&f.hflx_ptk, n1);
    if (g.energy) A(&f.Etot, n1);
    const bool diff_mid = prm->HyDiff && prm->HyDiffOrder >= 6;  // intermediate passes alternate two arrays
    if (diff_mid && !d.mixed) A(&f.diffx, 6 * n1);
    if (prm->HyDiff && prm->VertHyDiff && !d.mixed) A(&f.diffmv, n3);
    const long ncell = (long)P * nv;
    d.npart = (int)((ncell + 2047) / 2048);
    A(&d.partials, (size_t)5 * d.npart);
    if (e == hipSuccess) {
        e = hipMalloc((void **)&f.flags, sizeof(unsigned int));
        if (e == hipSuccess) {
            d.allocs.push_back(f.flags);
            e = hipMemset(f.flags, 0, sizeof(unsigned int));
            if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        }
    }
    if (e != hipSuccess) return cleanup_fail(fail(THOR_ERR_HIP, std::string("thor_dyn_create: alloc: ") + hipGetErrorString(e)));
    {
        // per-level constants of the chunked slow modes (both forms)
        double *slv = nullptr;
        if (e == hipSuccess) e = dalloc(d, &slv, (size_t)6 * nv);
        if (e == hipSuccess) {
            k_slow_levels<<<(nv + 63) / 64, 64, 0, d.stream>>>(g, slv);
            e = hipGetLastError();
            g.slv = slv;
        }
        if (e != hipSuccess)
            return cleanup_fail(fail(THOR_ERR_HIP, std::string("thor_dyn_create: level table: ") + hipGetErrorString(e)));
    }
    if (d.mixed) {
        auto F = [&](float **ptr, size_t n) {
            if (e == hipSuccess) e = falloc(d, ptr, n);
        };
        F(&f.mMhs, n3); F(&f.mAdv, n3); F(&f.mdiff, 6 * n1); F(&f.mdiffmh, n3);
        F(&f.mdiffw, n1); F(&f.mdiffrh, n1); F(&f.mdiffpr, n1); F(&f.mdivg, n3); F(&f.mSlowMh, n3);
        F(&f.mhpt, 2 * n1);
        F(&d.dmbuf, n3);
        if (diff_mid) F(&f.mdiffx, 6 * n1);
        float *wlap = nullptr, *wdd = nullptr, *udd = nullptr, *lvc = nullptr;
        F(&wlap, (size_t)7 * S); F(&wdd, (size_t)63 * S); F(&udd, (size_t)21 * S); F(&lvc, (size_t)4 * nv);
        if (e == hipSuccess) {
            // per-level constants, with the same double arithmetic as rscale_at and divm_w
            std::vector<float> lv(4 * (size_t)nv);
            for (int lev = 0; lev < nv; lev++) {
                const double rsd = prm->DeepModel ? prm->A / (gd->Altitude[lev] + prm->A) : 1.0;
                lv[lev] = (float)rsd;
                lv[nv + lev] = (float)(rsd * rsd);
                lv[2 * nv + lev] = (float)(1.0 / (gd->Altitudeh[lev] - gd->Altitudeh[lev + 1]));
                lv[3 * nv + lev] = (float)Kdhz[lev];
            }
            e = hipMemcpy(lvc, lv.data(), lv.size() * sizeof(float), hipMemcpyHostToDevice);
            g.lvc = lvc;
        }
        if (e == hipSuccess) {
            Geo gu = g;
            gu.deep = 0;  // rscale = 1
            k_probe_weights<<<(P + 255) / 256, 256, 0, d.stream>>>(gu, wlap, wdd, udd);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(d.stream);
            g.wlap = wlap;
            g.wdd = wdd;
            g.udd = udd;
        }
        if (e != hipSuccess)
            return cleanup_fail(fail(THOR_ERR_HIP, std::string("thor_dyn_create: mixed-precision set-up: ") + hipGetErrorString(e)));
    }
    *out = h;
    return THOR_OK;
}

extern "C" void thor_dyn_destroy(thor_dyn *h) {
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->sur) thor::mlp::detach(h->sur, h->stream);
    for (auto &pd : h->pending) {
        hipEventDestroy(pd.a);
        hipEventDestroy(pd.b);
    }
    for (auto ev : h->free_events) hipEventDestroy(ev);
    for (void *ptr : h->allocs) hipFree(ptr);
    if (h->ev_in) hipEventDestroy(h->ev_in);
    if (h->ev_out) hipEventDestroy(h->ev_out);
    for (auto ev : h->ev_s)
        if (ev) hipEventDestroy(ev);
    if (h->stream2) {
        hipStreamSynchronize(h->stream2);
        hipStreamDestroy(h->stream2);
    }
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

static thor_status state_io(thor_dyn *h, const thor_state *s, bool to_device) {
    using namespace thor;
    if (!h || !s) return fail(THOR_ERR_BAD_ARG, "state: null argument");
    double *user[5] = {s->Mh, s->Wh, s->W, s->Rho, s->pressure};
    const char *names[5] = {"Mh", "Wh", "W", "Rho", "pressure"};
    for (int i = 0; i < 5; i++) {
        if (!user[i]) return fail(THOR_ERR_BAD_ARG, "state: null field pointer");
        double *dev;
        int nlev, ncomp;
        field_ptr(*h, names[i], dev, nlev, ncomp);
        int rc = copy_field(*h, dev, nlev, ncomp, user[i], to_device, s->on_device != 0);
        if (rc != THOR_OK) return rc;
    }
    return THOR_OK;
}

extern "C" thor_status thor_dyn_set_state(thor_dyn *h, const thor_state *s) { return state_io(h, s, true); }
extern "C" thor_status thor_dyn_get_state(thor_dyn *h, const thor_state *s) { return state_io(h, s, false); }

extern "C" thor_status thor_dyn_thor(thor_dyn *h) {
    if (!h) return thor::fail(THOR_ERR_BAD_ARG, "thor_dyn_thor: null handle");
    return thor::thor_step(*h);
}

extern "C" thor_status thor_dyn_step(thor_dyn *h, const thor_state *state, const thor_forcing *forcing, double dt,
                                     int n_substeps, void *stream) {
    using namespace thor;
    if (!h || !state) return fail(THOR_ERR_BAD_ARG, "thor_dyn_step: null argument");
    if (!state->on_device) return fail(THOR_ERR_BAD_ARG, "thor_dyn_step: the state must be in device memory");
    double *user[5] = {state->Mh, state->Wh, state->W, state->Rho, state->pressure};
    for (double *u : user)
        if (!u) return fail(THOR_ERR_BAD_ARG, "thor_dyn_step: null state field");
    if (forcing && (!forcing->Qheat || !forcing->dMh || !forcing->dWh))
        return fail(THOR_ERR_BAD_ARG, "thor_dyn_step: null forcing field");
    if (n_substeps < 0 || (n_substeps > 0 && (n_substeps < 2 || n_substeps % 2)) || dt < 0 || std::isnan(dt))
        return fail(THOR_ERR_BAD_ARG, "thor_dyn_step: need dt >= 0 and n_substeps 0 or even >= 2");
    Dyn &d = *h;
    if (dt > 0 && dt != d.prm.timestep) {
        const int rc = set_timestep(d, dt);
        if (rc != THOR_OK) return rc;
    }
    if (n_substeps > 0) d.prm.n_substeps = n_substeps;
    // the caller's queued work (its ProfX, its modules) first
    const hipStream_t cs = (hipStream_t)stream;
    HIP_TRY(hipEventRecord(d.ev_in, cs));
    HIP_TRY(hipStreamWaitEvent(d.stream, d.ev_in, 0));
    const char *names[5] = {"Mh", "Wh", "W", "Rho", "pressure"};
    double *dev[5];
    int nlev[5], ncomp[5];
    for (int i = 0; i < 5; i++) {
        field_ptr(d, names[i], dev[i], nlev[i], ncomp[i]);
        transpose_async(d, user[i], dev[i], nlev[i], ncomp[i], true);
    }
    if (forcing) {
        transpose_async(d, forcing->Qheat, d.f.Qheat, d.nv, 1, true);
        transpose_async(d, forcing->dMh, d.f.dMh, d.nv, 3, true);
        transpose_async(d, forcing->dWh, d.f.dWh, d.nv + 1, 1, true);
    }
    HIP_TRY(hipGetLastError());
    const int rc = thor_step(d);
    if (rc != THOR_OK) return rc;
    for (int i = 0; i < 5; i++) {
        field_ptr(d, names[i], dev[i], nlev[i], ncomp[i]);  // the step swapped y(n) with the stage arrays
        transpose_async(d, dev[i], user[i], nlev[i], ncomp[i], false);
    }
    HIP_TRY(hipGetLastError());
    // the caller's later work on its stream sees the new state
    HIP_TRY(hipEventRecord(d.ev_out, d.stream));
    HIP_TRY(hipStreamWaitEvent(cs, d.ev_out, 0));
    return THOR_OK;
}

extern "C" thor_status thor_dyn_profx(thor_dyn *h, int forcing) {
    if (!h) return thor::fail(THOR_ERR_BAD_ARG, "thor_dyn_profx: null handle");
    return thor::profx(*h, forcing);
}

extern "C" thor_status thor_dyn_run(thor_dyn *h, int nsteps, int forcing) {
    if (!h || nsteps < 0) return thor::fail(THOR_ERR_BAD_ARG, "thor_dyn_run: bad argument");
    for (int i = 0; i < nsteps; i++) {
        int rc = thor::profx(*h, forcing);
        if (rc != THOR_OK) return rc;
        rc = thor::thor_step(*h);
        if (rc != THOR_OK) return rc;
    }
    return THOR_OK;
}

extern "C" thor_status thor_dyn_set_newtonian(thor_dyn *h, const thor_newtonian_params *q) {
    using namespace thor;
    if (!h || !q) return fail(THOR_ERR_BAD_ARG, "set_newtonian: null argument");
    if (!(q->p_top > 0 && q->p_bottom > q->p_top && q->p_contrast > 0 && q->p_contrast < q->p_bottom &&
          q->tau_top > 0 && q->tau_bottom > 0))
        return fail(THOR_ERR_BAD_ARG, "set_newtonian: need 0 < p_top < p_bottom, 0 < p_contrast < p_bottom, tau > 0");
    NRPar &n = h->nr;
    n.T_mean = q->T_mean;
    n.dT = q->dT;
    n.lon_sub = q->lon_sub;
    n.lnp_top = std::log(q->p_top);
    n.lnp_bot = std::log(q->p_bottom);
    n.lnp_con = std::log(q->p_contrast);
    n.lntau_top = std::log(q->tau_top);
    n.lntau_bot = std::log(q->tau_bottom);
    h->nr_set = true;
    return THOR_OK;
}

extern "C" thor_status thor_dyn_set_boundary_layer(thor_dyn *h, const thor_bl_params *q) {
    using namespace thor;
    if (!h) return fail(THOR_ERR_BAD_ARG, "set_boundary_layer: null handle");
    if (!q) {
        h->bl.on = 0;
        return THOR_OK;
    }
    if (!(q->surf_drag >= 0 && q->bl_sigma >= 0 && q->bl_sigma < 1))
        return fail(THOR_ERR_BAD_ARG, "set_boundary_layer: need surf_drag >= 0 and 0 <= bl_sigma < 1");
    h->bl.on = 1;
    h->bl.surf_drag = q->surf_drag;
    h->bl.bl_sigma = q->bl_sigma;
    return THOR_OK;
}

extern "C" thor_status thor_dyn_set_rt(thor_dyn *h, const thor_rt_params *q) {
    using namespace thor;
    if (!h || !q) return fail(THOR_ERR_BAD_ARG, "set_rt: null argument");
    if (!(q->kappa_sw >= 0 && q->kappa_lw >= 0 && q->f_lw > 0 && q->diff_ang > 0 && q->ps0 > 0 && q->incflx >= 0))
        return fail(THOR_ERR_BAD_ARG, "set_rt: need kappa >= 0, f_lw > 0, diff_ang > 0, ps0 > 0, incflx >= 0");
    if (!h->insol) {
        const size_t ni = (size_t)(h->nv + 1) * h->S;
        hipError_t e = dalloc(*h, &h->insol, h->S);
        if (e == hipSuccess) e = dalloc(*h, &h->asr, h->S);
        if (e == hipSuccess) e = dalloc(*h, &h->olr, h->S);
        if (e != hipSuccess) return fail(THOR_ERR_HIP, std::string("set_rt: alloc: ") + hipGetErrorString(e));
    }
    h->rt = RTPar{q->kappa_sw, q->kappa_lw, q->n_sw, q->n_lw, q->f_lw, q->diff_ang,
                  q->Tint, q->albedo, q->incflx, q->ps0, q->lon_sub};
    h->rt_set = true;
    return THOR_OK;
}

extern "C" thor_status thor_dyn_get_rt_column(thor_dyn *h, const char *name, double *dst, long n) {
    using namespace thor;
    if (!h || !name || !dst) return fail(THOR_ERR_BAD_ARG, "get_rt_column: null argument");
    if (!h->rt_set) return fail(THOR_ERR_BAD_ARG, "get_rt_column: RT not set");
    if (n != h->P) return fail(THOR_ERR_BAD_ARG, "get_rt_column: n must be the point count");
    const std::string s(name);
    const double *src = s == "insol" ? h->insol : s == "ASR" ? h->asr : s == "OLR" ? h->olr : nullptr;
    if (!src) return fail(THOR_ERR_BAD_ARG, "get_rt_column: name must be insol, ASR or OLR");
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost));
    return THOR_OK;
}

extern "C" thor_status thor_dyn_set_rt_surrogate(thor_dyn *h, thor_mlp *m, const thor_rt_surrogate_params *q) {
    using namespace thor;
    if (!h || !m || !q) return fail(THOR_ERR_BAD_ARG, "set_rt_surrogate: null argument");
    if (!h->rt_set) return fail(THOR_ERR_BAD_ARG, "set_rt_surrogate: set the RT parameters first (fallback and judge)");
    int nin = 0, nout = 0;
    mlp::dims(m, &nin, &nout);
    const int nv = h->nv;
    if (nin != 2 * nv + 1 || nout != nv)
        return fail(THOR_ERR_BAD_ARG, "set_rt_surrogate: the network must map 2 nv + 1 inputs to nv outputs");
    if (!q->in_mean || !q->in_std || !q->in_lo || !q->in_hi || !q->out_mean || !q->out_std || q->judge_every < 0)
        return fail(THOR_ERR_BAD_ARG, "set_rt_surrogate: missing normalisation arrays");
    if (!h->sur_mask) {
        hipError_t e = hipMalloc((void **)&h->sur_mask, (size_t)h->S * sizeof(int));
        if (e == hipSuccess) h->allocs.push_back(h->sur_mask);
        if (e == hipSuccess) e = hipMalloc((void **)&h->sur_nflag, sizeof(int));
        if (e == hipSuccess) h->allocs.push_back(h->sur_nflag);
        if (e == hipSuccess) e = hipMemset(h->sur_nflag, 0, sizeof(int));
        if (e == hipSuccess) e = hipMalloc((void **)&h->sur_list, (size_t)h->S * sizeof(int));
        if (e == hipSuccess) h->allocs.push_back(h->sur_list);
        if (e == hipSuccess) e = hipMalloc((void **)&h->sur_cnt, 2 * sizeof(int));
        if (e == hipSuccess) h->allocs.push_back(h->sur_cnt);
        if (e == hipSuccess) e = hipMemset(h->sur_cnt, 0, 2 * sizeof(int));
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);  // the fills before the handle's stream uses them
        if (e == hipSuccess) e = dalloc(*h, &h->qsur, (size_t)nv * h->S);
        h->err_nblk = (int)(((long)nv * h->S + 255) / 256);
        if (e == hipSuccess) e = dalloc(*h, &h->err_part, (size_t)3 * h->err_nblk);
        float *nrm = nullptr;
        if (e == hipSuccess) e = hipMalloc((void **)&nrm, (size_t)(4 * nin + 2 * nout) * sizeof(float));
        if (e == hipSuccess) h->allocs.push_back(nrm);
        if (e != hipSuccess) return fail(THOR_ERR_HIP, std::string("set_rt_surrogate: alloc: ") + hipGetErrorString(e));
        h->sp.in_mean = nrm;
        h->sp.in_std = nrm + nin;
        h->sp.in_lo = nrm + 2 * nin;
        h->sp.in_hi = nrm + 3 * nin;
        h->sp.out_mean = nrm + 4 * nin;
        h->sp.out_std = nrm + 4 * nin + nout;
    }
    const float *src[6] = {q->in_mean, q->in_std, q->in_lo, q->in_hi, q->out_mean, q->out_std};
    float *dst[6] = {(float *)h->sp.in_mean, (float *)h->sp.in_std, (float *)h->sp.in_lo, (float *)h->sp.in_hi,
                     (float *)h->sp.out_mean, (float *)h->sp.out_std};
    const int cnt[6] = {nin, nin, nin, nin, nout, nout};
    for (int i = 0; i < 6; i++)
        HIP_TRY(hipMemcpy(dst[i], src[i], cnt[i] * sizeof(float), hipMemcpyHostToDevice));
    h->sp.lon_sub = h->rt.lon_sub;
    if (h->sur && h->sur != m) mlp::detach(h->sur, h->stream);
    mlp::attach(m, h->stream);
    h->sur = m;
    h->judge_every = q->judge_every;
    h->sur_calls = h->sur_judged = 0;
    h->err_sq = h->err_max = h->err_cnt = 0.0;
    return THOR_OK;
}

extern "C" thor_status thor_dyn_surrogate_stats(thor_dyn *h, long *calls, long *judged, long *fallback_columns,
                                                double *rms_err, double *max_err) {
    using namespace thor;
    if (!h) return fail(THOR_ERR_BAD_ARG, "surrogate_stats: null handle");
    if (!h->sur) return fail(THOR_ERR_BAD_ARG, "surrogate_stats: no surrogate attached");
    int nflag = 0;
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(&nflag, h->sur_nflag, sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemset(h->sur_nflag, 0, sizeof(int)));
    HIP_TRY(hipStreamSynchronize(nullptr));  // cleared before the handle's stream counts again
    if (calls) *calls = h->sur_calls;
    if (judged) *judged = h->sur_judged;
    if (fallback_columns) *fallback_columns = nflag;
    if (rms_err) *rms_err = h->err_cnt > 0 ? std::sqrt(h->err_sq / h->err_cnt) : 0.0;
    if (max_err) *max_err = h->err_max;
    h->sur_calls = h->sur_judged = 0;
    h->err_sq = h->err_max = h->err_cnt = 0.0;
    return THOR_OK;
}

extern "C" thor_status thor_dyn_sync(thor_dyn *h) {
    using namespace thor;
    if (!h) return fail(THOR_ERR_BAD_ARG, "thor_dyn_sync: null handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    drain_timing(*h);
    unsigned int flags = 0;
    HIP_TRY(hipMemcpy(&flags, h->f.flags, sizeof(unsigned int), hipMemcpyDeviceToHost));
    if (flags) {
        HIP_TRY(hipMemset(h->f.flags, 0, sizeof(unsigned int)));
        HIP_TRY(hipStreamSynchronize(nullptr));  // cleared before the handle's stream can set it again
        if (flags & FLAG_NAN_T) return fail(THOR_ERR_NAN, "NaN in temperature after ProfX (profx_driver.cu:389-396)");
        return fail(THOR_ERR_NAN, "NaN in pressure in Density_Pressure_Eqs (thor_fastmodes.h:562-569)");
    }
    return THOR_OK;
}

extern "C" thor_status thor_dyn_globdiag(thor_dyn *h, double out[5]) {
    using namespace thor;
    if (!h || !out) return fail(THOR_ERR_BAD_ARG, "thor_dyn_globdiag: bad argument");
    Fields f = stage_fields(*h, 0);
    const long n = (long)h->P * h->nv;
    k_globdiag<<<h->npart, 1024, 0, h->stream>>>(h->g, f, h->partials, n);
    HIP_TRY(hipGetLastError());
    std::vector<double> part((size_t)5 * h->npart);
    HIP_TRY(hipMemcpyAsync(part.data(), h->partials, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int q = 0; q < 5; q++) {
        double s = 0.0;
        for (int b = 0; b < h->npart; b++) s += part[(size_t)q * h->npart + b];
        out[q] = s;
    }
    return THOR_OK;
}

// Totals of thor_dyn_globdiag_noalloc: thread q sums quantity q's block
// partials one after the other in block order, the host loop of
// thor_dyn_globdiag, so the totals carry the same bits.
__global__ void k_globdiag_totals(const double *partials, int nblk, double *totals) {
    const int q = threadIdx.x;
    if (q >= 5) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += partials[(size_t)q * nblk + b];
    totals[q] = s;
}

extern "C" thor_status thor_dyn_globdiag_noalloc(thor_dyn *h, double out[5]) {
    using namespace thor;
    if (!h || !out) return fail(THOR_ERR_BAD_ARG, "thor_dyn_globdiag_noalloc: bad argument");
    if (h->npart > kGdMaxBlocks)
        return fail(THOR_ERR_UNSUPPORTED, "thor_dyn_globdiag_noalloc: " + std::to_string((long)h->P * h->nv)
                                              + " cells exceed the static scratch's limit of "
                                              + std::to_string((long)kGdMaxBlocks * 2048)
                                              + " (use thor_dyn_globdiag)");
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != h->device)
        return fail(THOR_ERR_BAD_ARG, "thor_dyn_globdiag_noalloc: the current device is not the handle's");
    // every partial and total is overwritten below, so nothing is cleared
    Fields f = stage_fields(*h, 0);
    const long n = (long)h->P * h->nv;
    double *part = h->gd_scratch, *totals = h->gd_scratch + kGdTotals;
    k_globdiag<<<h->npart, 1024, 0, h->stream>>>(h->g, f, part, n);
    k_globdiag_totals<<<1, 64, 0, h->stream>>>(part, h->npart, totals);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, totals, 5 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return THOR_OK;
}

extern "C" thor_status thor_dyn_get_field(thor_dyn *h, const char *name, double *dst, long n) {
    using namespace thor;
    if (!h || !name || !dst) return fail(THOR_ERR_BAD_ARG, "get_field: null argument");
    {
        // mixed-precision operator weights, [P][ncoef] per point
        const std::string n_(name);
        const float *w = nullptr;
        int nc = 0;
        if (n_ == "wlap") { w = h->g.wlap; nc = 7; }
        else if (n_ == "wdd") { w = h->g.wdd; nc = 63; }
        else if (n_ == "udd") { w = h->g.udd; nc = 21; }
        if (nc) {
            if (!h->mixed) return fail(THOR_ERR_BAD_ARG, "get_field: operator weights exist in mixed precision only");
            if (n != (long)h->P * nc) return fail(THOR_ERR_BAD_ARG, "get_field: size mismatch");
            std::vector<float> tmp((size_t)nc * h->S);
            HIP_TRY(hipStreamSynchronize(h->stream));
            HIP_TRY(hipMemcpy(tmp.data(), w, tmp.size() * sizeof(float), hipMemcpyDeviceToHost));
            for (int p = 0; p < h->P; p++)
                for (int c = 0; c < nc; c++) dst[(size_t)p * nc + c] = tmp[(size_t)c * h->S + p];
            return THOR_OK;
        }
    }
    double *dev;
    int nlev, ncomp;
    if (!field_ptr(*h, name, dev, nlev, ncomp)) return fail(THOR_ERR_BAD_ARG, std::string("get_field: unknown ") + name);
    if (n != (long)h->P * nlev * ncomp) return fail(THOR_ERR_BAD_ARG, "get_field: size mismatch");
    HIP_TRY(hipStreamSynchronize(h->stream));
    return copy_field(*h, dev, nlev, ncomp, dst, false, false);
}

extern "C" thor_status thor_dyn_set_field(thor_dyn *h, const char *name, const double *src, long n) {
    using namespace thor;
    if (!h || !name || !src) return fail(THOR_ERR_BAD_ARG, "set_field: null argument");
    double *dev;
    int nlev, ncomp;
    if (!field_ptr(*h, name, dev, nlev, ncomp)) return fail(THOR_ERR_BAD_ARG, std::string("set_field: unknown ") + name);
    if (n != (long)h->P * nlev * ncomp) return fail(THOR_ERR_BAD_ARG, "set_field: size mismatch");
    HIP_TRY(hipStreamSynchronize(h->stream));
    return copy_field(*h, dev, nlev, ncomp, const_cast<double *>(src), true, false);
}

extern "C" thor_status thor_dyn_set_sponge(thor_dyn *h, const thor_sponge_params *q) {
    using namespace thor;
    if (!h) return fail(THOR_ERR_BAD_ARG, "thor_dyn_set_sponge: null handle");
    if (!q) {
        h->sponge_on = false;
        return THOR_OK;
    }
    if (q->nlat_bins < 2 || q->nlat_bins > 1000 || !(q->ns >= 0.0 && q->ns < 1.0) || q->Ruv < 0 || q->Rw < 0
        || (q->temp_sponge && q->RT_sponge < 0))
        return fail(THOR_ERR_BAD_ARG, "thor_dyn_set_sponge: need 2 <= nlat_bins <= 1000, 0 <= ns < 1, rates >= 0");
    if (q->raysp_calc_mode != 0 && q->raysp_calc_mode != 1 && q->raysp_calc_mode != 3)
        return fail(THOR_ERR_BAD_ARG, "thor_dyn_set_sponge: raysp_calc_mode must be 0 (imp), 1 (exp1) or 3 (exp3)");
    const int P = h->P, S = h->S, nv = h->nv, nlat = q->nlat_bins;
    // latitude bins (grid.cu:3215-3262): edges accumulated from -pi/2, index
    // within a bin in point order, points at or above the last edge in the top bin
    std::vector<double> lat((size_t)P);
    HIP_TRY(hipMemcpy(lat.data(), h->g.lonlat + S, P * sizeof(double), hipMemcpyDeviceToHost));
    const double des_lat = M_PI / nlat;
    std::vector<double> edge((size_t)nlat + 1);
    edge[0] = -M_PI / 2.0;
    for (int j = 1; j < nlat + 1; j++) edge[j] = edge[j - 1] + des_lat;
    std::vector<int> count((size_t)nlat, 0), tab((size_t)3 * S, 0);
    for (int i = 0; i < P; i++) {
        for (int j = 0; j < nlat; j++)
            if (lat[i] >= edge[j] && lat[i] < edge[j + 1]) {
                tab[i] = j;
                tab[2 * (size_t)S + i] = count[j]++;
                break;
            }
        if (lat[i] >= edge[nlat]) {
            tab[i] = nlat - 1;
            tab[2 * (size_t)S + i] = count[nlat - 1]++;
        }
    }
    int mc = 0;
    for (int i = 0; i < P; i++) {
        tab[(size_t)S + i] = count[tab[i]];
        mc = std::max(mc, tab[(size_t)S + i]);
    }
    const int max_count = (int)std::pow(2, std::ceil(std::log(mc) / std::log(2)));
    SpongePar sp{q->ns, q->Ruv, q->Rw, nlat, q->damp_uv_to_mean != 0, q->damp_w_to_mean != 0, max_count,
                 (max_count + 2047) / 2048, q->temp_sponge != 0, q->temp_sponge ? q->RT_sponge : 0.0,
                 q->raysp_calc_mode};
    // (re)allocate the tables; the zonal slots beyond each bin's count stay zero
    for (void *ptr : {(void *)h->ztab, (void *)h->ztmp, (void *)h->zout, (void *)h->vbar, (void *)h->ztrig,
                      (void *)h->zkf, (void *)h->zcnt, (void *)h->zdw}) {
        if (!ptr) continue;
        auto it = std::find(h->allocs.begin(), h->allocs.end(), ptr);
        if (it != h->allocs.end()) h->allocs.erase(it);
        (void)hipFree(ptr);
    }
    h->ztab = h->zcnt = nullptr;
    h->ztmp = h->zout = h->vbar = h->ztrig = h->zkf = h->zdw = nullptr;
    h->sponge_on = false;
    hipError_t e = hipMalloc((void **)&h->ztab, tab.size() * sizeof(int));
    if (e == hipSuccess) {
        h->allocs.push_back(h->ztab);
        e = hipMemcpy(h->ztab, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMalloc((void **)&h->zcnt, count.size() * sizeof(int));
    if (e == hipSuccess) {
        h->allocs.push_back(h->zcnt);
        e = hipMemcpy(h->zcnt, count.data(), count.size() * sizeof(int), hipMemcpyHostToDevice);
    }
    const int nc = sp.ncomp();
    if (e == hipSuccess) e = dalloc(*h, &h->ztmp, (size_t)nc * nlat * nv * max_count);
    if (e == hipSuccess) e = dalloc(*h, &h->zout, (size_t)nc * nlat * nv * sp.nb);
    if (e == hipSuccess) e = dalloc(*h, &h->vbar, (size_t)4 * nlat * nv);
    if (e == hipSuccess) e = dalloc(*h, &h->ztrig, (size_t)4 * S);
    if (e == hipSuccess) e = dalloc(*h, &h->zkf, (size_t)nv);
    if (e == hipSuccess && sp.mode != 0) e = dalloc(*h, &h->zdw, (size_t)nv * S);  // the explicit drag on W
    if (e == hipSuccess) {
        sp.trig = h->ztrig;
        sp.kf = h->zkf;
        sp.cnt = h->zcnt;
        k_sponge_tables<<<(std::max(P, nv) + 255) / 256, 256, 0, h->stream>>>(h->g, sp, h->ztrig, h->zkf);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    if (e != hipSuccess) return fail(THOR_ERR_HIP, std::string("thor_dyn_set_sponge: ") + hipGetErrorString(e));
    h->spg = sp;
    h->shrink_sponge = q->shrink_sponge;
    h->t_shrink = q->t_shrink;
    h->sponge_on = true;
    return THOR_OK;
}

extern "C" thor_status thor_dyn_set_hooks(thor_dyn *h, const thor_hooks *hooks) {
    using namespace thor;
    if (!h) return fail(THOR_ERR_BAD_ARG, "thor_dyn_set_hooks: null handle");
    h->hooks = hooks ? *hooks : thor_hooks{};
    return THOR_OK;
}

extern "C" thor_status thor_dyn_register_rk_array(thor_dyn *h, double *array_d, double *arrayk_d, double *arrayi_d,
                                                  int dimensions) {
    using namespace thor;
    if (!h || !array_d || !arrayk_d || !arrayi_d || dimensions < 1)
        return fail(THOR_ERR_BAD_ARG, "thor_dyn_register_rk_array: null array or dimensions < 1");
    thor::Dyn &d = *h;
    if (d.rka.n >= THOR_MAX_RK_ARRAYS)
        return fail(THOR_ERR_BAD_ARG, "thor_dyn_register_rk_array: at most THOR_MAX_RK_ARRAYS arrays");
    const int j = d.rka.n++;
    d.rka.a[j] = array_d;
    d.rka.ak[j] = arrayk_d;
    d.rka.ai[j] = arrayi_d;
    d.rka.count[j] = (long)d.P * d.nv * dimensions;
    return THOR_OK;
}

extern "C" thor_status thor_dyn_clear_rk_arrays(thor_dyn *h) {
    using namespace thor;
    if (!h) return fail(THOR_ERR_BAD_ARG, "thor_dyn_clear_rk_arrays: null handle");
    h->rka = thor::RKArrays{};
    return THOR_OK;
}

extern "C" thor_status thor_dyn_kernel_timing(thor_dyn *h, int enable) {
    using namespace thor;
    if (!h) return fail(THOR_ERR_BAD_ARG, "kernel_timing: null handle");
    drain_timing(*h);
    h->timing = enable != 0;
    if (enable) {
        h->stats.clear();
        h->stat_order.clear();
    }
    return THOR_OK;
}

extern "C" int thor_dyn_kernel_stat_count(thor_dyn *h) {
    if (!h) return 0;
    thor::drain_timing(*h);
    return (int)h->stat_order.size();
}

extern "C" thor_status thor_dyn_kernel_stat(thor_dyn *h, int i, const char **name, long *launches, double *total_ms,
                                            double *bytes_per_launch) {
    using namespace thor;
    if (!h || i < 0 || i >= (int)h->stat_order.size()) return fail(THOR_ERR_BAD_ARG, "kernel_stat: bad index");
    drain_timing(*h);
    const std::string &n = h->stat_order[i];
    const KernelStat &s = h->stats[n];
    if (name) *name = n.c_str();
    if (launches) *launches = s.launches;
    if (total_ms) *total_ms = s.total_ms;
    if (bytes_per_launch) *bytes_per_launch = s.launches ? s.bytes / s.launches : 0.0;
    return THOR_OK;
}

extern "C" thor_status thor_dyn_time_run(thor_dyn *h, int nsteps, int forcing, double *ms) {
    using namespace thor;
    if (!h || !ms || nsteps < 1) return fail(THOR_ERR_BAD_ARG, "time_run: bad argument");
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, h->stream));
    int rc = thor_dyn_run(h, nsteps, forcing);
    HIP_TRY(hipEventRecord(b, h->stream));
    HIP_TRY(hipEventSynchronize(b));
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = f;
    drain_timing(*h);
    return rc;
}

// ------------------------------------------------------------- tracer module
// THOR's chemistry module, its dynamical part (chemistry.cu:333-502): the
// tracer mass in three level-planar arrays advanced by the RK scheme like the
// state (tracer = y(n), tracerk = the stage's, tracers = the deviation), the
// diffusion tendency and the first diffusion pass; four hook callbacks.
struct thor_tracers {
    Dyn *d = nullptr;
    int ntr = 0;
    double *tr = nullptr, *trk = nullptr, *trs = nullptr, *dif = nullptr, *scr = nullptr;
    std::vector<void *> allocs;
};

namespace thor {
namespace {
TracerArgs tracer_args(thor_tracers *t, const thor_step_view *v) {
    TracerArgs a{};
    a.ntr = t->ntr;
    a.trk = t->trk;
    a.trs = t->trs;
    a.dif = t->dif;
    a.scr = t->scr;
    a.dif_out = t->dif;
    a.Rhok = v->Rhok;
    a.Rhos = v->Rhos;
    a.Mhk = v->Mhk;
    a.Whs = v->Whs;
    a.Whk = v->Whk;
    a.Mhs = v->Mhs;
    return a;
}
size_t tracer_bytes(const thor_tracers *t) { return (size_t)t->ntr * t->d->nv * t->d->S * sizeof(double); }

// dyn_core_loop_init: tracerk = tracer, tracers = 0 (chemistry.cu:333-343)
int tracer_init(void *user, void *stream, int, const thor_step_view *) {
    auto *t = (thor_tracers *)user;
    const hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(t->trk, t->tr, tracer_bytes(t), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(t->trs, 0, tracer_bytes(t), s);
    return e == hipSuccess ? 0 : 1;
}
// dyn_core_loop_slow_modes: the diffusion tendency of the stage (:345-427)
int tracer_slow(void *user, void *stream, int, double, const thor_step_view *v) {
    auto *t = (thor_tracers *)user;
    Dyn &d = *t->d;
    if (!d.prm.HyDiff) return 0;
    const TracerArgs a = tracer_args(t, v);
    const dim3 grid(8 * d.g.tiles_per_xcd * d.nv, t->ntr);
    k_tracer_diff<false><<<grid, kTile, 0, (hipStream_t)stream>>>(d.g, a);
    k_tracer_diff<true><<<grid, kTile, 0, (hipStream_t)stream>>>(d.g, a);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
// dyn_core_loop_fast_modes: Tracer_Eq of the substep (:429-492)
int tracer_fast(void *user, void *stream, int, double times, const thor_step_view *v) {
    auto *t = (thor_tracers *)user;
    Dyn &d = *t->d;
    const TracerArgs a = tracer_args(t, v);
    const dim3 grid(8 * d.g.tiles_per_xcd * d.nv, t->ntr);
    if (d.mixed) k_tracer_eq<true><<<grid, kTile, 0, (hipStream_t)stream>>>(d.g, a, times);
    else k_tracer_eq<false><<<grid, kTile, 0, (hipStream_t)stream>>>(d.g, a, times);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
// dyn_core_loop_end: tracer = tracerk (:495-502)
int tracer_end(void *user, void *stream, int, const thor_step_view *) {
    auto *t = (thor_tracers *)user;
    return hipMemcpyAsync(t->tr, t->trk, tracer_bytes(t), hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess
               ? 0 : 1;
}
} // namespace
} // namespace thor

extern "C" thor_status thor_tracers_create(thor_dyn *h, int ntr, thor_tracers **out) {
    using namespace thor;
    if (!h || !out) return fail(THOR_ERR_BAD_ARG, "thor_tracers_create: null argument");
    if (ntr < 1 || ntr > 6)
        return fail(THOR_ERR_BAD_ARG, "thor_tracers_create: 1 to 6 tracers (THOR's first diffusion pass keeps them in "
                                      "its six-variable diffusion array)");
    Dyn &d = *h;
    if (d.rka.n + 1 > THOR_MAX_RK_ARRAYS) return fail(THOR_ERR_BAD_ARG, "thor_tracers_create: RK-array table full");
    auto *t = new thor_tracers();
    t->d = &d;
    t->ntr = ntr;
    const size_t n = (size_t)ntr * d.nv * d.S;
    hipError_t e = hipSuccess;
    for (double **p : {&t->tr, &t->trk, &t->trs, &t->dif, &t->scr}) {
        if (e == hipSuccess) e = hipMalloc((void **)p, n * sizeof(double));
        if (e == hipSuccess) {
            t->allocs.push_back(*p);
            e = hipMemset(*p, 0, n * sizeof(double));
            if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        }
    }
    if (e != hipSuccess) {
        for (void *p : t->allocs) (void)hipFree(p);
        delete t;
        return fail(THOR_ERR_HIP, std::string("thor_tracers_create: ") + hipGetErrorString(e));
    }
    // advanced by the RK scheme (UpdateRK / UpdateRK2), the padding included
    const int j = d.rka.n++;
    d.rka.a[j] = t->trs;
    d.rka.ak[j] = t->trk;
    d.rka.ai[j] = t->tr;
    d.rka.count[j] = (long)n;
    *out = t;
    return THOR_OK;
}

extern "C" void thor_tracers_destroy(thor_tracers *t) {
    if (!t) return;
    Dyn &d = *t->d;
    (void)hipStreamSynchronize(d.stream);
    for (int j = 0; j < d.rka.n; j++)
        if (d.rka.a[j] == t->trs) {
            for (int k = j; k + 1 < d.rka.n; k++) {
                d.rka.a[k] = d.rka.a[k + 1];
                d.rka.ak[k] = d.rka.ak[k + 1];
                d.rka.ai[k] = d.rka.ai[k + 1];
                d.rka.count[k] = d.rka.count[k + 1];
            }
            d.rka.n--;
            break;
        }
    for (void *p : t->allocs) (void)hipFree(p);
    delete t;
}

static thor_status tracer_io(thor_tracers *t, double *host, long n, bool to_device) {
    using namespace thor;
    if (!t || !host) return fail(THOR_ERR_BAD_ARG, "thor_tracers: null argument");
    Dyn &d = *t->d;
    if (n != (long)d.P * d.nv * t->ntr) return fail(THOR_ERR_BAD_ARG, "thor_tracers: n must be P * nv * ntr");
    HIP_TRY(hipStreamSynchronize(d.stream));
    // THOR's layout [P][nv][ntr] <-> [ntr][nv][S]
    std::vector<double> tmp((size_t)t->ntr * d.nv * d.S, 0.0);
    if (!to_device) HIP_TRY(hipMemcpy(tmp.data(), t->tr, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int p = 0; p < d.P; p++)
        for (int lev = 0; lev < d.nv; lev++)
            for (int i = 0; i < t->ntr; i++) {
                double &dev = tmp[((size_t)i * d.nv + lev) * d.S + p];
                double &usr = host[((size_t)p * d.nv + lev) * t->ntr + i];
                if (to_device) dev = usr;
                else usr = dev;
            }
    if (to_device) HIP_TRY(hipMemcpy(t->tr, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
    return THOR_OK;
}

extern "C" thor_status thor_tracers_set(thor_tracers *t, const double *tracer, long n) {
    return tracer_io(t, const_cast<double *>(tracer), n, true);
}
extern "C" thor_status thor_tracers_get(thor_tracers *t, double *tracer, long n) { return tracer_io(t, tracer, n, false); }

extern "C" thor_status thor_tracers_hooks(thor_tracers *t, thor_hooks *hooks) {
    using namespace thor;
    if (!t || !hooks) return fail(THOR_ERR_BAD_ARG, "thor_tracers_hooks: null argument");
    hooks->user = t;
    hooks->dyn_core_loop_init = tracer_init;
    hooks->dyn_core_loop_slow_modes = tracer_slow;
    hooks->dyn_core_loop_fast_modes = tracer_fast;
    hooks->dyn_core_loop_end = tracer_end;
    return THOR_OK;
}
