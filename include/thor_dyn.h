/*
 * thor_dyn.h — C ABI of the MI355X-native HE-VI dynamical core.
 *
 * Drop-in boundary for THOR's dynamics step. A caller (the project's ESP
 * driver, a THOR build, or a Python/ctypes harness) owns the model state and
 * hands it over in THOR's own column-contiguous layout; the library keeps a
 * level-planar copy on the GPU and runs the step with hand-written gfx950
 * kernels. All functions return a thor_status (0 = OK); thor_last_error()
 * describes the most recent failure on the calling thread. The library never
 * calls exit().
 *
 * Reference interfaces replaced (paths relative to exoclime/THOR):
 *   thor_dyn_step          <- ESP::Thor(const SimulationSetup&, kernel_diagnostics&) on the
 *                             caller's device arrays, src/headers/esp.h:466,
 *                             src/ESP/thor_driver.cu:70-1466, called at src/esp.cu:1395-1399
 *   thor_dyn_thor          <- ESP::Thor on the handle's own copy of the state
 *                             src/headers/esp.h:466, src/ESP/thor_driver.cu:70-1466
 *   thor_dyn_profx         <- ESP::ProfX(...) with core_benchmark = HeldSuarez
 *                             src/headers/esp.h:468, src/ESP/profx_driver.cu:70-474
 *   thor_dyn_globdiag      <- ESP::globdiag(const SimulationSetup&)
 *                             src/ESP/profx_driver.cu:476-569
 *   thor_dyn_set_state /
 *   thor_dyn_get_state     <- the device arrays Mh_d, Wh_d, W_d, Rho_d, pressure_d owned by
 *                             class ESP (src/headers/esp.h:224-371) and their host copies
 *                             (ESP::copy_to_host, src/ESP/esp_initial.cu)
 *   thor_grid_build        <- Icogrid::Icogrid(...) src/ESP/grid.cu:85-302
 *   thor_ic_rest           <- ESP::initial_values(...) rest branch, src/ESP/esp_initial.cu:502-1139
 *   thor_dyn_create        <- ESP::ESP(...) + ESP::alloc_data + the Kd set-up of
 *                             ESP::initial_values (src/ESP/esp_initial.cu:68-483, 1292-1392)
 *   thor_dyn_set_sponge    <- RayleighSponge (raysp_calc_mode = imp) of ESP::ProfX,
 *                             src/ESP/profx_driver.cu:92-195, src/ESP/profx_sponge.cu:50-437
 *   thor_dyn_set_boundary_layer
 *                          <- the boundary_layer physics module, bl_type = RayleighHS
 *                             src/physics/modules/src/boundary_layer.cu:186-212,441-490
 *   thor_dyn_set_hooks     <- phy_modules_dyn_core_loop_init / _slow_modes / _fast_modes / _end
 *                             src/headers/phy_modules.h:76-86, called at thor_driver.cu:116,641,1290,1465
 *   thor_dyn_register_rk_array
 *                          <- device_RK_array_manager::register_array
 *                             src/headers/dyn/phy_modules_device.h:49-77, src/ESP/phy_modules_device.cu:53-73
 */
#ifndef THOR_DYN_H
#define THOR_DYN_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int thor_status;
enum {
    THOR_OK = 0,
    THOR_ERR_BAD_ARG = -1,
    THOR_ERR_HIP = -2,
    THOR_ERR_NAN = -3,
    THOR_ERR_NO_DEVICE = -4,
    THOR_ERR_IO = -5,
    THOR_ERR_UNSUPPORTED = -6,
    THOR_ERR_HOOK = -7
};

/* Message describing the last failure on this thread ("" if none). */
const char *thor_last_error(void);

/* ------------------------------------------------------------------ grid */

typedef struct thor_grid thor_grid;

/* Geometry in THOR's per-point layout (src/headers/grid.h:63-129):
 * point_local[P][6], maps[nr][18][18], div/grad[P][7][3], areas[P][6][3],
 * areasTr[P][6], areasT[P], nvecoa/nvecti/nvecte[P][6][3], func_r[P][3],
 * lonlat[P][2], Altitude[nv], Altitudeh[nv+1]. Poles are points P-2, P-1. */
typedef struct thor_grid_desc {
    int point_num, nv, nr, nl_region, glevel;
    int pent_ind[12];
    const int *point_local;
    const int *maps;
    const double *div, *grad, *areas, *areasTr, *areasT;
    const double *nvecoa, *nvecti, *nvecte, *func_r, *lonlat;
    const double *Altitude, *Altitudeh;
    const double *xyz, *xyzq;
} thor_grid_desc;

/* Builds the icosahedral grid (spring dynamics optional) and its operators.
 * If cache_dir is not NULL, the horizontal construction is read from / written
 * to cache_dir/thor_grid_g<level>_<spring>_<beta>.bin. */
thor_status thor_grid_build(int glevel, int nv, double radius, double top_altitude, int spring_dynamics,
                            double spring_beta, const char *cache_dir, thor_grid **out);
thor_status thor_grid_describe(const thor_grid *g, thor_grid_desc *out);
/* Iterations the spring dynamics took (0 if disabled or loaded from cache). */
int thor_grid_spring_iterations(const thor_grid *g);
void thor_grid_destroy(thor_grid *g);

/* ------------------------------------------------------------ parameters */

enum { THOR_PRECISION_FP64 = 0, THOR_PRECISION_MIXED = 1 };
/* thor_params.thermo_equation (thermo_equation_types, define.h:171) */
enum { THOR_THERMO_ENTROPY = 0, THOR_THERMO_ENERGY = 1 };

/* Scalars of SimulationSetup (src/headers/simulation_setup.h:54-119) and of
 * the ESP object that the dynamics step reads, plus the acoustic substep count
 * (the reference hard-codes 6, thor_driver.cu:91-92) and the precision mode. */
typedef struct thor_params {
    double A;        /* planet radius [m] */
    double Omega;    /* rotation rate [1/s] */
    double Gravit;   /* gravity [m/s2] */
    double P_Ref;    /* reference pressure [Pa] */
    double Rd;       /* gas constant [J/kg/K] */
    double Cp;       /* heat capacity [J/kg/K] */
    double Tmean;    /* initial isothermal temperature [K] */
    double Diffc;    /* hyperdiffusion coefficient */
    double DivDampc; /* divergence-damping coefficient */
    double timestep; /* large time step [s] */
    int glevel;
    int HyDiff, HyDiffOrder, DivDampP; /* HyDiffOrder: even, >= 4 */
    int DeepModel, NonHydro, GravHeightVar;
    int n_substeps; /* acoustic substeps per large step (reference: 6) */
    int precision;  /* THOR_PRECISION_FP64 or THOR_PRECISION_MIXED */
    /* Vertical hyperdiffusion (SimulationSetup::VertHyDiff, VertHyDiffOrder,
     * Diffc_v, simulation_setup.h:81-90): runs with HyDiff, as in THOR
     * (thor_driver.cu:590-612); Kdv6 = Diffc_v dz^order / timestep
     * (esp_initial.cu:1327-1334). VertHyDiffOrder: even, >= 4. */
    int VertHyDiff, VertHyDiffOrder;
    double Diffc_v;
    /* Diffusive sponge (SimulationSetup::DiffSponge, order_diff_sponge,
     * Dv_sponge, ns_diff_sponge; define.h:70-73), with HyDiff: strength
     * Dv_sponge sin^2(pi/2 (z/Top - ns) / (1 - ns)) above ns
     * (esp_initial.cu:1297-1321). order_diff_sponge: even, from
     * HyDiffOrder - 4 to HyDiffOrder and >= 2. */
    int DiffSponge, order_diff_sponge;
    double Dv_sponge, ns_diff_sponge;
    /* Thermodynamic equation of the fast modes (thermo_equation, esp.cu:441-443,
     * 696-703): THOR_THERMO_ENTROPY (potential temperature, THOR's default) or
     * THOR_THERMO_ENERGY (total energy; FP64 precision only, else
     * THOR_ERR_UNSUPPORTED). With the energy equation the pt and pth arrays of
     * the physics-hook view hold the flux variable h + ep + ek of the
     * density/pressure update. */
    int thermo_equation;
} thor_params;

/* Fills p with THOR's defaults for the Earth Held-Suarez case (ifile/earth_hstest.thr). */
void thor_params_default_earth(thor_params *p);

/* Rest initial condition: each column is the isothermal hydrostatic profile
 * of the reference's Newton solve (esp_initial.cu:520-587); then, if
 * temperature_noise > 0, T += U(-a, a) per cell at fixed pressure with rho
 * recomputed (precedent esp_initial.cu:1064-1071), and if wind_noise > 0 each
 * velocity component gets U(-w, w) with the radial part removed, Mh = rho*v.
 * Noise comes from std::mt19937_64(seed), u = (x >> 11) * 2^-53.
 * Outputs are host arrays in THOR's layout: Mh[P][nv][3], Wh[P][nv+1],
 * W[P][nv], Rho[P][nv], pressure[P][nv]. */
thor_status thor_ic_rest(const thor_grid_desc *g, const thor_params *p, unsigned long long seed,
                         double temperature_noise, double wind_noise, double *Mh, double *Wh, double *W,
                         double *Rho, double *pressure);

/* ------------------------------------------------------ dynamical core */

typedef struct thor_dyn thor_dyn;

/* Model state in THOR's layout (see thor_ic_rest). on_device = 0: host
 * pointers; 1: device pointers (HIP, same device as the handle). */
typedef struct thor_state {
    int on_device;
    double *Mh, *Wh, *W, *Rho, *pressure;
} thor_state;

/* Uploads the geometry, allocates state and scratch on the current HIP
 * device and computes the Kdh4/Kdhz profiles (esp_initial.cu:1292-1321). */
thor_status thor_dyn_create(const thor_grid_desc *g, const thor_params *p, thor_dyn **out);
void thor_dyn_destroy(thor_dyn *h);

thor_status thor_dyn_set_state(thor_dyn *h, const thor_state *s);
thor_status thor_dyn_get_state(thor_dyn *h, const thor_state *s);

/* One large step of the dynamical core (ESP::Thor). Asynchronous. */
thor_status thor_dyn_thor(thor_dyn *h);

/* THOR's forcing arrays, written by its ProfX and physics modules before the
 * dynamics step (ESP members profx_Qheat_d [P][nv] in W/m3, profx_dMh_d
 * [P][nv][3] and profx_dWh_d [P][nv+1], esp.h:224-371); device pointers. */
typedef struct thor_forcing {
    const double *Qheat, *dMh, *dWh;
} thor_forcing;

/* One large step on the caller's own arrays, as ESP::Thor works on the ESP
 * object's (thor_driver.cu:70-1466; the new state overwrites Mh_d, Wh_d,
 * W_d, Rho_d, pressure_d at :1458-1462): state must be device pointers in
 * THOR's layout (on_device = 1); they are read at the start and overwritten
 * with the new state. forcing gives THOR's profx_* arrays as device pointers
 * (read once, at the start); NULL keeps the handle's own forcing arrays (what
 * thor_dyn_profx or the previous call left there). dt > 0 and n_substeps > 0
 * replace the time step and substep count given at create (the Kd profiles
 * are recomputed, which waits for the handle's queued work); 0 keeps them.
 * Asynchronous: the step waits for the work queued on `stream` (a
 * hipStream_t, NULL for the default stream) before it reads the caller's
 * arrays, and work queued on `stream` afterwards waits for the new state.
 * Replaces ESP::Thor's signature esp.h:466 called from esp.cu:1395-1399. */
thor_status thor_dyn_step(thor_dyn *h, const thor_state *state, const thor_forcing *forcing, double dt,
                          int n_substeps, void *stream);

enum {
    THOR_FORCING_NONE = 0,
    THOR_FORCING_HELD_SUAREZ = 1, /* profx_held_suarez.h:51-117 */
    THOR_FORCING_NEWTONIAN = 2,   /* hot-Jupiter Newtonian relaxation, see thor_newtonian_params */
    THOR_FORCING_DOUBLE_GRAY = 3  /* double-gray two-stream RT heating, see thor_rt_params */
};
/* ESP::ProfX before the dynamics: W from Wh, T and theta from p and rho, the
 * benchmark forcing, then p = rho*Rd*T. With THOR_FORCING_DOUBLE_GRAY the
 * forcing is a heating rate (profx_Qheat) that the step integrates, as
 * THOR's physics modules do (phy_modules.h:76-86). Asynchronous. */
thor_status thor_dyn_profx(thor_dyn *h, int forcing);

/* Newtonian relaxation of temperature towards a day-night profile (no
 * reference counterpart; THOR's shallowHJ/deepHJ benchmarks,
 * profx_shallowHJ.h:51-117, have the same structure with other profiles):
 *   Teq(lon, lat, p) = T_mean + dT * w(p) * cos(lon - lon_sub) * cos(lat),
 *   w(p) = 1 for p <= p_contrast, falling linearly in ln p to 0 at p_bottom;
 *   tau_rad(p) log-spaced from tau_top at p_top to tau_bottom at p_bottom
 *   (clamped outside);  T <- (T + dt * Teq / tau) / (1 + dt / tau). */
typedef struct thor_newtonian_params {
    double T_mean, dT, lon_sub;
    double p_top, p_bottom, p_contrast;
    double tau_top, tau_bottom;
} thor_newtonian_params;
thor_status thor_dyn_set_newtonian(thor_dyn *h, const thor_newtonian_params *p);

/* Double-gray two-stream radiative transfer (rtm_dual_band, computetau,
 * radcsw, radclw, source_func_lin: src/physics/modules/inc/profx_RT.h:133-587;
 * insolation src/ESP/insolation.cu:128-207 for a synchronous rotator on a
 * circular orbit with zero obliquity: cos z = cos(lat) cos(lon - lon_sub),
 * clipped at 0). incflx is the stellar flux at the planet [W/m2], i.e. THOR's
 * incflx * r_orb^-2. No surface, latitude-independent kappa_lw. */
typedef struct thor_rt_params {
    double kappa_sw, kappa_lw; /* [m2/kg] */
    double n_sw, n_lw, f_lw;
    double diff_ang, Tint, albedo, incflx, ps0, lon_sub;
} thor_rt_params;
thor_status thor_dyn_set_rt(thor_dyn *h, const thor_rt_params *p);
/* Per-column RT diagnostics of the last double-gray ProfX: insolation
 * [W/m2], ASR and OLR [W] (reference ASR_d, OLR_d: flux x areasT). n = P. */
thor_status thor_dyn_get_rt_column(thor_dyn *h, const char *name, double *dst, long n);

/* Rayleigh sponge at the top of the model: THOR's RayleighSponge with
 * raysp_calc_mode = imp (profx_driver.cu:92-195; zonal_uv, zonal_w and
 * sponge_layer, profx_sponge.cu:50-437; latitude bins, grid.cu:3215-3262).
 * At the start of every ProfX with a forcing, in the layers with
 * Altitude / Top_altitude >= ns, the horizontal winds u, v and the vertical
 * wind w relax implicitly towards their zonal means over nlat_bins latitude
 * bands (towards zero when damp_*_to_mean = 0) at the rates Ruv, Rw [1/s]
 * times sin^2(pi/2 (n - ns) / (1 - ns)); Mh, W and the interior Wh are
 * rebuilt from the new winds. With temp_sponge (RayleighSpongeT) the
 * temperature relaxes the same way towards its zonal mean at the rate
 * RT_sponge, through the pressure (profx_sponge.cu:104-127, 388-404; THOR sums
 * slots of the mean that no point writes without clearing them, here they
 * are zero). p = NULL switches the sponge off. */
typedef struct thor_sponge_params {
    double ns, Ruv, Rw;
    int nlat_bins, damp_uv_to_mean, damp_w_to_mean;
    int temp_sponge;  /* appended: zero-filled callers keep the wind-only sponge */
    double RT_sponge; /* [1/s] */
    /* raysp_calc_mode: 0 imp (above, at the start of ProfX); 1 exp1 / 3 exp3:
     * in every RK stage before the slow modes, the drag -k (u - ubar) rho etc.
     * goes into profx_dMh / profx_dWh and the heating into profx_Qheat, with
     * the zonal means of the stage values formed in the first stage (exp1) or
     * in every stage (exp3) (thor_driver.cu:749-860). */
    int raysp_calc_mode;
    /* shrink_sponge: from step t_shrink on (steps counted from 1 per handle),
     * every sponge rate is scaled by exp(-(step - t_shrink) dt / (1000 dt))
     * (profx_driver.cu:153-160, thor_driver.cu:822-830). */
    int shrink_sponge, t_shrink;
} thor_sponge_params;
thor_status thor_dyn_set_sponge(thor_dyn *h, const thor_sponge_params *p);

/* nsteps x (ProfX + Thor) without host synchronisation. */
thor_status thor_dyn_run(thor_dyn *h, int nsteps, int forcing);

/* Waits for the handle's stream; reports THOR_ERR_NAN if the pressure NaN
 * flag was raised since the last sync (the reference exits instead,
 * thor_driver.cu:1390-1396). */
thor_status thor_dyn_sync(thor_dyn *h);

/* Global mass, total energy and angular momentum x, y, z (FP64 on the GPU,
 * per-cell values summed in the reference's tree order). Synchronous. */
thor_status thor_dyn_globdiag(thor_dyn *h, double out[5]);

/* thor_dyn_globdiag without allocation (no counterpart in the reference): for
 * callers that cannot allocate memory in the call (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The block
 * partial sums and the five totals go to a fixed device scratch of the
 * library (164 KiB per GPU, held until the process ends), the totals are
 * summed on the device in the same order as thor_dyn_globdiag, so `out` holds
 * the same bits; nothing is allocated on the device or the host. Limit:
 * point_num * nv <= 8 388 608 cells (THOR_ERR_UNSUPPORTED beyond). Call with
 * the handle's device current. All handles on one GPU share the scratch:
 * calls that overlap on one GPU (from several threads or streams) race.
 * Synchronous. */
thor_status thor_dyn_globdiag_noalloc(thor_dyn *h, double out[5]);

/* Copies a named field to / from the host in THOR's layout. Names: the state
 * ("Mh","Wh","W","Rho","pressure"), diagnostics of the last step
 * ("temperature","pt","h","hh","pth","gtil","gtilh"; "Etotal", the total
 * energy of the last acoustic substep, with THOR_THERMO_ENERGY only), forcing
 * inputs ("profx_Qheat","profx_dMh","profx_dWh"). n is the element count. */
thor_status thor_dyn_get_field(thor_dyn *h, const char *name, double *dst, long n);
thor_status thor_dyn_set_field(thor_dyn *h, const char *name, const double *src, long n);

/* Per-kernel timing with HIP events recorded on the handle's stream.
 * enable = 1 starts collection (resets counters). */
thor_status thor_dyn_kernel_timing(thor_dyn *h, int enable);
/* Number of distinct kernels timed so far; index i gives its name, launch
 * count, total milliseconds and algorithmic bytes per launch. */
int thor_dyn_kernel_stat_count(thor_dyn *h);
thor_status thor_dyn_kernel_stat(thor_dyn *h, int i, const char **name, long *launches, double *total_ms,
                                 double *bytes_per_launch);

/* Milliseconds of device time for n steps of thor_dyn_run, measured with HIP
 * events around the whole sequence on the handle's stream. */
thor_status thor_dyn_time_run(thor_dyn *h, int nsteps, int forcing, double *ms);

/* -------------------------------------------------------- physics modules */

/* What a physics module sees of the step when a hook runs (the reference
 * hands its modules `const ESP&`). Device pointers in the library's
 * level-planar layout: element (component c, level k, point p) of a field
 * with L levels is at [(c * L + k) * plane + p]; L = nv for layer fields,
 * nv + 1 for interface fields (Whk, Whs, pth). Stage values ("k") are those
 * of the current RK stage, deviations ("s") those of the current substep.
 * Mhs is FP32 when mhs_fp32 = 1 (THOR_PRECISION_MIXED), else FP64. */
typedef struct thor_step_view {
    long plane;
    int point_num, nv, mhs_fp32;
    const double *Mhk, *Whk, *Wk, *Rhok, *pressurek;
    const void *Mhs;
    const double *Whs, *Rhos, *pressures;
    const double *temperature, *h, *pt, *pth;
} thor_step_view;

/* Physics-module hook of the step (phy_modules.h:76-86). Each callback runs
 * on the calling host thread, in stream order, where the reference calls it:
 *   dyn_core_loop_init        before the first RK stage (thor_driver.cu:116);
 *   dyn_core_loop_slow_modes  in every RK stage after the hyperdiffusion and
 *                             before the slow tendencies (:641);
 *   dyn_core_loop_fast_modes  in every acoustic substep after the vertical
 *                             solve and before the density/pressure update (:1290);
 *   dyn_core_loop_end         after the new state is in place (:1465).
 * stream is the handle's hipStream_t (a module enqueues its kernels on it),
 * nstep the 1-based number of the Thor step this handle is taking, times the
 * RK stage's step (dt/3, dt/ns, dt/ns) as in the reference. A non-zero return
 * stops the step with THOR_ERR_HOOK (the state is then partly updated). Any
 * callback may be NULL; hooks = NULL removes them. */
typedef struct thor_hooks {
    void *user;
    int (*dyn_core_loop_init)(void *user, void *stream, int nstep, const thor_step_view *v);
    int (*dyn_core_loop_slow_modes)(void *user, void *stream, int nstep, double times, const thor_step_view *v);
    int (*dyn_core_loop_fast_modes)(void *user, void *stream, int nstep, double times, const thor_step_view *v);
    int (*dyn_core_loop_end)(void *user, void *stream, int nstep, const thor_step_view *v);
} thor_hooks;
thor_status thor_dyn_set_hooks(thor_dyn *h, const thor_hooks *hooks);

/* Registers a module array that the RK scheme advances like the state
 * (device_RK_array_manager::register_array; advanced inside UpdateRK and
 * UpdateRK2, thor_auxiliary.h:376-387 and :482-494): at the start of RK
 * stages 2 and 3, array = arrayi - arrayk; at the end of every stage,
 * arrayk += array. Caller-owned device arrays in THOR's layout
 * [(p * nv + lev) * dimensions + i]; the module itself clears array and sets
 * arrayk at dyn_core_loop_init and copies arrayk back at dyn_core_loop_end,
 * as THOR's modules do. At most THOR_MAX_RK_ARRAYS arrays
 * (NUM_PHY_MODULES_DYN_CORE_ARRAYS). */
#define THOR_MAX_RK_ARRAYS 10
thor_status thor_dyn_register_rk_array(thor_dyn *h, double *array_d, double *arrayk_d, double *arrayi_d,
                                       int dimensions);
thor_status thor_dyn_clear_rk_arrays(thor_dyn *h);

/* ------------------------------------------------ boundary-layer module */

/* THOR's boundary_layer module with bl_type = RayleighHS
 * (src/physics/modules/src/boundary_layer.cu:186-212, kernel rayleighHS
 * :441-490; keys surf_drag, bl_sigma): in every ProfX, after the forcing and
 * before the new pressure (phy_modules_phy_loop, src/ESP/profx_driver.cu:403-416),
 *   Mh <- Mh / (1 + kv dt),  kv = surf_drag * max(0, (p/ps - bl_sigma) / (1 - bl_sigma)),
 * with ps extrapolated from the two lowest layers (GravHeightVar as in the
 * step). NULL switches it off. */
typedef struct thor_bl_params {
    double surf_drag; /* [1/s], THOR's default 1/86400 */
    double bl_sigma;  /* THOR's default 0.7 */
} thor_bl_params;
thor_status thor_dyn_set_boundary_layer(thor_dyn *h, const thor_bl_params *p);

/* ------------------------------------------------- tracer module (chemistry) */

/* Passive tracers advected with the flow: the dynamical part of THOR's
 * chemistry module (src/physics/modules/src/chemistry.cu:333-502, kernels
 * src/physics/modules/inc/thor_chemistry.h). A consumer of the hooks above:
 * thor_tracers_create registers its tracer arrays with the RK scheme of h
 * (device_RK_array_manager::register_array, as chemistry.cu:115 does), and
 * thor_tracers_hooks fills a thor_hooks with its four callbacks, which the
 * caller installs with thor_dyn_set_hooks:
 *   init: tracerk = tracer, tracers = 0;
 *   slow modes (with HyDiff): the stage's fourth-order diffusion of the mixing
 *         ratio tracerk / Rhok (Tracer_Eq_Diffusion, two passes);
 *   fast modes: Tracer_Eq, the flux-form advection of the tracer mass by the
 *         stage momentum plus its deviation, into tracers;
 *   end: tracer = tracerk.
 * The tracer is THOR's tracer_d, the mass density rho q [P][nv][ntr] in its
 * layout; 1 <= ntr <= 6. The module must be destroyed before h. */
typedef struct thor_tracers thor_tracers;
thor_status thor_tracers_create(thor_dyn *h, int ntr, thor_tracers **out);
void thor_tracers_destroy(thor_tracers *t);
thor_status thor_tracers_set(thor_tracers *t, const double *tracer, long n);
thor_status thor_tracers_get(thor_tracers *t, double *tracer, long n);
thor_status thor_tracers_hooks(thor_tracers *t, thor_hooks *hooks);

#ifdef __cplusplus
}
#endif
#endif /* THOR_DYN_H */
