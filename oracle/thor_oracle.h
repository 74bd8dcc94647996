// ORACLE — test infrastructure only. Serial FP64 host restatement of THOR's
// HE-VI dynamical core step (src/ESP/thor_driver.cu:70-1466 and the kernels of
// src/headers/dyn/*.h), the Held-Suarez ProfX step (src/ESP/profx_driver.cu,
// src/headers/phy/profx_held_suarez.h, profx_auxiliary.h) and the global
// diagnostics (src/ESP/profx_globdiag.cu, src/headers/reduction_add.h).
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
// this library; it is the checker, never the product path.
//
// Every array uses the reference's column-contiguous layout:
//   scalars [P][nv], interfaces [P][nv+1], vectors [P][nv][3],
//   div/grad [P][7][3], areas/nvec* [P][6][3], areasTr [P][6], maps [nr][18][18].
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_grid {
    int point_num, nv, nr, nl_region;
    const int *point_local, *maps;
    const double *div, *grad, *areas, *areasTr, *areasT;
    const double *nvecoa, *nvecti, *nvecte, *func_r, *lonlat;
    const double *Altitude, *Altitudeh;
} oracle_grid;

typedef struct oracle_params {
    double A, Omega, Gravit, P_Ref, Rd, Cp;
    double Diffc, DivDampc, timestep;
    int glevel, HyDiffOrder, n_substeps;
    int DeepModel, NonHydro, HyDiff, DivDampP, GravHeightVar;
    int VertHyDiff, VertHyDiffOrder;
    double Diffc_v;
    int DiffSponge, order_diff_sponge;
    double Dv_sponge, ns_diff_sponge;
    int energy_equation;  // thermo_equation = energy
} oracle_params;

typedef struct oracle oracle;

oracle *oracle_create(const oracle_grid *g, const oracle_params *p);
void oracle_destroy(oracle *o);
// name is one of the reference array names ("Mh", "Wh", "W", "Rho", "pressure",
// "temperature", "profx_Qheat", "profx_dMh", "profx_dWh", "Mhk", "Adv", ...);
// returns the number of doubles or -1.
long oracle_array_size(oracle *o, const char *name);
int oracle_set(oracle *o, const char *name, const double *src, long n);
int oracle_get(oracle *o, const char *name, double *dst, long n);
// ESP::Thor: one large step of the dynamical core.
void oracle_thor(oracle *o);
// ESP::ProfX with the Held-Suarez benchmark (no physics modules).
void oracle_profx_held_suarez(oracle *o);
// One restated kernel (with its pole kernel, where THOR has one) on the
// current arrays, for kernel-by-kernel comparison with THOR's own kernels:
//   "advec_cori1"  Compute_Advec_Cori1 + _Poles   (thor_adv_cor.h:51-362)
//   "advec_cori2"  Compute_Advec_Cori2            (thor_adv_cor.h:366-536)
//   "diffusion"    Diffusion_Op + _Poles, a0 = stepnum, a1 = laststep (thor_diff.h:54-880)
//   "vertical_diff" vertical_diff                 (thor_diff.h:1198-1426)
//   "sponge_explicit" explicit Rayleigh sponge of RK stage a0 (thor_driver.cu:749-860)
//   "correct_horizontal"                          (thor_diff.h:882-915)
//   "divm"         DivM_Op + _Poles on Mhk/Whk, a0 = laststep (thor_div.h:49-535)
//   "slow_modes"   Compute_Slow_Modes + _Poles    (thor_slowmodes.h:48-683)
//   "held_suarez"  held_suarez                    (profx_held_suarez.h:51-117)
//   "vertical_eq"  Vertical_Eq, a0 = substep length in ms, a1 != 0 keeps each
//                  column's system as "tri_a" .. "tri_d" [P][nv+1]
//                  (a x_{l-1} + b x_l + c x_{l+1} = d at interfaces 1..nv-1)
//                                                 (thor_vertical_int.h:52-563)
//   "temperature_h_pt_geff"                       (thor_auxiliary.h:55-309)
//   "update_rk" / "update_rk2"                    (thor_auxiliary.h:311-529)
//   "momentum_eq"  Momentum_Eq + _Poles, a0 = substep length in ms (thor_fastmodes.h:54-253)
//   "prepare_implicit_vertical" + _Poles          (thor_vertical_int.h:566-771)
//   "density_pressure" Density_Pressure_Eqs + _Poles, a0 = substep length in ms
//                                                 (thor_fastmodes.h:255-847)
//   "tracer_diffusion" Tracer_Eq_Diffusion + _Poles, a0 = laststep; first pass
//                  into "trdiff" [P][nv][ntr] (thor_chemistry.h:364-760)
//   "tracer_eq"    Tracer_Eq + _Poles, a0 = substep length in ms (thor_chemistry.h:31-361)
// Returns 0, or -1 for an unknown name. "Kdh4"/"Kdhz" are readable arrays.
int oracle_kernel(oracle *o, const char *name, int a0, int a1);
// ProfX + Thor, n times.
void oracle_steps(oracle *o, int n);
// Host threads of the block and column loops (process-wide; n < 1 leaves the
// OpenMP default, which honours OMP_NUM_THREADS). Results are the same for
// every thread count; 1 is the serial loop.
void oracle_set_threads(int n);
int oracle_get_threads(void);
// Chemistry tracers (thor_chemistry.h, driven through the physics hooks as in
// chemistry.cu:333-502): ntr tracer masses [P][nv][ntr] in the arrays
// "tracer", "tracerk", "tracers", "difftr" (0 switches them off).
void oracle_set_tracers(oracle *o, int ntr);
// Forcings other than Held-Suarez (forcing codes as THOR_FORCING_* of the
// library): 2 Newtonian relaxation with q = {T_mean, dT, lon_sub, p_top,
// p_bottom, p_contrast, tau_top, tau_bottom}; 3 double-gray RT with q =
// {kappa_sw, kappa_lw, n_sw, n_lw, f_lw, diff_ang, Tint, albedo, incflx, ps0,
// lon_sub} (per-column results in the arrays "insol", "ASR", "OLR").
void oracle_set_newtonian(oracle *o, const double q[8]);
void oracle_set_rt(oracle *o, const double q[11]);
// ProfX with forcing 0..3; returns -1 for an unknown forcing.
/* Rayleigh sponge in ProfX (implicit mode): q = {ns, Ruv, Rw, nlat_bins,
 * damp_uv_to_mean, damp_w_to_mean}; NULL switches it off. */
void oracle_set_sponge(oracle *o, const double q[6]);
/* The sponge's latitude-bin table [P][3] (bin, points in bin, index in bin). */
int oracle_zonal_table(oracle *o, int *dst, long n, int *max_count);
/* Boundary-layer module (RayleighHS) after the ProfX forcing: q = {surf_drag,
 * bl_sigma}; NULL switches it off. */
void oracle_set_boundary_layer(oracle *o, const double q[2]);
int oracle_profx(oracle *o, int forcing);
// n x (ProfX with forcing + Thor).
int oracle_run(oracle *o, int n, int forcing);
// Global mass, total energy, angular momentum (x, y, z): per-cell values summed
// with the reference's 1024-thread tree reduction order.
void oracle_globdiag(oracle *o, double out[5]);
// Tree-reduction sum of an arbitrary array in the reference order (reduction_add.h:59-111).
double oracle_tree_sum(const double *d, long n);
// Elementary functions used by the restatement (fn 0 pow, 1 log, 2 sin, 3 cos, 4 exp).
double oracle_math(int fn, double x, double y);

#ifdef __cplusplus
}
#endif
