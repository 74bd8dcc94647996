/*
 * thor_io.h — THOR's file formats around the dynamical core (libthor_io.so).
 *
 * Kept out of libthor_amd.so so the core library needs no HDF5. Used by the
 * CLI driver (thor_esp) and by tests.
 *
 * Reference interfaces replaced (paths relative to exoclime/THOR):
 *   thor_config_*      <- config_file::parse_config and the keys registered in
 *                         main (src/utils/config_file.cpp:71-131, src/esp.cu:252-475)
 *   thor_h5_*          <- storage::append_table / append_value / read_table
 *                         (src/headers/storage.h, src/utils/storage.cpp): one
 *                         1-D dataset per name, IEEE F64 LE or STD I32 LE, with
 *                         string attributes "Variable" (description) and "units"
 *   thor_output_*      <- ESP::output (src/ESP/esp_output.cu:136-486): the grid,
 *                         planet and per-output files
 *   thor_restart_read  <- the restart branch of ESP::initial_values
 *                         (src/ESP/esp_initial.cu:1140-1283)
 */
#ifndef THOR_IO_H
#define THOR_IO_H

#include "thor_dyn.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ .thr files */

typedef struct thor_config thor_config;

/* Parses a THOR configuration file: "key = value" lines, '#' comments,
 * blank lines. A malformed line is an error (THOR_ERR_IO, with the line
 * number in thor_last_error); unknown keys are kept and reported by
 * thor_config_unknown. */
thor_status thor_config_load(const char *path, thor_config **out);
void thor_config_free(thor_config *c);
/* The raw value of key (trimmed), or NULL if the file does not set it. */
const char *thor_config_get(const thor_config *c, const char *key);
/* Number of keys that are not THOR keys known to this build; i-th name. */
int thor_config_unknown_count(const thor_config *c);
const char *thor_config_unknown(const thor_config *c, int i);

/* Run settings of a configuration, with THOR's defaults (src/headers/define.h,
 * src/ESP/simulation_setup.cu:56-110) for keys the file does not set. */
typedef struct thor_run_config {
    thor_params params;       /* physics and numerics of the step */
    double top_altitude;      /* Top_altitude [m] */
    int vlevel, spring_dynamics;
    double spring_beta;
    int num_steps, n_out, globdiag, rest, gcm_off;
    int forcing;              /* THOR_FORCING_* from core_benchmark / radiative_transfer */
    char simulation_ID[160];
    char results_path[512];
    char initial[512];        /* initial conditions file when rest = false */
    int GPU_ID_N;
    /* Rayleigh sponge (RayleighSponge, RayleighSpongeT, ...) */
    int RayleighSponge;
    thor_sponge_params sponge;
    /* double-gray radiative transfer (radiative_transfer = true) */
    thor_rt_params rt;
    double Tstar, radius_star, planet_star_dist; /* [K], [R_sun], [au] */
    /* boundary-layer module (boundary_layer = true, bl_type = RayleighHS) */
    int boundary_layer;
    thor_bl_params bl;
} thor_run_config;

/* Fills r from c; keys with values this build does not run (another
 * thermo_equation, a benchmark other than HeldSuarez / NoBenchmark,
 * vert_refined, ...) give THOR_ERR_UNSUPPORTED naming the key. */
thor_status thor_config_run(const thor_config *c, thor_run_config *r);

/* ---------------------------------------------------------------- HDF5 */

/* Appends a 1-D FP64 dataset `name` (leading "/" optional) with the
 * "Variable" and "units" attributes to the file, created (truncated) when
 * create != 0. */
thor_status thor_h5_write(const char *file, int create, const char *name, const double *data, long n,
                          const char *unit, const char *description);
/* Number of elements of a dataset (-1 if absent). */
long thor_h5_size(const char *file, const char *name);
/* Reads a dataset of n elements as FP64 (integer datasets are converted). */
thor_status thor_h5_read(const char *file, const char *name, double *data, long n);
/* The root group's dataset names, newline-separated, into buf (len bytes). */
int thor_h5_names(const char *file, char *buf, int len);
/* 1 for an integer dataset, 0 for floating point, -1 if absent. */
int thor_h5_is_int(const char *file, const char *name);
/* Reads a string attribute of a dataset into buf (len bytes, NUL-terminated). */
thor_status thor_h5_attr(const char *file, const char *name, const char *attr, char *buf, int len);

/* ----------------------------------------------------- THOR output files */

/* esp_output_grid_<ID>.h5 and esp_output_planet_<ID>.h5 in dir (written at
 * step 0 in THOR). */
thor_status thor_output_grid(const char *dir, const thor_run_config *r, const thor_grid_desc *g);
thor_status thor_output_planet(const char *dir, const thor_run_config *r);
/* esp_output_<ID>_<fidx>.h5: /nstep, /simulation_time, /Rho, /Pressure,
 * /Mh, /Wh, /Qheat, /Rd, /Cp, and with globdiag the global sums
 * (/GlobalMass, /GlobalE, /GlobalAMx, /GlobalAMy, /GlobalAMz). Host arrays in
 * THOR's layout; glob may be NULL. */
thor_status thor_output_state(const char *dir, const thor_run_config *r, int fidx, int nstep, double simulation_time,
                              const thor_grid_desc *g, const double *Mh, const double *Wh, const double *Rho,
                              const double *pressure, const double *Qheat, const double glob[5]);
/* Restart: checks the planet file against r (A, Top_altitude, glevel,
 * vlevel) and reads /nstep, /simulation_time, /Rho, /Pressure, /Mh, /Wh
 * (and /Rd, /Cp, which must be uniform and equal to the configuration's);
 * W is interpolated from Wh at the layer centres as THOR does. */
thor_status thor_restart_read(const char *state_file, const char *planet_file, const thor_run_config *r,
                              const thor_grid_desc *g, int *nstep, double *simulation_time, double *Mh, double *Wh,
                              double *W, double *Rho, double *pressure);

#ifdef __cplusplus
}
#endif
#endif /* THOR_IO_H */
