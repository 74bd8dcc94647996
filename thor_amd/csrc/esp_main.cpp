// thor_esp: the simulation driver around the dynamical core, with THOR's
// command line (esp.cu:164-230, 500-580), configuration keys and output
// files. Per step it runs ProfX (the configured benchmark forcing or the
// double-gray radiative transfer, with the Rayleigh sponge) and the dynamics
// step, as the main loop of esp.cu:1379-1470 does; every n_out steps it writes
// esp_output_<ID>_<n>.h5 (the initial state is file 0 of a new run).
//
//   thor_esp <config.thr> [-o output_dir] [-i initial.h5 | -c continue.h5]
//            [-N num_steps] [-g gpu_id] [--grid-cache dir]
#include "../../include/thor_dyn.h"
#include "../../include/thor_io.h"

#include <hip/hip_runtime_api.h>

#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <regex>
#include <string>
#include <vector>

namespace {

int die(const char *what, int rc) {
    std::fprintf(stderr, "thor_esp: %s: %s\n", what, thor_last_error());
    return rc ? 1 : 0;
}

bool exists(const std::string &p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

// The planet file of an initial-conditions file (directories.cpp:314-362):
// <dir>/esp_output_<ID>_<n>.h5 -> <dir>/esp_output_planet_<ID>.h5, else
// <dir>/<stem>_planet.h5; n is the file's output index.
bool planet_of(const std::string &init, std::string &planet, int &file_number) {
    const size_t slash = init.rfind('/');
    const std::string dir = slash == std::string::npos ? "." : init.substr(0, slash);
    const std::string base = slash == std::string::npos ? init : init.substr(slash + 1);
    std::smatch m;
    static const std::regex scheme("esp_output_(.+)_([0-9]+)\\.h5");
    if (std::regex_match(base, m, scheme)) {
        planet = dir + "/esp_output_planet_" + m[1].str() + ".h5";
        file_number = std::atoi(m[2].str().c_str());
        return true;
    }
    const size_t dot = base.rfind('.');
    planet = dir + "/" + base.substr(0, dot) + "_planet.h5";
    file_number = -1;
    return false;
}

} // namespace

int main(int argc, char **argv) {
    std::string config, out_dir, initial, cont, grid_cache;
    int nsteps_arg = -1, gpu = -1;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto next = [&](const char *what) -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "thor_esp: %s needs a value\n", what);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-o" || a == "--output_dir") out_dir = next("-o");
        else if (a == "-i" || a == "--initial") initial = next("-i");
        else if (a == "-c" || a == "--continue") cont = next("-c");
        else if (a == "-N" || a == "--numsteps") nsteps_arg = std::atoi(next("-N").c_str());
        else if (a == "-g" || a == "--gpu_id") gpu = std::atoi(next("-g").c_str());
        else if (a == "--grid-cache") grid_cache = next("--grid-cache");
        else if (a == "-h" || a == "--help") {
            std::printf("usage: thor_esp <config.thr> [-o output_dir] [-i initial.h5 | -c continue.h5] "
                        "[-N num_steps] [-g gpu_id] [--grid-cache dir]\n");
            return 0;
        } else if (!a.empty() && a[0] == '-') {
            std::fprintf(stderr, "thor_esp: unknown option %s\n", a.c_str());
            return 2;
        } else config = a;
    }
    if (config.empty()) {
        std::fprintf(stderr, "thor_esp: no configuration file\n");
        return 2;
    }
    if (!initial.empty() && !cont.empty()) {
        std::fprintf(stderr, "thor_esp: --continue and --initial options set, options are exclusive\n");
        return 2;
    }
    thor_config *cfg = nullptr;
    if (thor_config_load(config.c_str(), &cfg) != THOR_OK) return die("config", 1);
    for (int i = 0; i < thor_config_unknown_count(cfg); i++)
        std::fprintf(stderr, "WARNING: config file key [%s] does not exist, skipping\n", thor_config_unknown(cfg, i));
    thor_run_config r;
    if (thor_config_run(cfg, &r) != THOR_OK) return die("config", 1);
    thor_config_free(cfg);
    if (nsteps_arg > 0) r.num_steps = nsteps_arg;
    if (out_dir.empty()) out_dir = r.results_path;
    if (gpu < 0) gpu = r.GPU_ID_N;
    ::mkdir(out_dir.c_str(), 0755);
    if (hipSetDevice(gpu) != hipSuccess) {
        std::fprintf(stderr, "thor_esp: cannot use GPU %d\n", gpu);
        return 1;
    }

    // grid and state
    thor_grid *grid = nullptr;
    if (thor_grid_build(r.params.glevel, r.vlevel, r.params.A, r.top_altitude, r.spring_dynamics, r.spring_beta,
                        grid_cache.empty() ? nullptr : grid_cache.c_str(), &grid) != THOR_OK)
        return die("grid", 1);
    thor_grid_desc gd;
    thor_grid_describe(grid, &gd);
    const int P = gd.point_num, nv = gd.nv;
    const size_t n1 = (size_t)P * nv, ni = (size_t)P * (nv + 1);
    std::vector<double> Mh(3 * n1), Wh(ni), W(n1), Rho(n1), pr(n1), Q(n1, 0.0);
    int step_idx = 1, fidx = 0;
    double t0 = 0.0;
    const std::string init = !cont.empty() ? cont : (!initial.empty() ? initial : (r.rest ? "" : r.initial));
    if (init.empty()) {
        if (thor_ic_rest(&gd, &r.params, 0, 0.0, 0.0, Mh.data(), Wh.data(), W.data(), Rho.data(), pr.data()) != THOR_OK)
            return die("initial state", 1);
    } else {
        std::string planet;
        int file_number = -1;
        const bool numbered = planet_of(init, planet, file_number);
        if (!cont.empty() && !numbered) {
            std::fprintf(stderr, "thor_esp: could not recognise the output numbering of %s\n", init.c_str());
            return 1;
        }
        if (!exists(init) || !exists(planet)) {
            std::fprintf(stderr, "thor_esp: %s or its planet file %s not found\n", init.c_str(), planet.c_str());
            return 1;
        }
        int nstep0 = 0;
        if (thor_restart_read(init.c_str(), planet.c_str(), &r, &gd, &nstep0, &t0, Mh.data(), Wh.data(), W.data(),
                              Rho.data(), pr.data()) != THOR_OK)
            return die("restart", 1);
        std::printf("thor_esp: loaded %s (step %d, t = %.1f s)\n", init.c_str(), nstep0, t0);
        if (!cont.empty()) {
            step_idx = nstep0 + 1;
            fidx = file_number + 1;
        }
    }

    thor_dyn *dyn = nullptr;
    if (thor_dyn_create(&gd, &r.params, &dyn) != THOR_OK) return die("dynamical core", 1);
    thor_state st = {0, Mh.data(), Wh.data(), W.data(), Rho.data(), pr.data()};
    if (thor_dyn_set_state(dyn, &st) != THOR_OK) return die("state", 1);
    if (r.RayleighSponge && thor_dyn_set_sponge(dyn, &r.sponge) != THOR_OK) return die("sponge", 1);
    if (r.boundary_layer && thor_dyn_set_boundary_layer(dyn, &r.bl) != THOR_OK) return die("boundary layer", 1);
    if (r.forcing == THOR_FORCING_DOUBLE_GRAY && thor_dyn_set_rt(dyn, &r.rt) != THOR_OK) return die("RT", 1);
    const int forcing = r.forcing;

    auto output = [&](int idx, int nstep, double time) -> int {
        if (thor_dyn_get_state(dyn, &st) != THOR_OK) return die("get_state", 1);
        if (thor_dyn_get_field(dyn, "profx_Qheat", Q.data(), (long)n1) != THOR_OK) return die("Qheat", 1);
        double glob[5];
        if (r.globdiag && thor_dyn_globdiag(dyn, glob) != THOR_OK) return die("globdiag", 1);
        if (thor_output_state(out_dir.c_str(), &r, idx, nstep, time, &gd, Mh.data(), Wh.data(), Rho.data(), pr.data(),
                              Q.data(), r.globdiag ? glob : nullptr) != THOR_OK)
            return die("output", 1);
        std::printf("thor_esp: step %d, t = %.1f s (%.3f days): esp_output_%s_%d.h5\n", nstep, time, time / 86400.0,
                    r.simulation_ID, idx);
        std::fflush(stdout);
        return 0;
    };
    if (cont.empty()) {
        if (thor_output_grid(out_dir.c_str(), &r, &gd) != THOR_OK) return die("grid output", 1);
        if (thor_output_planet(out_dir.c_str(), &r) != THOR_OK) return die("planet output", 1);
        if (output(0, 0, t0)) return 1;
        fidx = 1;
    }
    const auto wall0 = std::chrono::steady_clock::now();
    int done = 0;
    for (int nstep = step_idx; nstep <= r.num_steps; nstep++) {
        if (thor_dyn_profx(dyn, forcing) != THOR_OK) return die("ProfX", 1);
        if (!r.gcm_off && thor_dyn_thor(dyn) != THOR_OK) return die("Thor", 1);
        done++;
        const double time = t0 + (nstep - step_idx + 1) * r.params.timestep;
        if (nstep % r.n_out == 0) {
            if (thor_dyn_sync(dyn) != THOR_OK) return die("step", 1);
            if (output(fidx, nstep, time)) return 1;
            fidx++;
        }
    }
    if (thor_dyn_sync(dyn) != THOR_OK) return die("step", 1);
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    if (done > 0)
        std::printf("thor_esp: %d steps in %.2f s (%.1f simulated days per wall-day)\n", done, wall,
                    done * r.params.timestep / wall);
    thor_dyn_destroy(dyn);
    thor_grid_destroy(grid);
    return 0;
}
