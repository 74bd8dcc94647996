// THOR's file formats around the dynamical core: the .thr configuration
// (config_file.cpp:71-131 and the keys of esp.cu:252-475), the HDF5 storage
// layout of storage.h (one 1-D dataset per name with "Variable" and "units"
// string attributes), the output files of ESP::output (esp_output.cu:136-486)
// and the restart branch of ESP::initial_values (esp_initial.cu:1140-1283).
// Written against the HDF5 C API; libthor_amd.so itself does not use HDF5.
#include "../../include/thor_io.h"
#include "common.h"

#include <hdf5.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <string>
#include <vector>

struct thor_config {
    std::map<std::string, std::string> kv;
    std::vector<std::string> unknown;
};

namespace thor {
namespace {

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

// Keys of THOR's configuration files (esp.cu:252-475, the radiative-transfer
// module's, define.h) and the two this build adds (n_substeps, precision).
const std::set<std::string> &known_keys() {
    static const std::set<std::string> k = {
        "config_version", "num_steps", "timestep", "n_out", "custom_log_n_out", "log_n_out", "results_path",
        "globdiag", "custom_global_n_out", "global_n_out", "output_mean", "simulation_ID", "radius",
        "rotation_rate", "gravitation", "GravHeightVar", "Rd", "Cp", "Tmean", "P_ref", "Top_altitude", "glevel",
        "vlevel", "spring_dynamics", "spring_beta", "vert_refined", "lowest_layer_thickness", "transition_altitude",
        "NonHydro", "DeepModel", "rest", "initial", "init_PT_profile", "core_benchmark", "gcm_off", "conv_adj",
        "conv_adj_iter", "soft_adjustment", "conv_adj_type", "thermo_equation", "ultrahot_heating",
        "ultrahot_thermo", "HyDiff", "HyDiffOrder", "DivDampP", "VertHyDiff", "VertHyDiffOrder", "Diffc",
        "DivDampc", "Diffc_v", "increased_damping_for_n_steps", "RayleighSponge", "RayleighSpongeT",
        "damp_uv_to_mean", "damp_w_to_mean", "nlat_bins", "ns_ray_sponge", "Ruv_sponge", "Rw_sponge", "RT_sponge",
        "raysp_calc_mode", "shrink_sponge", "t_shrink", "DiffSponge", "Dv_sponge", "ns_diff_sponge",
        "order_diff_sponge", "radiative_transfer", "rt_type", "Tstar", "planet_star_dist", "radius_star", "albedo",
        "diff_ang", "n_lw", "n_sw", "kappa_lw", "kappa_sw", "f_lw", "Tint", "latf_lw", "kappa_lw_pole", "surface",
        "Csurf", "rt1Dmode", "sync_rot", "boundary_layer", "bl_type", "surf_drag", "bl_sigma", "GPU_ID_N", "output_diffusion",
        "out_interm_momentum", "single_column", "MetStar", "bv_freq", "a_dense", "b_dense", "tausw", "taulw",
        "dgrt_spinup_start", "dgrt_spinup_stop", "dgrt_spindown_stop", "n_substeps", "precision"};
    return k;
}

bool parse_bool(const std::string &v, bool &out) {
    if (v == "true" || v == "1") out = true;
    else if (v == "false" || v == "0") out = false;
    else return false;
    return true;
}

// ---------------------------------------------------------------- HDF5
struct H5Quiet {
    H5Quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};

struct Hid {
    hid_t id;
    herr_t (*close)(hid_t);
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid() {
        if (id >= 0) close(id);
    }
    operator hid_t() const { return id; }
};

std::string dname(const char *name) { return name[0] == '/' ? std::string(name) : "/" + std::string(name); }

// One string attribute of storage::append_table: a fixed-length C string of
// exactly the text's length.
bool put_attr(hid_t ds, const char *attr, const std::string &text) {
    Hid st(H5Tcopy(H5T_C_S1), H5Tclose);
    if (st < 0 || H5Tset_size(st, text.empty() ? 1 : text.size()) < 0) return false;
    Hid as(H5Screate(H5S_SCALAR), H5Sclose);
    Hid a(H5Acreate2(ds, attr, st, as, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return a >= 0 && H5Awrite(a, st, text.empty() ? "" : text.c_str()) >= 0;
}

// Appends a dataset of FP64 (file type IEEE_F64LE) or int (STD_I32LE) values.
template <class T>
int write_table(hid_t f, const char *name, const T *data, long n, const std::string &unit, const std::string &desc) {
    const bool is_int = std::is_same<T, int>::value;
    const hsize_t dims[1] = {(hsize_t)n};
    Hid sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
    Hid ds(H5Dcreate2(f, dname(name).c_str(), is_int ? H5T_STD_I32LE : H5T_IEEE_F64LE, sp, H5P_DEFAULT,
                      H5P_DEFAULT, H5P_DEFAULT),
           H5Dclose);
    if (ds < 0) return fail(THOR_ERR_IO, std::string("HDF5: cannot create dataset ") + name);
    if (H5Dwrite(ds, is_int ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        return fail(THOR_ERR_IO, std::string("HDF5: cannot write dataset ") + name);
    if (!put_attr(ds, "Variable", desc) || !put_attr(ds, "units", unit))
        return fail(THOR_ERR_IO, std::string("HDF5: cannot write the attributes of ") + name);
    return THOR_OK;
}

hid_t open_file(const char *file, bool create) {
    static H5Quiet quiet;
    (void)quiet;
    return create ? H5Fcreate(file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT) : H5Fopen(file, H5F_ACC_RDWR, H5P_DEFAULT);
}

long table_size(hid_t f, const char *name) {
    if (H5Lexists(f, dname(name).c_str(), H5P_DEFAULT) <= 0) return -1;
    Hid ds(H5Dopen2(f, dname(name).c_str(), H5P_DEFAULT), H5Dclose);
    if (ds < 0) return -1;
    Hid sp(H5Dget_space(ds), H5Sclose);
    return sp < 0 ? -1 : (long)H5Sget_simple_extent_npoints(sp);
}

int read_table(hid_t f, const char *name, double *data, long n) {
    const long m = table_size(f, name);
    if (m < 0) return fail(THOR_ERR_IO, std::string("HDF5: no dataset ") + name);
    if (m != n)
        return fail(THOR_ERR_IO, std::string("HDF5: dataset ") + name + " has " + std::to_string(m) +
                                     " elements, expected " + std::to_string(n));
    Hid ds(H5Dopen2(f, dname(name).c_str(), H5P_DEFAULT), H5Dclose);
    if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        return fail(THOR_ERR_IO, std::string("HDF5: cannot read ") + name);
    return THOR_OK;
}

// Output file writer: datasets appended one after the other.
struct Out {
    Hid f;
    int rc = THOR_OK;
    explicit Out(const std::string &path) : f(open_file(path.c_str(), true), H5Fclose) {
        if (f < 0) rc = fail(THOR_ERR_IO, "HDF5: cannot create " + path);
    }
    void table(const char *name, const double *d, long n, const char *unit, const char *desc) {
        if (rc == THOR_OK) rc = write_table(f, name, d, n, unit, desc);
    }
    void itable(const char *name, const int *d, long n, const char *unit, const char *desc) {
        if (rc == THOR_OK) rc = write_table(f, name, d, n, unit, desc);
    }
    void value(const char *name, double v, const char *unit, const char *desc) { table(name, &v, 1, unit, desc); }
    void ivalue(const char *name, int v, const char *unit, const char *desc) { itable(name, &v, 1, unit, desc); }
};

std::string path_of(const char *dir, const std::string &file) {
    std::string d(dir ? dir : ".");
    if (!d.empty() && d.back() != '/') d += '/';
    return d + file;
}

} // namespace
} // namespace thor

using namespace thor;

extern "C" thor_status thor_config_load(const char *path, thor_config **out) {
    if (!path || !out) return fail(THOR_ERR_BAD_ARG, "thor_config_load: null argument");
    std::ifstream in(path);
    if (!in) return fail(THOR_ERR_IO, std::string("thor_config_load: cannot open ") + path);
    auto *c = new thor_config();
    int lineno = 0;
    for (std::string line; std::getline(in, line);) {
        lineno++;
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        const size_t eq = t.find('=');
        std::string key = eq == std::string::npos ? "" : trim(t.substr(0, eq));
        std::string val = eq == std::string::npos ? "" : trim(t.substr(eq + 1));
        const size_t hash = val.find('#');
        if (hash != std::string::npos) val = trim(val.substr(0, hash));
        bool ok = !key.empty() && !val.empty();
        for (char ch : key) ok = ok && (std::isalnum((unsigned char)ch) || ch == '_');
        if (!ok) {
            delete c;
            return fail(THOR_ERR_IO, std::string("thor_config_load: ") + path + " line " + std::to_string(lineno) +
                                         ": not a key = value pair: [" + line + "]");
        }
        if (!known_keys().count(key)) c->unknown.push_back(key);
        c->kv[key] = val;
    }
    *out = c;
    return THOR_OK;
}

extern "C" void thor_config_free(thor_config *c) { delete c; }

extern "C" const char *thor_config_get(const thor_config *c, const char *key) {
    if (!c || !key) return nullptr;
    auto it = c->kv.find(key);
    return it == c->kv.end() ? nullptr : it->second.c_str();
}

extern "C" int thor_config_unknown_count(const thor_config *c) { return c ? (int)c->unknown.size() : 0; }
extern "C" const char *thor_config_unknown(const thor_config *c, int i) {
    return c && i >= 0 && i < (int)c->unknown.size() ? c->unknown[i].c_str() : nullptr;
}

extern "C" thor_status thor_config_run(const thor_config *c, thor_run_config *r) {
    if (!c || !r) return fail(THOR_ERR_BAD_ARG, "thor_config_run: null argument");
    std::memset(r, 0, sizeof(*r));
    thor_params &p = r->params;
    // SimulationSetup defaults (simulation_setup.cu:56-110) and define.h
    p.A = 72427000.0;
    p.Omega = 9.09E-5;
    p.Gravit = 47.0;
    p.Rd = 3714;
    p.Cp = 13000;
    p.Tmean = 1400;
    p.P_Ref = 10000000.0;
    r->top_altitude = 1235376.0;
    p.Diffc = 0.009973;
    p.DivDampc = p.Diffc;
    p.Diffc_v = 0.0;
    p.timestep = 1800;
    p.glevel = 4;
    p.HyDiff = 1;
    p.HyDiffOrder = 4;
    p.DivDampP = 1;
    p.DeepModel = 1;
    p.NonHydro = 1;
    p.GravHeightVar = 0;
    p.VertHyDiff = 0;
    p.VertHyDiffOrder = 6;
    p.DiffSponge = 0;
    p.order_diff_sponge = 2;
    p.Dv_sponge = 0.01;
    p.ns_diff_sponge = 0.75;
    p.n_substeps = 6;  // thor_driver.cu:91-92
    p.precision = THOR_PRECISION_FP64;
    p.thermo_equation = THOR_THERMO_ENTROPY;
    r->vlevel = 32;
    r->spring_dynamics = 1;
    r->spring_beta = 1.15;
    r->num_steps = 48000;
    r->n_out = 1000;
    r->globdiag = 0;
    r->rest = 1;
    r->forcing = THOR_FORCING_HELD_SUAREZ;
    std::snprintf(r->simulation_ID, sizeof r->simulation_ID, "%s", "Earth");
    std::snprintf(r->results_path, sizeof r->results_path, "%s", "results");
    std::snprintf(r->initial, sizeof r->initial, "%s", "ifile/esp_initial.h5");
    thor_sponge_params &s = r->sponge;
    s.ns = 0.75;
    s.Ruv = 1e-4;
    s.Rw = 1e-4;
    s.nlat_bins = 20;
    s.damp_uv_to_mean = 1;
    s.damp_w_to_mean = 1;
    s.RT_sponge = 1e-4;
    s.t_shrink = 144000;
    thor_rt_params &rt = r->rt;
    rt.kappa_lw = 0.002;
    rt.kappa_sw = 0.001;
    rt.Tint = 100.0;
    rt.f_lw = 0.5;
    rt.n_sw = 1.0;
    rt.n_lw = 2.0;
    rt.diff_ang = 0.5;
    rt.albedo = 0.0;
    r->Tstar = 5778.0;
    r->radius_star = 1.0;
    r->planet_star_dist = 1.0;

    std::string bad;
    auto get = [&](const char *k) -> const char * { return thor_config_get(c, k); };
    auto D = [&](const char *k, double &v) {
        if (const char *s = get(k)) {
            char *end = nullptr;
            const double x = std::strtod(s, &end);
            if (end == s || *end) bad = k;
            else v = x;
        }
    };
    auto I = [&](const char *k, int &v) {
        if (const char *s = get(k)) {
            char *end = nullptr;
            const long x = std::strtol(s, &end, 10);
            if (end == s || *end) bad = k;
            else v = (int)x;
        }
    };
    auto B = [&](const char *k, int &v) {
        if (const char *s = get(k)) {
            bool b;
            if (!parse_bool(s, b)) bad = k;
            else v = b ? 1 : 0;
        }
    };
    auto S = [&](const char *k, char *dst, size_t n) {
        if (const char *s = get(k)) std::snprintf(dst, n, "%s", s);
    };
    I("num_steps", r->num_steps);
    D("timestep", p.timestep);
    I("n_out", r->n_out);
    S("results_path", r->results_path, sizeof r->results_path);
    B("globdiag", r->globdiag);
    S("simulation_ID", r->simulation_ID, sizeof r->simulation_ID);
    D("radius", p.A);
    D("rotation_rate", p.Omega);
    D("gravitation", p.Gravit);
    B("GravHeightVar", p.GravHeightVar);
    D("Rd", p.Rd);
    D("Cp", p.Cp);
    D("Tmean", p.Tmean);
    D("P_ref", p.P_Ref);
    D("Top_altitude", r->top_altitude);
    I("glevel", p.glevel);
    I("vlevel", r->vlevel);
    B("spring_dynamics", r->spring_dynamics);
    D("spring_beta", r->spring_beta);
    B("NonHydro", p.NonHydro);
    B("DeepModel", p.DeepModel);
    B("rest", r->rest);
    S("initial", r->initial, sizeof r->initial);
    B("gcm_off", r->gcm_off);
    B("HyDiff", p.HyDiff);
    I("HyDiffOrder", p.HyDiffOrder);
    B("DivDampP", p.DivDampP);
    B("VertHyDiff", p.VertHyDiff);
    I("VertHyDiffOrder", p.VertHyDiffOrder);
    D("Diffc", p.Diffc);
    D("DivDampc", p.DivDampc);
    D("Diffc_v", p.Diffc_v);
    B("RayleighSponge", r->RayleighSponge);
    B("RayleighSpongeT", s.temp_sponge);
    B("damp_uv_to_mean", s.damp_uv_to_mean);
    B("damp_w_to_mean", s.damp_w_to_mean);
    I("nlat_bins", s.nlat_bins);
    D("ns_ray_sponge", s.ns);
    D("Ruv_sponge", s.Ruv);
    D("Rw_sponge", s.Rw);
    D("RT_sponge", s.RT_sponge);
    B("shrink_sponge", s.shrink_sponge);
    I("t_shrink", s.t_shrink);
    B("DiffSponge", p.DiffSponge);
    D("Dv_sponge", p.Dv_sponge);
    D("ns_diff_sponge", p.ns_diff_sponge);
    I("order_diff_sponge", p.order_diff_sponge);
    D("Tstar", r->Tstar);
    D("planet_star_dist", r->planet_star_dist);
    D("radius_star", r->radius_star);
    D("albedo", rt.albedo);
    D("diff_ang", rt.diff_ang);
    D("n_lw", rt.n_lw);
    D("n_sw", rt.n_sw);
    D("kappa_lw", rt.kappa_lw);
    D("kappa_sw", rt.kappa_sw);
    D("f_lw", rt.f_lw);
    D("Tint", rt.Tint);
    I("GPU_ID_N", r->GPU_ID_N);
    I("n_substeps", p.n_substeps);
    if (!bad.empty()) return fail(THOR_ERR_BAD_ARG, "thor_config_run: bad value for key " + bad);
    if (const char *m = get("raysp_calc_mode")) {
        const std::string v(m);
        if (v == "imp") s.raysp_calc_mode = 0;
        else if (v == "exp1") s.raysp_calc_mode = 1;
        else if (v == "exp3") s.raysp_calc_mode = 3;
        else return fail(THOR_ERR_BAD_ARG, "thor_config_run: raysp_calc_mode must be imp, exp1 or exp3");
    }
    if (const char *m = get("precision")) {
        const std::string v(m);
        if (v == "fp64") p.precision = THOR_PRECISION_FP64;
        else if (v == "mixed") p.precision = THOR_PRECISION_MIXED;
        else return fail(THOR_ERR_BAD_ARG, "thor_config_run: precision must be fp64 or mixed");
    }
    // the radiative-transfer module replaces the benchmark forcing
    int radiative_transfer = 0;
    B("radiative_transfer", radiative_transfer);
    const std::string bench = get("core_benchmark") ? get("core_benchmark") : "HeldSuarez";
    if (radiative_transfer) {
        if (bench != "NoBenchmark")
            return fail(THOR_ERR_UNSUPPORTED, "thor_config_run: radiative_transfer needs core_benchmark = NoBenchmark");
        r->forcing = THOR_FORCING_DOUBLE_GRAY;
        // incident flux at the planet (radiative_transfer.cu:1431-1432), R_sun and au in m
        const double R_SUN = 695700000.0, AU = 149597870700.0, SIGMA = 5.6703744191844314e-08;
        const double q = (r->radius_star * R_SUN) / (r->planet_star_dist * AU);
        rt.incflx = q * q * SIGMA * std::pow(r->Tstar, 4);
        rt.ps0 = p.P_Ref;
        rt.lon_sub = 0.0;
    } else if (bench == "HeldSuarez") {
        r->forcing = THOR_FORCING_HELD_SUAREZ;
    } else if (bench == "NoBenchmark") {
        r->forcing = THOR_FORCING_NONE;
    } else {
        return fail(THOR_ERR_UNSUPPORTED, "thor_config_run: core_benchmark " + bench +
                                              " is not run by this build (HeldSuarez, NoBenchmark)");
    }
    auto must = [&](const char *k, const char *want) -> bool {
        const char *v = get(k);
        return !v || std::string(v) == want;
    };
    // thermo_equation (esp.cu:441-443, 696-710): entropy (default) or energy
    if (const char *v = get("thermo_equation")) {
        if (std::string(v) == "entropy") p.thermo_equation = THOR_THERMO_ENTROPY;
        else if (std::string(v) == "energy") p.thermo_equation = THOR_THERMO_ENERGY;
        else return fail(THOR_ERR_BAD_ARG, std::string("thor_config_run: thermo_equation not recognised: ") + v);
    }
    if (!must("vert_refined", "false"))
        return fail(THOR_ERR_UNSUPPORTED, "thor_config_run: vert_refined must be false (uniform vertical grid)");
    if (!must("conv_adj", "false")) return fail(THOR_ERR_UNSUPPORTED, "thor_config_run: conv_adj must be false");
    if (!must("surface", "false")) return fail(THOR_ERR_UNSUPPORTED, "thor_config_run: surface must be false");
    // boundary-layer module (boundary_layer.cu:339-362, defaults boundary_layer.h:104-105)
    B("boundary_layer", r->boundary_layer);
    r->bl.surf_drag = 1.0 / 86400.0;
    r->bl.bl_sigma = 0.7;
    if (r->boundary_layer) {
        if (!must("bl_type", "RayleighHS"))
            return fail(THOR_ERR_UNSUPPORTED, "thor_config_run: bl_type must be RayleighHS (LocalMixL is not run)");
        D("surf_drag", r->bl.surf_drag);
        D("bl_sigma", r->bl.bl_sigma);
        if (!bad.empty()) return fail(THOR_ERR_BAD_ARG, "thor_config_run: bad value for key " + bad);
    }
    if (!must("ultrahot_thermo", "none") || !must("ultrahot_heating", "none"))
        return fail(THOR_ERR_UNSUPPORTED, "thor_config_run: ultrahot_thermo / ultrahot_heating must be none");
    if (!must("output_mean", "false") && get("output_mean"))
        std::fprintf(stderr, "thor_config_run: output_mean is not written by this build\n");
    if (r->num_steps < 1 || r->n_out < 1 || p.timestep <= 0 || r->vlevel < 2)
        return fail(THOR_ERR_BAD_ARG, "thor_config_run: need num_steps, n_out > 0, timestep > 0, vlevel >= 2");
    return THOR_OK;
}

extern "C" thor_status thor_h5_write(const char *file, int create, const char *name, const double *data, long n,
                                     const char *unit, const char *description) {
    if (!file || !name || !data || n < 1) return fail(THOR_ERR_BAD_ARG, "thor_h5_write: bad argument");
    Hid f(open_file(file, create != 0), H5Fclose);
    if (f < 0) return fail(THOR_ERR_IO, std::string("thor_h5_write: cannot open ") + file);
    return write_table(f, name, data, n, unit ? unit : "", description ? description : "");
}

extern "C" long thor_h5_size(const char *file, const char *name) {
    if (!file || !name) return -1;
    static H5Quiet quiet;
    (void)quiet;
    Hid f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    return f < 0 ? -1 : table_size(f, name);
}

extern "C" thor_status thor_h5_read(const char *file, const char *name, double *data, long n) {
    if (!file || !name || !data) return fail(THOR_ERR_BAD_ARG, "thor_h5_read: null argument");
    static H5Quiet quiet;
    (void)quiet;
    Hid f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f < 0) return fail(THOR_ERR_IO, std::string("thor_h5_read: cannot open ") + file);
    return read_table(f, name, data, n);
}

// Names of the root group's datasets, newline-separated, in creation order
// where the file tracks it (else by name).
extern "C" int thor_h5_names(const char *file, char *buf, int len) {
    if (!file || !buf || len < 1) return fail(THOR_ERR_BAD_ARG, "thor_h5_names: bad argument");
    static H5Quiet quiet;
    (void)quiet;
    Hid f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f < 0) return fail(THOR_ERR_IO, std::string("thor_h5_names: cannot open ") + file);
    std::string names;
    auto cb = [](hid_t, const char *name, const H5L_info_t *, void *op) -> herr_t {
        auto *s = static_cast<std::string *>(op);
        s->append(name);
        s->push_back('\n');
        return 0;
    };
    if (H5Literate(f, H5_INDEX_NAME, H5_ITER_INC, nullptr, cb, &names) < 0)
        return fail(THOR_ERR_IO, "thor_h5_names: iteration failed");
    if ((int)names.size() + 1 > len) return fail(THOR_ERR_BAD_ARG, "thor_h5_names: buffer too small");
    std::memcpy(buf, names.c_str(), names.size() + 1);
    return THOR_OK;
}

// 1 for an integer dataset, 0 for floating point, -1 if absent.
extern "C" int thor_h5_is_int(const char *file, const char *name) {
    if (!file || !name) return -1;
    static H5Quiet quiet;
    (void)quiet;
    Hid f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f < 0 || H5Lexists(f, dname(name).c_str(), H5P_DEFAULT) <= 0) return -1;
    Hid ds(H5Dopen2(f, dname(name).c_str(), H5P_DEFAULT), H5Dclose);
    Hid t(H5Dget_type(ds), H5Tclose);
    return H5Tget_class(t) == H5T_INTEGER ? 1 : 0;
}

extern "C" thor_status thor_h5_attr(const char *file, const char *name, const char *attr, char *buf, int len) {
    if (!file || !name || !attr || !buf || len < 1) return fail(THOR_ERR_BAD_ARG, "thor_h5_attr: bad argument");
    static H5Quiet quiet;
    (void)quiet;
    Hid f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f < 0) return fail(THOR_ERR_IO, std::string("thor_h5_attr: cannot open ") + file);
    Hid ds(H5Dopen2(f, dname(name).c_str(), H5P_DEFAULT), H5Dclose);
    if (ds < 0) return fail(THOR_ERR_IO, std::string("thor_h5_attr: no dataset ") + name);
    Hid a(H5Aopen(ds, attr, H5P_DEFAULT), H5Aclose);
    if (a < 0) return fail(THOR_ERR_IO, std::string("thor_h5_attr: no attribute ") + attr);
    Hid t(H5Aget_type(a), H5Tclose);
    const size_t sz = H5Tget_size(t);
    if (H5Tget_class(t) != H5T_STRING || H5Tis_variable_str(t) > 0)
        return fail(THOR_ERR_IO, "thor_h5_attr: not a fixed-length string");
    std::vector<char> tmp(sz + 1, 0);
    if (H5Aread(a, t, tmp.data()) < 0) return fail(THOR_ERR_IO, "thor_h5_attr: read failed");
    std::snprintf(buf, (size_t)len, "%s", tmp.data());
    return THOR_OK;
}

extern "C" thor_status thor_output_grid(const char *dir, const thor_run_config *r, const thor_grid_desc *g) {
    if (!r || !g) return fail(THOR_ERR_BAD_ARG, "thor_output_grid: null argument");
    Out o(path_of(dir, std::string("esp_output_grid_") + r->simulation_ID + ".h5"));
    const int P = g->point_num, nv = g->nv;
    o.table("/Altitude", g->Altitude, nv, "m", "Altitude");
    o.table("/Altitudeh", g->Altitudeh, nv + 1, "m", "Altitude at the interfaces");
    o.table("/areasT", g->areasT, P, "m^2", "Main cells areas");
    o.table("/lonlat", g->lonlat, 2L * P, "-", "Longitudes and latitudes");
    o.value("/point_num", (double)P, "-", "Number of grid points in one level");
    o.value("/nv", (double)nv, "-", "Number of vertical layers");
    o.itable("/pntloc", g->point_local, 6L * P, "-", "Neighbours indexes");
    o.table("/grad", g->grad, 21L * P, "m^-1", "Horizontal gradient operator");
    o.table("/div", g->div, 21L * P, "m^-1", "Horizontal divergence operator");
    return o.rc;
}

extern "C" thor_status thor_output_planet(const char *dir, const thor_run_config *r) {
    if (!r) return fail(THOR_ERR_BAD_ARG, "thor_output_planet: null argument");
    Out o(path_of(dir, std::string("esp_output_planet_") + r->simulation_ID + ".h5"));
    const thor_params &p = r->params;
    o.ivalue("/glevel", p.glevel, "-", "Horizontal subdivision level");
    o.ivalue("/vlevel", r->vlevel, "-", "Vertical subdivision level");
    o.value("/spring_dynamics", r->spring_dynamics ? 1.0 : 0.0, "-", "Spring dynamics");
    o.value("/spring_beta", r->spring_beta, "-", "Spring Beta");
    o.value("/A", p.A, "m", "Planet radius");
    o.value("/Rd", p.Rd, "J/(Kg K)", "Gas constant");
    o.value("/Omega", p.Omega, "1/s", "Rotation rate");
    o.value("/Gravit", p.Gravit, "m/s^2", "Surface gravity");
    o.value("/P_Ref", p.P_Ref, "Pa", "Reference pressure");
    o.value("/Top_altitude", r->top_altitude, "m", "Top of the model's domain");
    o.value("/Cp", p.Cp, "J/(Kg K)", "Specific heat capacity");
    o.value("/RayleighSponge", r->RayleighSponge ? 1.0 : 0.0, "-", "Using Rayleigh SpongeLayer?");
    o.value("/RayleighSpongeT", r->sponge.temp_sponge ? 1.0 : 0.0, "-", "Using thermal SpongeLayer?");
    o.value("/DiffSponge", p.DiffSponge ? 1.0 : 0.0, "-", "Using Diffusive SpongeLayer?");
    o.value("/DeepModel", p.DeepModel ? 1.0 : 0.0, "-", "Using Deep Model");
    o.value("/NonHydro", p.NonHydro ? 1.0 : 0.0, "-", "Using Non Hydrostatic parameter");
    o.value("/output_mean", 0.0, "-", "outputting mean quantities");
    o.value("/output_diffusion", 0.0, "-", "outputting diffusion terms");
    o.value("/out_interm_momentum", 0.0, "-", "outputting intermediate momentum values");
    o.value("/DivDampP", p.DivDampP ? 1.0 : 0.0, "-", "Using Divergence-damping");
    o.value("/HyDiff", p.HyDiff ? 1.0 : 0.0, "-", "Using Hyper diffusion");
    o.value("/Diffc", p.Diffc, "-", "Hyperdiffusion strength");
    o.value("/Diffc_v", p.Diffc_v, "-", "Vertical hyperdiffusion strength");
    o.value("/Tmean", p.Tmean, "K", "Mean atmospheric temperature");
    o.value("/conv_adj", 0.0, "-", "Using convection adjustment");
    o.value("/gcm_off", r->gcm_off ? 1.0 : 0.0, "-", "GCM off");
    o.value("/single_column", 0.0, "-", "single column mode");
    o.value("/rest", r->rest ? 1.0 : 0.0, "-", "Starting from rest");
    const int bench = r->forcing == THOR_FORCING_HELD_SUAREZ ? 1 : 0;  // define.h:151-160
    o.ivalue("/core_benchmark", bench, "-", "Using benchmark forcing");
    if (r->RayleighSponge) {
        o.ivalue("/nlat_bins", r->sponge.nlat_bins, "-", "number of lat rings for Rayleigh sponge layer");
        o.value("/ns_ray_sponge", r->sponge.ns, "-", "Bottom of rayleigh sponge layer");
        o.value("/Ruv_sponge", r->sponge.Ruv, "1/s", "Strength of Rayleigh sponge layer (uv components)");
        o.value("/Rw_sponge", r->sponge.Rw, "1/s", "Strength of Rayleigh sponge layer (w component)");
    }
    if (p.DiffSponge) {
        o.ivalue("order_diff_sponge", p.order_diff_sponge, "-", "Order of diffusive sponge");
        o.value("/ns_diff_sponge", p.ns_diff_sponge, "-", "Bottom of diffusive sponge layer");
        o.value("/Dv_sponge", p.Dv_sponge, "-", "Strength of diffusive sponge layer");
    }
    o.value("/Tint", r->rt.Tint, "K", "Temperature of interior heat flux");
    o.value("/f_lw", r->rt.f_lw, "-", "fraction of taulw in well-mixed absorber");
    o.value("/kappa_lw", r->rt.kappa_lw, "-", "longwave opacity");
    o.value("/kappa_sw", r->rt.kappa_sw, "-", "shortwave opacity");
    o.value("/phy_module", 0.0, "-",
            r->forcing == THOR_FORCING_DOUBLE_GRAY ? "radiative_transfer" : "none");
    o.value("/surface", 0.0, "-", "include solid/liquid surface");
    o.value("/Csurf", 1e7, "J/K/m^2", "heat capacity of surface by area");
    return o.rc;
}

extern "C" thor_status thor_output_state(const char *dir, const thor_run_config *r, int fidx, int nstep,
                                         double simulation_time, const thor_grid_desc *g, const double *Mh,
                                         const double *Wh, const double *Rho, const double *pressure,
                                         const double *Qheat, const double glob[5]) {
    if (!r || !g || !Mh || !Wh || !Rho || !pressure || !Qheat)
        return fail(THOR_ERR_BAD_ARG, "thor_output_state: null argument");
    Out o(path_of(dir, std::string("esp_output_") + r->simulation_ID + "_" + std::to_string(fidx) + ".h5"));
    const long N = (long)g->point_num * g->nv;
    o.ivalue("/nstep", nstep, "-", "Step number");
    o.value("/simulation_time", simulation_time, "s", "Simulation time");
    o.table("/Rho", Rho, N, "kg/m^3", "Density");
    o.table("/Pressure", pressure, N, "Pa", "Pressure");
    o.table("/Mh", Mh, 3 * N, "kg m^-2 s^-1", "Horizontal Momentum");
    o.table("/Wh", Wh, (long)g->point_num * (g->nv + 1), "kg m^-2 s^-1", "Vertical Momentum");
    if (glob) {
        o.value("/GlobalE", glob[1], "kg m^2/s^2", "Global Total Energy");
        o.value("/GlobalMass", glob[0], "kg", "Global Mass");
        o.value("/GlobalAMx", glob[2], "kg m^2/s", "Global AngMomX");
        o.value("/GlobalAMy", glob[3], "kg m^2/s", "Global AngMomY");
        o.value("/GlobalAMz", glob[4], "kg m^2/s", "Global AngMomZ");
    }
    o.table("/Qheat", Qheat, N, "W m^-3", "Physics module Qheat");
    std::vector<double> rd((size_t)N, r->params.Rd), cp((size_t)N, r->params.Cp);
    o.table("/Rd", rd.data(), N, "J/K/kg", "Local gas constant");
    o.table("/Cp", cp.data(), N, "J/K/kg", "Local heat capacity");
    return o.rc;
}

extern "C" thor_status thor_restart_read(const char *state_file, const char *planet_file, const thor_run_config *r,
                                         const thor_grid_desc *g, int *nstep, double *simulation_time, double *Mh,
                                         double *Wh, double *W, double *Rho, double *pressure) {
    if (!state_file || !planet_file || !r || !g || !nstep || !simulation_time || !Mh || !Wh || !W || !Rho || !pressure)
        return fail(THOR_ERR_BAD_ARG, "thor_restart_read: null argument");
    static H5Quiet quiet;
    (void)quiet;
    {
        // the planet must be the configured one (esp_initial.cu:1150-1207)
        Hid f(H5Fopen(planet_file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        if (f < 0) return fail(THOR_ERR_IO, std::string("thor_restart_read: cannot open ") + planet_file);
        const std::pair<const char *, double> want[] = {{"/A", r->params.A},
                                                        {"/Top_altitude", r->top_altitude},
                                                        {"/glevel", (double)r->params.glevel},
                                                        {"/vlevel", (double)r->vlevel},
                                                        {"/surface", 0.0}};
        for (const auto &w : want) {
            double v = 0.0;
            const int rc = read_table(f, w.first, &v, 1);
            if (rc != THOR_OK) return rc;
            if (v != w.second)
                return fail(THOR_ERR_BAD_ARG, std::string("thor_restart_read: ") + w.first + " of the planet file (" +
                                                  std::to_string(v) + ") differs from the configuration (" +
                                                  std::to_string(w.second) + ")");
        }
    }
    Hid f(H5Fopen(state_file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f < 0) return fail(THOR_ERR_IO, std::string("thor_restart_read: cannot open ") + state_file);
    const int P = g->point_num, nv = g->nv;
    const long N = (long)P * nv;
    double ns = 0.0;
    int rc = read_table(f, "/nstep", &ns, 1);
    if (rc == THOR_OK) rc = read_table(f, "/simulation_time", simulation_time, 1);
    if (rc == THOR_OK) rc = read_table(f, "/Rho", Rho, N);
    if (rc == THOR_OK) rc = read_table(f, "/Pressure", pressure, N);
    if (rc == THOR_OK) rc = read_table(f, "/Mh", Mh, 3 * N);
    if (rc == THOR_OK) rc = read_table(f, "/Wh", Wh, (long)P * (nv + 1));
    if (rc != THOR_OK) return rc;
    std::vector<double> rd((size_t)N), cp((size_t)N);
    rc = read_table(f, "/Rd", rd.data(), N);
    if (rc == THOR_OK) rc = read_table(f, "/Cp", cp.data(), N);
    if (rc != THOR_OK) return rc;
    for (long i = 0; i < N; i++)
        if (rd[i] != r->params.Rd || cp[i] != r->params.Cp)
            return fail(THOR_ERR_UNSUPPORTED, "thor_restart_read: /Rd and /Cp must be uniform and equal to the "
                                              "configuration's (per-cell values are not run by this build)");
    *nstep = (int)ns;
    // W at the layer centres from Wh (esp_initial.cu:1263-1277)
    for (int i = 0; i < P; i++)
        for (int lev = 0; lev < nv; lev++) {
            const double xi = g->Altitude[lev], xim1 = g->Altitudeh[lev], xip1 = g->Altitudeh[lev + 1];
            const double a = (xi - xip1) / (xim1 - xip1), b = (xi - xim1) / (xip1 - xim1);
            W[(size_t)i * nv + lev] = Wh[(size_t)i * (nv + 1) + lev] * a + Wh[(size_t)i * (nv + 1) + lev + 1] * b;
        }
    return THOR_OK;
}
