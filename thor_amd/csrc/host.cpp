// Host-side entry points of the C ABI: error reporting, grid construction,
// default parameters and the rest initial condition.
#include "../../include/thor_dyn.h"
#include "common.h"
#include "grid.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace thor {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
} // namespace thor

extern "C" const char *thor_last_error(void) { return thor::g_last_error.c_str(); }

struct thor_grid {
    thor::IcoGrid g;
};

extern "C" thor_status thor_grid_build(int glevel, int nv, double radius, double top_altitude,
                                       int spring_dynamics, double spring_beta, const char *cache_dir,
                                       thor_grid **out) {
    if (!out) return thor::fail(THOR_ERR_BAD_ARG, "thor_grid_build: null output pointer");
    if (glevel < 4 || glevel > 8)
        return thor::fail(THOR_ERR_BAD_ARG, "thor_grid_build: glevel must be in 4..8 (got " + std::to_string(glevel) + ")");
    if (nv < 2 || radius <= 0 || top_altitude <= 0)
        return thor::fail(THOR_ERR_BAD_ARG, "thor_grid_build: need nv >= 2, radius > 0 and top_altitude > 0");
    try {
        thor_grid *tg = new thor_grid();
        bool loaded = false;
        std::string path;
        if (cache_dir && cache_dir[0]) {
            char buf[512];
            std::snprintf(buf, sizeof buf, "%s/thor_grid_g%d_%d_%.6f.bin", cache_dir, glevel,
                          spring_dynamics ? 1 : 0, spring_beta);
            path = buf;
            loaded = thor::load_horizontal(tg->g, path, glevel, spring_dynamics != 0, spring_beta);
        }
        if (!loaded) {
            thor::build_horizontal(tg->g, glevel, spring_dynamics != 0, spring_beta);
            if (!path.empty()) thor::save_horizontal(tg->g, path);
        }
        thor::finish_grid(tg->g, radius, top_altitude, nv);
        *out = tg;
        return THOR_OK;
    } catch (const std::exception &e) {
        return thor::fail(THOR_ERR_BAD_ARG, std::string("thor_grid_build: ") + e.what());
    }
}

extern "C" thor_status thor_grid_describe(const thor_grid *tg, thor_grid_desc *d) {
    if (!tg || !d) return thor::fail(THOR_ERR_BAD_ARG, "thor_grid_describe: null argument");
    const thor::IcoGrid &g = tg->g;
    d->point_num = g.point_num;
    d->nv = g.nv;
    d->nr = g.nr;
    d->nl_region = g.nl_region;
    d->glevel = g.glevel;
    for (int k = 0; k < 12; k++) d->pent_ind[k] = g.pent_ind[k];
    d->point_local = g.point_local.data();
    d->maps = g.maps.data();
    d->div = g.div.data();
    d->grad = g.grad.data();
    d->areas = g.areas.data();
    d->areasTr = g.areasTr.data();
    d->areasT = g.areasT.data();
    d->nvecoa = g.nvecoa.data();
    d->nvecti = g.nvecti.data();
    d->nvecte = g.nvecte.data();
    d->func_r = g.func_r.data();
    d->lonlat = g.lonlat.data();
    d->Altitude = g.Altitude.data();
    d->Altitudeh = g.Altitudeh.data();
    d->xyz = g.xyz.data();
    d->xyzq = g.xyzq.data();
    return THOR_OK;
}

extern "C" int thor_grid_spring_iterations(const thor_grid *tg) { return tg ? tg->g.spring_iterations : -1; }

extern "C" void thor_grid_destroy(thor_grid *tg) { delete tg; }

extern "C" void thor_params_default_earth(thor_params *p) {
    // ifile/earth_hstest.thr
    std::memset(p, 0, sizeof(*p));
    p->A = 6371000.0;
    p->Omega = 7.292E-5;
    p->Gravit = 9.8;
    p->P_Ref = 100000.0;
    p->Rd = 287.0;
    p->Cp = 1005.0;
    p->Tmean = 300.0;
    p->Diffc = 0.0048;
    p->DivDampc = 0.0048;
    p->timestep = 1000;
    p->glevel = 4;
    p->HyDiff = 1;
    p->HyDiffOrder = 4;
    p->DivDampP = 1;
    p->DeepModel = 1;
    p->NonHydro = 1;
    p->GravHeightVar = 0;
    p->n_substeps = 6;
    p->precision = THOR_PRECISION_FP64;
    p->VertHyDiff = 0;  // define.h:84-85
    p->VertHyDiffOrder = 6;
    p->Diffc_v = 0.0;
    p->DiffSponge = 0;  // define.h:70-73
    p->order_diff_sponge = 2;
    p->Dv_sponge = 0.01;
    p->ns_diff_sponge = 0.75;
}

extern "C" thor_status thor_ic_rest(const thor_grid_desc *g, const thor_params *p, unsigned long long seed,
                                    double temperature_noise, double wind_noise, double *Mh, double *Wh,
                                    double *W, double *Rho, double *pressure) {
    if (!g || !p || !Mh || !Wh || !W || !Rho || !pressure)
        return thor::fail(THOR_ERR_BAD_ARG, "thor_ic_rest: null argument");
    const int P = g->point_num, nv = g->nv;
    const double *Alt = g->Altitude, *Alth = g->Altitudeh;
    const double Rd = p->Rd, eps = 1e-15, A = p->A;
    const int it_max = 100;
    // Column 0: Newton-Raphson solve of the discrete hydrostatic equation,
    // level by level from the surface (esp_initial.cu:520-587).
    std::vector<double> pc(nv), rc(nv), tc(nv);
    for (int lev = 0; lev < nv; lev++) {
        double P_L, rho_L, T_L, dz, g_L, l_int, r_int;
        if (lev == 0) {
            P_L = p->P_Ref;
            rho_L = p->P_Ref / (Rd * p->Tmean);
            T_L = p->Tmean;
            dz = Alt[0];
            g_L = p->GravHeightVar ? p->Gravit * std::pow(A / (A + Alt[0]), 2) : p->Gravit;
            l_int = 0.5;
            r_int = 0.5;
        } else {
            P_L = pc[lev - 1];
            rho_L = rc[lev - 1];
            T_L = tc[lev - 1];
            dz = Alt[lev] - Alt[lev - 1];
            g_L = p->GravHeightVar ? p->Gravit * std::pow(A / (A + Alt[lev - 1]), 2) : p->Gravit;
            l_int = (Alt[lev] - Alth[lev]) / dz;
            r_int = (Alth[lev] - Alt[lev - 1]) / dz;
        }
        pc[lev] = P_L;
        rc[lev] = rho_L;
        tc[lev] = T_L;
        double ptmp = pc[lev] + 2 * eps * pc[lev];
        double gg = p->GravHeightVar ? p->Gravit * std::pow(A / (A + Alt[lev]), 2) : p->Gravit;
        for (int it = 0; it < it_max && (ptmp - pc[lev]) / pc[lev] > eps; it++) {
            ptmp = pc[lev];
            double f = (pc[lev] - P_L) / dz + (gg * r_int + g_L * l_int) * (rc[lev] * r_int + rho_L * l_int);
            double df = 1.0 / dz + (gg * r_int + g_L * l_int) * r_int / (Rd * tc[lev]);
            pc[lev] = pc[lev] - f / df;
            rc[lev] = pc[lev] / (Rd * tc[lev]);
        }
    }
    std::mt19937_64 rng(seed);
    auto uniform = [&](double a) {
        double u = (double)(rng() >> 11) * (1.0 / 9007199254740992.0);
        return a * (2.0 * u - 1.0);
    };
    for (int i = 0; i < P; i++) {
        for (int lev = 0; lev < nv; lev++) {
            double pr = pc[lev], T = tc[lev], rho = rc[lev];
            if (temperature_noise > 0) {
                T = T + uniform(temperature_noise);
                rho = pr / (Rd * T);
            }
            pressure[i * nv + lev] = pr;
            Rho[i * nv + lev] = rho;
            W[i * nv + lev] = 0.0;
            Wh[i * (nv + 1) + lev] = 0.0;
            double v[3] = {0.0, 0.0, 0.0};
            if (wind_noise > 0) {
                for (int k = 0; k < 3; k++) v[k] = uniform(wind_noise);
                const double *f = g->func_r + i * 3;
                double vr = v[0] * f[0] + v[1] * f[1] + v[2] * f[2];
                for (int k = 0; k < 3; k++) v[k] = v[k] - vr * f[k];
            }
            for (int k = 0; k < 3; k++) Mh[(i * nv + lev) * 3 + k] = rho * v[k];
        }
        Wh[i * (nv + 1) + nv] = 0.0;
    }
    return THOR_OK;
}
