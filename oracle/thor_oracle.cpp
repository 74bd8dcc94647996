// ORACLE — test infrastructure only (see thor_oracle.h).
//
// Each routine below restates one THOR kernel as a serial loop over the same
// (block, thread) indices the CUDA launch uses, reading neighbours through the
// rhombus maps exactly as the kernel's shared-memory halo load does
// (src/headers/kernel_halo_helpers.h:47-128). Build with -ffp-contract=off so
// every product and sum is rounded as written.
//
// Loops over independent blocks (rhombus tiles, columns) may run on several
// host threads (oracle_set_threads; one thread by default, the faithful
// serial loop). Every output element is written by exactly one iteration with
// the same arithmetic, so the results do not depend on the thread count.
#include "thor_oracle.h"
#include "../thor_amd/csrc/portable_math.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include <omp.h>

// independent iterations of an outer block or column loop
#define ORACLE_PAR _Pragma("omp parallel for schedule(dynamic)")

double oracle_tree_sum(const double *d, long n);

namespace {

const int NL = 16, NHL = 18, NHL2 = NHL * NHL;

// Shared-memory positions a 16x16 block loads: its interior, the four halo
// sides and the two corners (-1,-1), (16,16) (kernel_halo_helpers.h:73-120).
std::vector<int> make_load_positions() {
    std::vector<int> pos;
    for (int y = 0; y < NL; y++)
        for (int x = 0; x < NL; x++) pos.push_back((y + 1) * NHL + x + 1);
    for (int x = 0; x < NL; x++) {
        pos.push_back((x + 1) * NHL);                 // y==0 threads: left column
        pos.push_back((x + 1) * NHL + NHL - 3 + 2);   // y==3 threads: right column
        pos.push_back(x + 1);                         // y==7 threads: bottom row
        pos.push_back((NHL - 3 + 2) * NHL + (x + 1)); // y==11 threads: top row
    }
    pos.push_back(0);
    pos.push_back((NHL - 3 + 2) * NHL + (NHL - 3 + 2));
    return pos;
}
const std::vector<int> kLoad = make_load_positions();

struct Oracle {
    oracle_grid g;
    oracle_params p;
    int P, nv, nvi, nr;
    std::map<std::string, std::vector<double>> arr;
    std::vector<double> Kdh4, Kdhz, Kdv6, Kdh2;
    double nrq[8] = {0}, rtq[11] = {0};  // forcing parameters
    // Rayleigh sponge: ns, Ruv, Rw, nlat_bins, damp_uv_to_mean, damp_w_to_mean,
    // temp_sponge, RT_sponge
    bool bl_on = false;
    double blq[2] = {0};  // boundary layer (RayleighHS): surf_drag, bl_sigma
    bool sponge_on = false;
    double spq[11] = {0};  // ..., [8] raysp_calc_mode: 0 imp, 1 exp1, 3 exp3; [9] shrink_sponge, [10] t_shrink
    long nstep = 0;        // Thor steps completed (the current step is nstep + 1)
    std::vector<double> vbar, tbar;  // the sponge's zonal means (kept over the stages in exp1)
    bool sponge_imp() const { return sponge_on && (int)spq[8] == 0; }
    std::vector<int> ztab;  // [P][3]: latitude bin, points in the bin, index within the bin
    int max_count = 0;
    // vertical solve: keep each column's tridiagonal system (a, b, c, d per
    // interface, as "tri_a" .. "tri_d" [P][nv+1]) for the residual checks
    bool capture_tri = false;
    // chemistry tracers (thor_chemistry.h) driven through the physics hooks
    int ntr = 0;

    double *A_(const char *n) { return arr[n].data(); }
};

void sponge_explicit(Oracle &o, int rk);

// Fills a shared tile: f(global index) for loaded positions whose map entry is
// valid, 0.0 where the map holds -1 (missing pentagon corner).
template <class F> inline void fill(double *sh, const int *m, F f) {
    for (int pos : kLoad) {
        int gi = m[pos];
        sh[pos] = gi >= 0 ? f(gi) : 0.0;
    }
}

struct Stencil {
    int ir, id, iri, pt[7]; // pt[0] = ir, pt[1..6] = pt1..pt6
    bool pent;
};
inline Stencil stencil(const int *m, int x, int y) {
    Stencil s;
    s.ir = (y + 1) * NHL + x + 1;
    s.id = m[s.ir];
    s.iri = y * NL + x;
    s.pent = (x == 0 && y == 0 && m[0] == -1);
    s.pt[0] = s.ir;
    s.pt[1] = (y + 2) * NHL + x + 1;
    s.pt[2] = (y + 2) * NHL + x + 2;
    s.pt[3] = (y + 1) * NHL + x + 2;
    s.pt[4] = (y)*NHL + x + 1;
    s.pt[5] = s.pent ? ((y + 1) * NHL + x) : ((y)*NHL + x);
    s.pt[6] = (y + 1) * NHL + x;
    return s;
}
// The three ring neighbours of triangle j used by Diffusion_Op / DivM_Op
// (thor_diff.h:257-289, thor_div.h:174-205).
inline void tri_pts(int x, int y, int j, bool pent, int &p1, int &p2, int &p3) {
    const int n1 = (y + 2) * NHL + x + 1, n2 = (y + 2) * NHL + x + 2, n3 = (y + 1) * NHL + x + 2;
    const int n4 = (y)*NHL + x + 1, n5 = (y)*NHL + x, n6 = (y + 1) * NHL + x;
    if (j == 0) { p1 = n1; p2 = n2; p3 = n3; }
    else if (j == 1) { p1 = n2; p2 = n3; p3 = n4; }
    else if (j == 2) { p1 = n3; p2 = n4; p3 = pent ? n6 : n5; }
    else if (j == 3) { p1 = n4; p2 = pent ? n6 : n5; p3 = pent ? n1 : n6; }
    else if (j == 4) { p1 = pent ? n6 : n5; p2 = pent ? n1 : n6; p3 = pent ? n2 : n1; }
    else { p1 = n6; p2 = n1; p3 = n2; }
}

inline double rscale_of(const Oracle &o, int lev) {
    return o.p.DeepModel ? o.p.A / (o.g.Altitude[lev] + o.p.A) : 1.0;
}

// ---------------------------------------------------------------- advection
// Compute_Advec_Cori1 (thor_adv_cor.h:51-235).
void advec_cori1(Oracle &o) {
    const int nv = o.nv;
    double *Adv = o.A_("Adv"), *v = o.A_("v"), *Mh = o.A_("Mhk"), *W = o.A_("Wk"), *Rho = o.A_("Rhok");
    const double *div = o.g.div, *fr = o.g.func_r;
    ORACLE_PAR
    for (int ib = 0; ib < o.nr; ib++) {
        double Ms[3][NHL2], vs[3][NHL2];
        const int *m = o.g.maps + ib * NHL2;
        for (int lev = 0; lev < nv; lev++) {
            for (int pos : kLoad) {
                int gi = m[pos];
                if (gi >= 0) {
                    double rho = Rho[gi * nv + lev], w = W[gi * nv + lev];
                    for (int k = 0; k < 3; k++) Ms[k][pos] = Mh[(gi * nv + lev) * 3 + k];
                    for (int k = 0; k < 3; k++) vs[k][pos] = Ms[k][pos] / rho + (w / rho) * fr[gi * 3 + k];
                } else {
                    for (int k = 0; k < 3; k++) Ms[k][pos] = vs[k][pos] = 0.0;
                }
            }
            double rs = rscale_of(o, lev);
            for (int y = 0; y < NL; y++)
                for (int x = 0; x < NL; x++) {
                    Stencil s = stencil(m, x, y);
                    const double *d = div + s.id * 21;
                    double nf[3] = {0.0, 0.0, 0.0};
                    for (int kk = 0; kk < 3; kk++)     // divergence component
                        for (int c = 0; c < 3; c++) {  // advected component
                            double sum = d[0 * 3 + kk] * vs[c][s.pt[0]] * Ms[kk][s.pt[0]];
                            for (int j = 1; j < 7; j++)
                                sum = sum + d[j * 3 + kk] * vs[c][s.pt[j]] * Ms[kk][s.pt[j]];
                            nf[c] += rs * sum;
                        }
                    for (int c = 0; c < 3; c++) {
                        Adv[(s.id * nv + lev) * 3 + c] = nf[c];
                        v[(s.id * nv + lev) * 3 + c] = vs[c][s.ir];
                    }
                }
        }
    }
}

// Compute_Advec_Cori_Poles (thor_adv_cor.h:241-362).
void advec_poles(Oracle &o) {
    const int nv = o.nv, P = o.P;
    double *Adv = o.A_("Adv"), *v = o.A_("v"), *Mh = o.A_("Mhk"), *W = o.A_("Wk"), *Rho = o.A_("Rhok");
    const double *div = o.g.div, *fr = o.g.func_r;
    for (int id = P - 2; id < P; id++) {
        int loc[6];
        loc[0] = id;
        for (int i = 0; i < 5; i++) loc[i + 1] = o.g.point_local[id * 6 + i];
        for (int lev = 0; lev < nv; lev++) {
            double vp[6][3], Mp[6][3];
            for (int i = 0; i < 6; i++) {
                int gi = loc[i];
                double rho = Rho[gi * nv + lev], w = W[gi * nv + lev];
                for (int k = 0; k < 3; k++) {
                    Mp[i][k] = Mh[(gi * nv + lev) * 3 + k];
                    vp[i][k] = Mp[i][k] / rho + (w / rho) * fr[gi * 3 + k];
                }
            }
            double rs = rscale_of(o, lev);
            double nf[3] = {0.0, 0.0, 0.0};
            for (int k = 0; k < 3; k++)
                for (int c = 0; c < 3; c++) {
                    double sum = div[id * 21 + 0 * 3 + k] * vp[0][c] * Mp[0][k];
                    for (int i = 1; i < 6; i++) sum = sum + div[id * 21 + i * 3 + k] * vp[i][c] * Mp[i][k];
                    nf[c] += rs * sum;
                }
            for (int c = 0; c < 3; c++) {
                Adv[(id * nv + lev) * 3 + c] = nf[c];
                v[(id * nv + lev) * 3 + c] = vp[0][c];
            }
        }
    }
}

// Compute_Advec_Cori2 (thor_adv_cor.h:366-536).
void advec_cori2(Oracle &o) {
    const int nv = o.nv;
    double *Adv = o.A_("Adv"), *v = o.A_("v"), *Wh = o.A_("Whk"), *Rho = o.A_("Rhok");
    const double *Alt = o.g.Altitude, *Alth = o.g.Altitudeh, *fr = o.g.func_r;
    const double A = o.p.A, Omega = o.p.Omega;
    ORACLE_PAR
    for (int id = 0; id < o.P; id++) {
        double altht = 0, althl = 0, altt = 0, alt = 0, altl = 0;
        double wht = 0, whl = 0, vxl = 0, vx = 0, vxt = 0, vyl = 0, vy = 0, vyt = 0, vzl = 0, vz = 0, vzt = 0;
        double dvwxl, dvwyl, dvwzl, dvwxt, dvwyt, dvwzt, intt, intl, xi, xim, xip;
        for (int lev = 0; lev < nv; lev++) {
            if (lev == 0) {
                dvwxl = 0.0; dvwyl = 0.0; dvwzl = 0.0;
                altht = Alth[lev + 1]; althl = Alth[lev]; alt = Alt[lev]; altt = Alt[lev + 1];
                vx = v[(id * nv + lev) * 3 + 0]; vxt = v[(id * nv + lev + 1) * 3 + 0];
                vy = v[(id * nv + lev) * 3 + 1]; vyt = v[(id * nv + lev + 1) * 3 + 1];
                vz = v[(id * nv + lev) * 3 + 2]; vzt = v[(id * nv + lev + 1) * 3 + 2];
                wht = Wh[id * (nv + 1) + lev + 1];
                xi = altht; xim = alt; xip = altt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                dvwxt = (vx * intt + vxt * intl) * wht;
                dvwyt = (vy * intt + vyt * intl) * wht;
                dvwzt = (vz * intt + vzt * intl) * wht;
            } else if (lev == nv - 1) {
                dvwxt = 0.0; dvwyt = 0.0; dvwzt = 0.0;
                xi = althl; xim = altl; xip = alt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                dvwxl = (vxl * intt + vx * intl) * whl;
                dvwyl = (vyl * intt + vy * intl) * whl;
                dvwzl = (vzl * intt + vz * intl) * whl;
            } else {
                xi = althl; xim = altl; xip = alt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                dvwxl = (vxl * intt + vx * intl) * whl;
                dvwyl = (vyl * intt + vy * intl) * whl;
                dvwzl = (vzl * intt + vz * intl) * whl;
                xi = altht; xim = alt; xip = altt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                dvwxt = (vx * intt + vxt * intl) * wht;
                dvwyt = (vy * intt + vyt * intl) * wht;
                dvwzt = (vz * intt + vzt * intl) * wht;
            }
            double r2p, r2m, r2l;
            if (o.p.DeepModel) {
                r2p = MPOW(altht + A, 2.0);
                r2m = MPOW(alt + A, 2.0);
                r2l = MPOW(althl + A, 2.0);
            } else {
                r2p = r2m = r2l = 1.0;
            }
            double dz = 1.0 / ((althl - altht) * r2m);
            double davx = (dvwxl * r2l - dvwxt * r2p) * dz;
            double davy = (dvwyl * r2l - dvwyt * r2p) * dz;
            double davz = (dvwzl * r2l - dvwzt * r2p) * dz;
            double rho = Rho[id * nv + lev];
            double Cx, Cy, Cz;
            if (o.p.DeepModel) {
                Cx = -2.0 * Omega * vy * rho;
                Cy = 2.0 * Omega * vx * rho;
                Cz = 0.0;
            } else {
                double f0 = fr[id * 3 + 0], f1 = fr[id * 3 + 1], f2 = fr[id * 3 + 2];
                Cx = 2 * Omega * (vz * f2 * f1 - vy * MPOW(f2, 2)) * rho;
                Cy = -2 * Omega * (vz * f2 * f0 - vx * MPOW(f2, 2)) * rho;
                Cz = 2 * Omega * (vy * f2 * f0 - vx * f2 * f1) * rho;
            }
            Adv[(id * nv + lev) * 3 + 0] += davx + Cx;
            Adv[(id * nv + lev) * 3 + 1] += davy + Cy;
            Adv[(id * nv + lev) * 3 + 2] += davz + Cz;
            if (lev < nv - 1) {
                althl = altht;
                altht = Alth[lev + 2];
                altl = alt;
                alt = altt;
                if (lev < nv - 2) altt = Alt[lev + 2];
                whl = wht;
                if (lev < nv - 2) wht = Wh[id * (nv + 1) + lev + 2];
                vxl = vx; vx = vxt;
                if (lev < nv - 2) vxt = v[(id * nv + lev + 2) * 3 + 0];
                vyl = vy; vy = vyt;
                if (lev < nv - 2) vyt = v[(id * nv + lev + 2) * 3 + 1];
                vzl = vz; vz = vzt;
                if (lev < nv - 2) vzt = v[(id * nv + lev + 2) * 3 + 2];
            }
        }
    }
}

// ------------------------------------------------ thermodynamic diagnostics
// Compute_Temperature_H_Pt_Geff, entropy form, calcT = true (thor_auxiliary.h:55-309).
void temperature_h_pt_geff(Oracle &o) {
    const int nv = o.nv;
    double *T = o.A_("temperature"), *pk = o.A_("pressurek"), *Rho = o.A_("Rhok"), *h_d = o.A_("h"),
           *hh = o.A_("hh"), *pt_d = o.A_("pt"), *pth = o.A_("pth"), *gtil = o.A_("gtil"),
           *gtilh_d = o.A_("gtilh"), *Wh = o.A_("Whk"), *Cp = o.A_("Cp"), *Rd = o.A_("Rd");
    // energy equation (thor_auxiliary.h:143-161, 212-230, 281-287): potential
    // and kinetic energy at centres and interfaces, the stage's total energy
    const bool energy = o.p.energy_equation;
    double *ep_d = o.A_("epotential"), *ek_d = o.A_("ekinetic"), *Et_d = o.A_("Etotal"),
           *eph_d = o.A_("epotentialh"), *ekh_d = o.A_("ekinetich"), *Mk = o.A_("Mhk"), *Wk = o.A_("Wk");
    const double *Alt = o.g.Altitude, *Alth = o.g.Altitudeh;
    const double P_Ref = o.p.P_Ref, Gravit = o.p.Gravit, A = o.p.A;
    const bool ghv = o.p.GravHeightVar;
    ORACLE_PAR
    for (int id = 0; id < o.P; id++) {
        double temperature, rho = 0, rhoh, rhol = 0, rhop, pressure = 0, h = 0, hl = 0, pt = 0, ptl = 0,
                                     pl = 0, pp, dpdz;
        double gtilh = 0, gtilht, dz, dp_dz, Cv, CvoCp, xi, xim, xip, intt, intl, extr;
        double alt = 0, altl = 0, alth = 0, altht = 0, ep = 0, epl = 0, ek = 0, ekl = 0;
        for (int lev = 0; lev < nv + 1; lev++) {
            if (lev < nv) {
                Cv = Cp[id * nv + lev] - Rd[id * nv + lev];
                CvoCp = Cv / Cp[id * nv + lev];
                if (lev > 0) {
                    pl = pressure; rhol = rho; hl = h; altl = alt;
                    if (energy) { epl = ep; ekl = ek; }
                    else ptl = pt;
                }
                pressure = pk[id * nv + lev];
                rho = Rho[id * nv + lev];
                temperature = pressure / (rho * Rd[id * nv + lev]);
                T[id * nv + lev] = temperature;
                alt = Alt[lev];
                alth = Alth[lev];
                altht = Alth[lev + 1];
                h = Cv * temperature + pressure / rho;
                h_d[id * nv + lev] = h;
                if (energy) {
                    if (ghv) ep = Gravit * MPOW(A / (A + Alt[lev]), 2) * Alt[lev];
                    else ep = Gravit * Alt[lev];
                    ep_d[id * nv + lev] = ep;
                    const long e = (long)id * nv + lev;
                    ek = 0.5 * (Mk[e * 3 + 0] * Mk[e * 3 + 0] + Mk[e * 3 + 1] * Mk[e * 3 + 1] + Mk[e * 3 + 2] * Mk[e * 3 + 2]
                                + Wk[e] * Wk[e])
                         / (Rho[e] * Rho[e]);
                    ek_d[e] = ek;
                    Et_d[e] = Rho[e] * (ep + ek) + Cv / Rd[e] * pressure;
                } else {
                    pt = (P_Ref / (Rd[id * nv + lev] * rho)) * MPOW(pressure / P_Ref, CvoCp);
                    pt_d[id * nv + lev] = pt;
                }
            }
            if (lev == 0) {
                extr = (-alt - Alt[lev + 1]) / (Alt[lev + 1] - alt);
                hh[id * (nv + 1) + 0] = h;
                if (ghv)
                    pl = pk[id * nv + 1] - rho * Gravit * MPOW(A / (A + Alt[0]), 2) * (-alt - Alt[1]);
                else
                    pl = pk[id * nv + 1] - rho * Gravit * (-alt - Alt[1]);
                rhol = Rho[id * nv + 1] + (Rho[id * nv + 1] - rho) * extr;
                rhoh = 0.5 * (rho + rhol);
                dz = 2.0 * Alt[0];
                dpdz = (pk[id * nv + 0] - pl) / dz;
                gtilh_d[id * (nv + 1) + 0] = -(1.0 / rhoh) * dpdz;
                gtilh = 0.0;
            } else if (lev == nv) {
                extr = (2 * Alth[nv] - alt - Alt[nv - 2]) / (Alt[nv - 1] - Alt[nv - 2]);
                Cv = Cp[id * nv + lev - 1] - Rd[id * nv + lev - 1];
                CvoCp = Cv / Cp[id * nv + lev - 1];
                hh[id * (nv + 1) + nv] = h;
                if (ghv)
                    pp = pk[id * nv + nv - 2]
                         - rho * Gravit * MPOW(A / (A + Alt[lev]), 2) * (2 * Alth[nv] - alt - Alt[nv - 2]);
                else
                    pp = pk[id * nv + nv - 2] - rho * Gravit * (2 * Alth[nv] - alt - Alt[nv - 2]);
                if (pp < 0) pp = 0;
                rhop = Rho[id * nv + nv - 2] + (Rho[id * nv + nv - 1] - Rho[id * nv + nv - 2]) * extr;
                rhoh = 0.5 * (rho + rhop);
                if (energy) {
                    if (ghv) {
                        eph_d[id * (nv + 1) + nv] = Gravit * MPOW(A / (A + Alth[nv]), 2) * Alth[nv];
                        eph_d[id * (nv + 1) + 0] = Gravit * MPOW(A / (A + Alth[0]), 2) * Alth[0];
                    } else {
                        eph_d[id * (nv + 1) + nv] = Gravit * Alth[nv];
                        eph_d[id * (nv + 1) + 0] = Gravit * Alth[0];
                    }
                    ekh_d[id * (nv + 1) + nv] =
                        ek_d[id * nv + nv - 2] + (ek_d[id * nv + nv - 1] - ek_d[id * nv + nv - 2]) * extr;
                    ekh_d[id * (nv + 1) + 0] = ek_d[id * nv + 1]
                                               + (ek_d[id * nv + 1] - ek_d[id * nv + 0]) * (Alth[0] - Alt[1])
                                                     / (Alt[1] - Alt[0]);
                } else {
                    pth[id * (nv + 1) + nv] =
                        pt_d[id * nv + nv - 2] + (pt_d[id * nv + nv - 1] - pt_d[id * nv + nv - 2]) * extr;
                    pth[id * (nv + 1) + 0] = pt_d[id * nv + 1]
                                             + (pt_d[id * nv + 1] - pt_d[id * nv + 0]) * (Alth[0] - Alt[1])
                                                   / (Alt[1] - Alt[0]);
                }
                dz = 2.0 * (Alth[nv] - Alt[nv - 1]);
                if (ghv)
                    pp = pk[id * nv + nv - 2]
                         - Rho[id * nv + nv - 1] * Gravit * MPOW(A / (A + Alt[nv - 1]), 2)
                               * (2 * Alth[nv] - Alt[nv - 1] - Alt[nv - 2]);
                else
                    pp = pk[id * nv + nv - 2]
                         - Rho[id * nv + nv - 1] * Gravit * (2 * Alth[nv] - Alt[nv - 1] - Alt[nv - 2]);
                dpdz = (pp - pk[id * nv + nv - 1]) / dz;
                gtilh_d[id * (nv + 1) + nv] = -(1.0 / rhoh) * dpdz;
                gtilht = 0.0;
                xi = alt; xim = alth; xip = altht;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                gtil[id * nv + nv - 1] = gtilh * intt + gtilht * intl;
            } else {
                dz = alt - altl;
                dp_dz = (pressure - pl) / dz;
                xi = alth; xim = altl; xip = alt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                hh[id * (nv + 1) + lev] = hl * intt + h * intl;
                if (energy) {
                    eph_d[id * (nv + 1) + lev] = epl * intt + ep * intl;
                    ekh_d[id * (nv + 1) + lev] = ekl * intt + ek * intl;
                } else {
                    pth[id * (nv + 1) + lev] = ptl * intt + pt * intl;
                }
                rhoh = rhol * intt + rho * intl;
                gtilht = -(Wh[id * (nv + 1) + lev] / rhoh) * dp_dz;
                xi = alt; xim = alth; xip = altht;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                gtil[id * nv + lev - 1] = gtilh * intt + gtilht * intl;
                gtilh = gtilht;
                gtilh_d[id * (nv + 1) + lev] = -(1.0 / rhoh) * dp_dz;
            }
        }
    }
}

// ---------------------------------------------------------------- diffusion
// One Laplacian around the control volume (thor_diff.h:359-494): returns the
// accumulated lap for the six triangles (five for pentagons).
template <class GetA>
inline double laplacian(const Oracle &o, int id, double rscale, int n_faces, GetA a, const int ring[6],
                        int aself) {
    (void)aself;
    const double *areas = o.g.areas + id * 18, *areasTr = o.g.areasTr + id * 6;
    const double *nti = o.g.nvecti + id * 18, *nte = o.g.nvecte + id * 18, *noa = o.g.nvecoa + id * 18;
    double lap = 0.0, lap1 = 0, lap2 = 0, lap3 = 0, lap4, lap5, lap6;
    double lapx1 = 0, lapx2 = 0, lapy1 = 0, lapy2 = 0, lapz1 = 0, lapz2 = 0, AT1 = 0, AT2, astar;
    const double a0 = a(-1);
    for (int j = 0; j < 6; j++) {
        int jp1 = (j + 1) % n_faces, jp2 = (j + 2) % n_faces;
        double a1 = a(ring[j % n_faces]), a2 = a(ring[jp1]), a3 = a(ring[jp2]);
        double vdiff = 0.5 * rscale;
        if (j == 0) {
            AT1 = rscale / areasTr[j];
            astar = (areas[j * 3 + 0] * a0 + areas[j * 3 + 1] * a1 + areas[j * 3 + 2] * a2) / areasTr[j];
            lap1 = (0.5 * (a1 + a0) - astar);
            lap2 = (0.5 * (a1 + a2) - astar);
            lap3 = (0.5 * (a0 + a2) - astar);
        }
        AT2 = rscale / areasTr[jp1];
        astar = (areas[jp1 * 3 + 0] * a0 + areas[jp1 * 3 + 1] * a2 + areas[jp1 * 3 + 2] * a3) / areasTr[jp1];
        lap4 = (0.5 * (a0 + a2) - astar);
        lap5 = (0.5 * (a2 + a3) - astar);
        lap6 = (0.5 * (a0 + a3) - astar);
        if (j == 0) {
            lapx1 = (-lap1 * nti[j * 3 + 0] + lap2 * nte[j * 3 + 0] + lap3 * nti[jp1 * 3 + 0]) * AT1;
            lapx2 = (-lap4 * nti[jp1 * 3 + 0] + lap5 * nte[jp1 * 3 + 0] + lap6 * nti[jp2 * 3 + 0]) * AT2;
            lapy1 = (-lap1 * nti[j * 3 + 1] + lap2 * nte[j * 3 + 1] + lap3 * nti[jp1 * 3 + 1]) * AT1;
            lapy2 = (-lap4 * nti[jp1 * 3 + 1] + lap5 * nte[jp1 * 3 + 1] + lap6 * nti[jp2 * 3 + 1]) * AT2;
            lapz1 = (-lap1 * nti[j * 3 + 2] + lap2 * nte[j * 3 + 2] + lap3 * nti[jp1 * 3 + 2]) * AT1;
            lapz2 = (-lap4 * nti[jp1 * 3 + 2] + lap5 * nte[jp1 * 3 + 2] + lap6 * nti[jp2 * 3 + 2]) * AT2;
        } else {
            lapx1 = lapx2;
            lapx2 = (-lap4 * nti[jp1 * 3 + 0] + lap5 * nte[jp1 * 3 + 0] + lap6 * nti[jp2 * 3 + 0]) * AT2;
            lapy1 = lapy2;
            lapy2 = (-lap4 * nti[jp1 * 3 + 1] + lap5 * nte[jp1 * 3 + 1] + lap6 * nti[jp2 * 3 + 1]) * AT2;
            lapz1 = lapz2;
            lapz2 = (-lap4 * nti[jp1 * 3 + 2] + lap5 * nte[jp1 * 3 + 2] + lap6 * nti[jp2 * 3 + 2]) * AT2;
        }
        lap += ((lapx1 + lapx2) * noa[j * 3 + 0] + (lapy1 + lapy2) * noa[j * 3 + 1]
                + (lapz1 + lapz2) * noa[j * 3 + 2])
               * vdiff;
        if (n_faces == 5 && j == 4) break;
    }
    return lap;
}

// Diffusion_Op + Diffusion_Op_Poles for one pass (thor_diff.h:54-880).
// stepnum == 0: argument is the fluid property; otherwise the previous result
// (times rho on the last pass for var > 0). Reads a frozen copy of diff.
void diffusion_pass(Oracle &o, int stepnum, bool laststep) {
    const int nv = o.nv, P = o.P;
    double *diff = o.A_("diff"), *Mh = o.A_("Mhk"), *Rho = o.A_("Rhok"), *T = o.A_("temperature"),
           *W = o.A_("Wk"), *Rd = o.A_("Rd"), *Cp = o.A_("Cp");
    double *diffmh = o.A_("diffmh"), *diffw = o.A_("diffw"), *diffrh = o.A_("diffrh"), *diffpr = o.A_("diffpr");
    std::vector<double> diff_in(diff, diff + (size_t)6 * P * nv);
    auto arg = [&](int gi, int lev, int var) -> double {
        double rho = Rho[gi * nv + lev];
        if (stepnum == 0) {
            double a;
            if (var == 0) a = rho;
            else if (var <= 3) a = Mh[(gi * nv + lev) * 3 + var - 1];
            else if (var == 4) a = W[gi * nv + lev];
            else if (o.p.energy_equation) a = (Cp[gi * nv + lev] - Rd[gi * nv + lev]) * T[gi * nv + lev];
            else a = Rd[gi * nv + lev] * T[gi * nv + lev];
            if (var >= 1 && var <= 4) a = a / rho;
            return a;
        } else if (laststep) {
            if (var == 0) return diff_in[(gi * nv + lev) * 6 + var];
            return diff_in[(gi * nv + lev) * 6 + var] * rho;
        }
        return diff_in[(gi * nv + lev) * 6 + var];
    };
    const double sign = std::pow(-1.0, o.p.HyDiffOrder / 2);
    // diffusive sponge (thor_diff.h:453-509): at pass order_diff_sponge / 2 the
    // pass input at the point, times (-1)^(order/2 + 1) rho Kdh2, is kept for
    // the momentum and w fields and added to the last pass's result
    const bool sponge_on = o.p.DiffSponge && o.p.order_diff_sponge != o.p.HyDiffOrder;
    const double ssp = MPOW(-1.0, o.p.order_diff_sponge / 2 + 1);
    double *dsp = o.A_("diff_sponge");
    auto sponge = [&](int id, int lev, int var, double &lap) {
        if (!sponge_on || var < 1 || var > 4) return;
        const long e = ((long)id * nv + lev) * 6 + var;
        if (stepnum == o.p.order_diff_sponge / 2) dsp[e] = ssp * Rho[id * nv + lev] * o.Kdh2[lev] * diff_in[e];
        if (laststep) lap += dsp[e];
    };
    auto store = [&](int id, int lev, int var, double lap) {
        if (laststep) {
            if (var == 0) diffrh[id * nv + lev] = lap;
            else if (var <= 3) diffmh[(id * nv + lev) * 3 + var - 1] = lap;
            else if (var == 4) diffw[id * nv + lev] = lap;
            else diffpr[id * nv + lev] = lap;
        } else {
            diff[(id * nv + lev) * 6 + var] = lap;
        }
    };
    ORACLE_PAR
    for (int vb = 0; vb < 6 * o.nr; vb++) {
        const int var = vb / o.nr, ib = vb % o.nr;
        {
            double as[NHL2];
            const int *m = o.g.maps + ib * NHL2;
            for (int lev = 0; lev < nv; lev++) {
                fill(as, m, [&](int gi) { return arg(gi, lev, var); });
                double rs = rscale_of(o, lev);
                double sdiff = -o.Kdh4[lev] * sign;
                for (int y = 0; y < NL; y++)
                    for (int x = 0; x < NL; x++) {
                        Stencil s = stencil(m, x, y);
                        int nf = s.pent ? 5 : 6;
                        int ring[6];
                        for (int j = 0; j < 6; j++) {
                            int p1, p2, p3;
                            tri_pts(x, y, j, s.pent, p1, p2, p3);
                            ring[j] = p1;
                        }
                        double lap = laplacian(o, s.id, rs, nf,
                                               [&](int hp) { return hp < 0 ? as[s.ir] : as[hp]; }, ring, s.ir);
                        if (laststep) lap *= sdiff;
                        sponge(s.id, lev, var, lap);
                        store(s.id, lev, var, lap);
                    }
            }
        }
    }
    // Poles.
    for (int var = 0; var < 6; var++)
        for (int id = P - 2; id < P; id++)
            for (int lev = 0; lev < nv; lev++) {
                int ring[6];
                for (int i = 0; i < 5; i++) ring[i] = o.g.point_local[id * 6 + i];
                ring[5] = ring[4];
                double rs = rscale_of(o, lev);
                double sdiff = -o.Kdh4[lev] * sign;
                double lap = laplacian(o, id, rs, 5,
                                       [&](int gi) { return gi < 0 ? arg(id, lev, var) : arg(gi, lev, var); },
                                       ring, id);
                if (laststep) lap = sdiff * lap;
                sponge(id, lev, var, lap);
                store(id, lev, var, lap);
            }
}

// Correct_Horizontal (thor_diff.h:882-915): the radial component is removed
// from the horizontal and the vertical momentum diffusion (the latter is zero
// unless VertHyDiff).
void correct_horizontal(Oracle &o) {
    const int nv = o.nv;
    const double *fr = o.g.func_r;
    for (const char *name : {"diffmh", "diffmv"}) {
        double *dm = o.A_(name);
        ORACLE_PAR
        for (int id = 0; id < o.P; id++)
            for (int lev = 0; lev < nv; lev++) {
                double *d = dm + (id * nv + lev) * 3;
                double r = fr[id * 3 + 0] * d[0] + fr[id * 3 + 1] * d[1] + fr[id * 3 + 2] * d[2];
                d[0] += -fr[id * 3 + 0] * r;
                d[1] += -fr[id * 3 + 1] * r;
                d[2] += -fr[id * 3 + 2] * r;
            }
    }
}

// vertical_diff (thor_diff.h:1198-1426, called at thor_driver.cu:590-612):
// per column and variable (0 rho, 1-3 the momentum components over rho,
// 4 W over rho, 5 T), a vertical Laplacian in flux form with r^2 weights,
// ghost levels below the first and above the last layer holding copies of
// the layer next to them, applied VertHyDiffOrder / 2 times. Before the last
// application the field is multiplied by sign Kdv6 (and by rho for
// variables 1-5), sign = (-1)^(order/2 + 1); the T result is multiplied by
// Rd. The density result is not used (diffrv = 0, include_rho = false).
void vertical_diff(Oracle &o) {
    const int nv = o.nv, P = o.P, order = o.p.VertHyDiffOrder;
    const double *Alt = o.g.Altitude, *Alth = o.g.Altitudeh, *Kv = o.Kdv6.data();
    const double *Rho = o.A_("Rhok"), *Mh = o.A_("Mhk"), *W = o.A_("Wk"), *T = o.A_("temperature");
    double *dmv = o.A_("diffmv"), *dwv = o.A_("diffwv"), *drv = o.A_("diffrv"), *dpv = o.A_("diffprv");
    const double sign = MPOW(-1.0, order / 2 + 1), Rd = o.p.Rd;
    // one application at layer lev of the ghost-extended column a[0..nv+1]
    // (a[lev + 1] is layer lev): flux difference over the layer, / r^2
    auto flux_div = [&](const std::vector<double> &a, int lev) {
        const double dzi = 1.0 / (Alth[lev + 1] - Alth[lev]);
        double ri = 1.0, rup = 1.0, rlow = 1.0;
        if (o.p.DeepModel) {
            ri = o.p.A + Alt[lev];
            rup = o.p.A + Alth[lev + 1];
            rlow = o.p.A + Alth[lev];
        }
        double dzup, dzlow;
        if (lev == 0) {
            dzup = 1.0 / (Alt[1] - Alt[0]);
            dzlow = 1.0 / (2.0 * Alt[0]);
        } else if (lev == nv - 1) {
            dzup = 1.0 / (2 * (Alth[lev + 1] - Alt[lev]));
            dzlow = 1.0 / (Alt[lev] - Alt[lev - 1]);
        } else {
            dzup = 1.0 / (Alt[lev + 1] - Alt[lev]);
            dzlow = 1.0 / (Alt[lev] - Alt[lev - 1]);
        }
        const double flux = MPOW(rup, 2) * (a[lev + 2] - a[lev + 1]) * dzup - MPOW(rlow, 2) * (a[lev + 1] - a[lev]) * dzlow;
        return std::make_pair(flux * dzi, MPOW(ri, 2));
    };
    ORACLE_PAR
    for (int id = 0; id < P; id++) {
        std::vector<double> a(nv + 2), d(nv);
        for (int var = 0; var < 6; var++) {
            auto value = [&](int lev) {
                const long e = (long)id * nv + lev;
                const double r = Rho[e];
                if (var == 0) return r;
                if (var <= 3) return Mh[e * 3 + var - 1] / r;
                if (var == 4) return W[e] / r;
                return T[e];
            };
            for (int lev = 0; lev < nv; lev++) a[lev + 1] = value(lev);
            a[0] = a[1];
            a[nv + 1] = a[nv];
            const int passes = order / 2 - 1;
            for (int l = 0; l < passes; l++) {
                for (int lev = 0; lev < nv; lev++) {
                    auto q = flux_div(a, lev);
                    d[lev] = q.first / q.second;
                }
                for (int lev = 0; lev < nv; lev++) {
                    if (l < passes - 1) a[lev + 1] = d[lev];
                    else if (var == 0) a[lev + 1] = sign * Kv[lev] * d[lev];
                    else a[lev + 1] = sign * Rho[(long)id * nv + lev] * Kv[lev] * d[lev];
                }
                a[0] = a[1];
                a[nv + 1] = a[nv];
            }
            for (int lev = 0; lev < nv; lev++) {
                auto q = flux_div(a, lev);
                const double out = var != 5 ? q.first / q.second : q.first * Rd / q.second;
                const long e = (long)id * nv + lev;
                if (var == 0) drv[e] = 0.0;
                else if (var <= 3) dmv[e * 3 + var - 1] = out;
                else if (var == 4) dwv[e] = out;
                else dpv[e] = out;
            }
        }
    }
}

// ------------------------------------------------------- divergence damping
// DivM_Op + DivM_Op_Poles (thor_div.h:49-535). M/Wh are Mhk/Whk in the slow
// part and Mhs/Whs in the acoustic substeps.
void divm_pass(Oracle &o, const char *Mname, const char *Whname, bool laststep) {
    const int nv = o.nv, P = o.P;
    double *DivM = o.A_("DivM"), *divg = o.A_("divg_Mh"), *M = o.A_(Mname), *Wh = o.A_(Whname);
    const double *areas = o.g.areas, *areasTr = o.g.areasTr, *nti = o.g.nvecti, *nte = o.g.nvecte,
                 *noa = o.g.nvecoa, *fr = o.g.func_r, *Alth = o.g.Altitudeh;
    std::vector<double> divg_in;
    if (laststep) divg_in.assign(divg, divg + (size_t)3 * P * nv);

    // Core of the per-point computation for one ring (pentagon-aware).
    auto point = [&](int id, int lev, int nfaces, bool pole, auto get_a, auto get_wl, auto get_wt, double out[3]) {
        double rs = rscale_of(o, lev);
        double AT0 = 0.5 * rs, lapx = 0.0, lapy = 0.0, lapz = 0.0;
        double dz = laststep ? 0.0 : 1.0 / (Alth[lev] - Alth[lev + 1]);
        double a01x = 0, a01y = 0, a01z = 0, a12x = 0, a12y = 0, a12z = 0, AT1 = 0, lap1 = 0, lap2 = 0;
        const double *ar = areas + id * 18, *aT = areasTr + id * 6, *ti = nti + id * 18, *te = nte + id * 18,
                     *oa = noa + id * 18;
        for (int j = 0; j < nfaces; j++) {
            int jp1 = (j + 1) % nfaces, jp2 = (j + 2) % nfaces;
            // ring slots: 0 = centre, j+1.. = ring neighbours
            double dwdz, dwdz2;
            if (!laststep) {
                double w1 = pole ? ar[j * 3 + 1] : ar[j * 3 + 2];
                double w2 = pole ? ar[jp1 * 3 + 1] : ar[jp1 * 3 + 2];
                double meanwl = (ar[j * 3 + 0] * get_wl(0) + ar[j * 3 + 1] * get_wl(j + 1) + w1 * get_wl(jp1 + 1)) / aT[j];
                double meanwt = (ar[j * 3 + 0] * get_wt(0) + ar[j * 3 + 1] * get_wt(j + 1) + w1 * get_wt(jp1 + 1)) / aT[j];
                double meanwl2 = (ar[jp1 * 3 + 0] * get_wl(0) + ar[jp1 * 3 + 1] * get_wl(jp1 + 1) + w2 * get_wl(jp2 + 1)) / aT[jp1];
                double meanwt2 = (ar[jp1 * 3 + 0] * get_wt(0) + ar[jp1 * 3 + 1] * get_wt(jp1 + 1) + w2 * get_wt(jp2 + 1)) / aT[jp1];
                dwdz = (meanwl - meanwt) * dz;
                dwdz2 = (meanwl2 - meanwt2) * dz;
            } else {
                dwdz = 0.0;
                dwdz2 = 0.0;
            }
            if (j == 0) {
                a01x = get_a(0, 0) + get_a(j + 1, 0);
                a01y = get_a(0, 1) + get_a(j + 1, 1);
                a01z = get_a(0, 2) + get_a(j + 1, 2);
                a12x = get_a(j + 1, 0) + get_a(jp1 + 1, 0);
                a12y = get_a(j + 1, 1) + get_a(jp1 + 1, 1);
                a12z = get_a(j + 1, 2) + get_a(jp1 + 1, 2);
                AT1 = 0.5 * rs / aT[j];
            }
            double AT2 = 0.5 * rs / aT[jp1];
            double a02x = get_a(0, 0) + get_a(jp1 + 1, 0), a02y = get_a(0, 1) + get_a(jp1 + 1, 1),
                   a02z = get_a(0, 2) + get_a(jp1 + 1, 2);
            double a23x = get_a(jp1 + 1, 0) + get_a(jp2 + 1, 0), a23y = get_a(jp1 + 1, 1) + get_a(jp2 + 1, 1),
                   a23z = get_a(jp1 + 1, 2) + get_a(jp2 + 1, 2);
            double a03x = get_a(0, 0) + get_a(jp2 + 1, 0), a03y = get_a(0, 1) + get_a(jp2 + 1, 1),
                   a03z = get_a(0, 2) + get_a(jp2 + 1, 2);
            if (j == 0) {
                lap1 = (-a01x * ti[j * 3 + 0] - a01y * ti[j * 3 + 1] - a01z * ti[j * 3 + 2] + a12x * te[j * 3 + 0]
                        + a12y * te[j * 3 + 1] + a12z * te[j * 3 + 2] + a02x * ti[jp1 * 3 + 0]
                        + a02y * ti[jp1 * 3 + 1] + a02z * ti[jp1 * 3 + 2])
                       * AT1;
            } else {
                lap1 = lap2;
            }
            lap2 = (-a02x * ti[jp1 * 3 + 0] - a02y * ti[jp1 * 3 + 1] - a02z * ti[jp1 * 3 + 2]
                    + a23x * te[jp1 * 3 + 0] + a23y * te[jp1 * 3 + 1] + a23z * te[jp1 * 3 + 2]
                    + a03x * ti[jp2 * 3 + 0] + a03y * ti[jp2 * 3 + 1] + a03z * ti[jp2 * 3 + 2])
                   * AT2;
            lapx += (lap1 + dwdz + lap2 + dwdz2) * oa[j * 3 + 0] * AT0;
            lapy += (lap1 + dwdz + lap2 + dwdz2) * oa[j * 3 + 1] * AT0;
            lapz += (lap1 + dwdz + lap2 + dwdz2) * oa[j * 3 + 2] * AT0;
        }
        out[0] = lapx; out[1] = lapy; out[2] = lapz;
    };
    auto finish = [&](int id, int lev, double l[3]) {
        if (laststep) {
            double sdiff = o.Kdhz[lev];
            double fx = fr[id * 3 + 0], fy = fr[id * 3 + 1], fz = fr[id * 3 + 2];
            double lapr = fx * l[0] + fy * l[1] + fz * l[2];
            double lx = l[0] - fx * lapr, ly = l[1] - fy * lapr, lz = l[2] - fz * lapr;
            DivM[(id * nv + lev) * 3 + 0] = -lx * sdiff;
            DivM[(id * nv + lev) * 3 + 1] = -ly * sdiff;
            DivM[(id * nv + lev) * 3 + 2] = -lz * sdiff;
        } else {
            for (int k = 0; k < 3; k++) divg[(id * nv + lev) * 3 + k] = l[k];
        }
    };
    ORACLE_PAR
    for (int ib = 0; ib < o.nr; ib++) {
        double as[3][NHL2], wl[NHL2], wt[NHL2];
        const int *m = o.g.maps + ib * NHL2;
        for (int lev = 0; lev < nv; lev++) {
            for (int pos : kLoad) {
                int gi = m[pos];
                for (int k = 0; k < 3; k++)
                    as[k][pos] = gi < 0 ? 0.0 : (laststep ? divg_in[(gi * nv + lev) * 3 + k] : M[(gi * nv + lev) * 3 + k]);
                wl[pos] = (gi < 0 || laststep) ? 0.0 : Wh[gi * (nv + 1) + lev];
                wt[pos] = (gi < 0 || laststep) ? 0.0 : Wh[gi * (nv + 1) + lev + 1];
            }
            for (int y = 0; y < NL; y++)
                for (int x = 0; x < NL; x++) {
                    Stencil s = stencil(m, x, y);
                    int nfc = s.pent ? 5 : 6;
                    int slot[7];
                    slot[0] = s.ir;
                    for (int j = 0; j < 6; j++) {
                        int p1, p2, p3;
                        tri_pts(x, y, j, s.pent, p1, p2, p3);
                        slot[j + 1] = p1;
                    }
                    double l[3];
                    point(s.id, lev, nfc, false, [&](int sl, int k) { return as[k][slot[sl]]; },
                          [&](int sl) { return wl[slot[sl]]; }, [&](int sl) { return wt[slot[sl]]; }, l);
                    finish(s.id, lev, l);
                }
        }
    }
    for (int id = P - 2; id < P; id++)
        for (int lev = 0; lev < nv; lev++) {
            int gslot[6];
            gslot[0] = id;
            for (int i = 0; i < 5; i++) gslot[i + 1] = o.g.point_local[id * 6 + i];
            double l[3];
            point(id, lev, 5, true,
                  [&](int sl, int k) {
                      int gi = gslot[sl];
                      return laststep ? divg_in[(gi * nv + lev) * 3 + k] : M[(gi * nv + lev) * 3 + k];
                  },
                  [&](int sl) { return Wh[gslot[sl] * (nv + 1) + lev]; },
                  [&](int sl) { return Wh[gslot[sl] * (nv + 1) + lev + 1]; }, l);
            finish(id, lev, l);
        }
}

// -------------------------------------------------------------- slow modes
// Compute_Slow_Modes + _Poles (thor_slowmodes.h:48-683), entropy form, with
// the vertical-diffusion terms diffmv/diffwv/diffrv/diffprv (zero unless
// VertHyDiff).
void slow_modes(Oracle &o) {
    const int nv = o.nv, P = o.P;
    double *SlowMh = o.A_("SlowMh"), *SlowWh = o.A_("SlowWh"), *SlowRho = o.A_("SlowRho"),
           *Slowp = o.A_("Slowpressure"), *Mh = o.A_("Mhk"), *Wh = o.A_("Whk"), *Rho = o.A_("Rhok"),
           *Adv = o.A_("Adv"), *DivM = o.A_("DivM"), *diffmh = o.A_("diffmh"), *diffw = o.A_("diffw"),
           *diffrh = o.A_("diffrh"), *diffpr = o.A_("diffpr"), *pk = o.A_("pressurek"), *h_d = o.A_("h"),
           *hh = o.A_("hh"), *gtil = o.A_("gtil"), *Cp = o.A_("Cp"), *Rd = o.A_("Rd"),
           *dMh = o.A_("profx_dMh"), *dWh = o.A_("profx_dWh"), *Qh = o.A_("profx_Qheat");
    const double *diffmv = o.A_("diffmv"), *diffwv = o.A_("diffwv"), *diffrv = o.A_("diffrv"),
                 *diffprv = o.A_("diffprv");
    const double *fr = o.g.func_r, *Alt = o.g.Altitude, *Alth = o.g.Altitudeh, *div = o.g.div, *grad = o.g.grad;
    const double A = o.p.A, Gravit = o.p.Gravit;

    // Per-point body shared by tiles and poles; nb(k) gives the global index of
    // stencil slot k (0 = self). pole selects the pole variant of dw/dz.
    auto body = [&](int id, int lev, int nslots, bool pole, auto nb) {
        double Cv = Cp[id * nv + lev] - Rd[id * nv + lev];
        double fx = fr[id * 3 + 0], fy = fr[id * 3 + 1], fz = fr[id * 3 + 2];
        double advx = Adv[(id * nv + lev) * 3 + 0], advy = Adv[(id * nv + lev) * 3 + 1],
               advz = Adv[(id * nv + lev) * 3 + 2];
        double rho = Rho[id * nv + lev];
        double advrl = 0, rhol = 0, altl = 0, pressurel = 0;
        if (lev != 0) {
            double ax = Adv[(id * nv + lev - 1) * 3 + 0], ay = Adv[(id * nv + lev - 1) * 3 + 1],
                   az = Adv[(id * nv + lev - 1) * 3 + 2];
            advrl = ax * fx + ay * fy + az * fz;
            rhol = Rho[id * nv + lev - 1];
            altl = Alt[lev - 1];
            pressurel = pk[id * nv + lev - 1];
        }
        double advr = advx * fx + advy * fy + advz * fz;
        advx += -advr * fx;
        advy += -advr * fy;
        advz += -advr * fz;
        double altht = Alth[lev + 1], althl = Alth[lev], alt = Alt[lev];
        if (!o.p.NonHydro) {
            const double *M0 = Mh + (id * nv + lev) * 3;
            advr = (MPOW(M0[0], 2) + MPOW(M0[1], 2) + MPOW(M0[2], 2)) / (rho * (alt + A));
            if (lev != 0) {
                const double *M1 = Mh + (id * nv + lev - 1) * 3;
                advrl = (MPOW(M1[0], 2) + MPOW(M1[1], 2) + MPOW(M1[2], 2)) / (rhol * (altl + A));
            }
        }
        double dz = altht - althl, r2p, r2m, r2l, rscale, dr2dz;
        if (o.p.DeepModel) {
            r2p = MPOW(altht + A, 2.0);
            r2m = MPOW(alt + A, 2.0);
            r2l = MPOW(althl + A, 2.0);
            rscale = A / (alt + A);
            dr2dz = (MPOW(altht + A, 3.0) - MPOW(althl + A, 3.0)) / 3;
        } else {
            r2p = r2m = r2l = 1.0;
            rscale = 1.0;
            dr2dz = r2m * dz;
        }
        double nflxv[3], nflxr = 0.0, nflxp = 0.0;
        for (int k = 0; k < 3; k++) {
            double sg = grad[id * 21 + 0 * 3 + k] * pk[nb(0) * nv + lev];
            double sr = div[id * 21 + 0 * 3 + k] * Mh[(nb(0) * nv + lev) * 3 + k];
            double sp = div[id * 21 + 0 * 3 + k] * Mh[(nb(0) * nv + lev) * 3 + k] * h_d[nb(0) * nv + lev];
            for (int j = 1; j < nslots; j++) {
                int gi = nb(j);
                sg = sg + grad[id * 21 + j * 3 + k] * pk[gi * nv + lev];
                sr = sr + div[id * 21 + j * 3 + k] * Mh[(gi * nv + lev) * 3 + k];
                sp = sp + div[id * 21 + j * 3 + k] * Mh[(gi * nv + lev) * 3 + k] * h_d[gi * nv + lev];
            }
            nflxv[k] = rscale * sg;
            nflxr += rscale * sr;
            nflxp += rscale * sp;
        }
        double whl, wht, hhl, hht;
        if (lev == 0) {
            whl = 0.0; wht = Wh[id * (nv + 1) + lev + 1]; hhl = 0.0; hht = hh[id * (nv + 1) + lev + 1];
        } else {
            whl = Wh[id * (nv + 1) + lev]; wht = Wh[id * (nv + 1) + lev + 1];
            hhl = hh[id * (nv + 1) + lev]; hht = hh[id * (nv + 1) + lev + 1];
        }
        double dwdz = pole ? (wht * r2p - whl * r2l) / (dz * r2m) : (wht * r2p - whl * r2l) / (dr2dz);
        double dwhdz = (wht * r2p * hht - whl * r2l * hhl) / (dz * r2m);
        double dpr = nflxv[0] * fx + nflxv[1] * fy + nflxv[2] * fz;
        const double fv[3] = {fx, fy, fz}, ad[3] = {advx, advy, advz};
        for (int k = 0; k < 3; k++) {
            int e = (id * nv + lev) * 3 + k;
            SlowMh[e] = -(nflxv[k] - dpr * fv[k]) - ad[k] + DivM[e] + diffmh[e] + diffmv[e] + dMh[e];
        }
        if (lev == 0) {
            SlowWh[id * (nv + 1) + 0] = 0.0;
            SlowWh[id * (nv + 1) + nv] = 0.0;
        } else {
            double xi = althl, xim = altl, xip = alt;
            double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            double swr;
            if (o.p.DeepModel) {
                if (o.p.GravHeightVar)
                    swr = (advrl + rhol * Gravit * MPOW(A / (A + altl), 2)) * intt
                          + (advr + rho * Gravit * MPOW(A / (A + alt), 2)) * intl;
                else
                    swr = (advrl + rhol * Gravit) * intt + (advr + rho * Gravit) * intl;
            } else {
                if (o.p.GravHeightVar)
                    swr = (rhol * Gravit * MPOW(A / (A + altl), 2)) * intt
                          + (rho * Gravit * MPOW(A / (A + alt), 2)) * intl;
                else
                    swr = (rhol * Gravit) * intt + (rho * Gravit) * intl;
            }
            if (o.p.NonHydro)
                swr += (-diffw[id * nv + lev - 1] - diffwv[id * nv + lev - 1]) * intt
                       + (-diffw[id * nv + lev] - diffwv[id * nv + lev]) * intl;
            double dzl = alt - altl;
            double dpdz = (pk[id * nv + lev] - pressurel) / dzl;
            swr += dpdz;
            swr -= dWh[id * (nv + 1) + lev];
            SlowWh[id * (nv + 1) + lev] = -swr;
        }
        nflxr += dwdz;
        SlowRho[id * nv + lev] = -nflxr + diffrh[id * nv + lev] + diffrv[id * nv + lev];
        double r = 1.0 / rho;
        const double *M0 = Mh + (id * nv + lev) * 3;
        double vgp = (nflxv[0] - dpr * fx) * M0[0] * r + (nflxv[1] - dpr * fy) * M0[1] * r
                     + (nflxv[2] - dpr * fz) * M0[2] * r;
        Slowp[id * nv + lev] = (Rd[id * nv + lev] / Cv)
                               * (-nflxp - gtil[id * nv + lev] - dwhdz + vgp + Qh[id * nv + lev]);
        if (o.p.energy_equation)  // thor_slowmodes.h:375-381
            Slowp[id * nv + lev] += Rd[id * nv + lev] / Cv * (diffpr[id * nv + lev] + diffprv[id * nv + lev]);
        else
            Slowp[id * nv + lev] += diffpr[id * nv + lev] + diffprv[id * nv + lev];
    };
    ORACLE_PAR
    for (int ib = 0; ib < o.nr; ib++) {
        const int *m = o.g.maps + ib * NHL2;
        for (int lev = 0; lev < nv; lev++)
            for (int y = 0; y < NL; y++)
                for (int x = 0; x < NL; x++) {
                    Stencil s = stencil(m, x, y);
                    // A missing pentagon corner reads zeros through the halo; its
                    // coefficient div6/grad6 is zero and pt5 is remapped, so the
                    // -1 entry is never dereferenced.
                    body(s.id, lev, 7, false, [&](int k) { return m[s.pt[k]]; });
                }
    }
    for (int id = P - 2; id < P; id++) {
        int gslot[6];
        gslot[0] = id;
        for (int i = 0; i < 5; i++) gslot[i + 1] = o.g.point_local[id * 6 + i];
        for (int lev = 0; lev < nv; lev++) body(id, lev, 6, true, [&](int k) { return gslot[k]; });
    }
}

// ------------------------------------------------------- RK bookkeeping
// UpdateRK (thor_auxiliary.h:311-418): deviations = y(n) - y(k).
void update_rk(Oracle &o) {
    const int nv = o.nv;
    double *Ms = o.A_("Mhs"), *Mk = o.A_("Mhk"), *Mi = o.A_("Mh"), *Whs = o.A_("Whs"), *Whk = o.A_("Whk"),
           *Whi = o.A_("Wh"), *Ws = o.A_("Ws"), *Rs = o.A_("Rhos"), *Rk = o.A_("Rhok"), *Ri = o.A_("Rho"),
           *ps = o.A_("pressures"), *pk = o.A_("pressurek"), *pi = o.A_("pressure");
    const double *fr = o.g.func_r, *Alt = o.g.Altitude, *Alth = o.g.Altitudeh;
    ORACLE_PAR
    for (int id = 0; id < o.P; id++) {
        double fx = fr[id * 3 + 0], fy = fr[id * 3 + 1], fz = fr[id * 3 + 2];
        for (int lev = 0; lev < nv; lev++) {
            int e = (id * nv + lev) * 3;
            double Mx = Mi[e + 0] - Mk[e + 0], My = Mi[e + 1] - Mk[e + 1], Mz = Mi[e + 2] - Mk[e + 2];
            double aux = Mx * fx + My * fy + Mz * fz;
            Ms[e + 0] = Mx - aux * fx;
            Ms[e + 1] = My - aux * fy;
            Ms[e + 2] = Mz - aux * fz;
            Rs[id * nv + lev] = Ri[id * nv + lev] - Rk[id * nv + lev];
            ps[id * nv + lev] = pi[id * nv + lev] - pk[id * nv + lev];
        }
        double altht = 0, althl, wht = 0, whl;
        for (int lev = 0; lev < nv + 1; lev++) {
            if (lev == 0) {
                Whs[id * (nv + 1) + 0] = 0.0;
                altht = Alth[lev];
                wht = 0.0;
            } else {
                whl = wht;
                if (lev == nv) wht = 0.0;
                else wht = Whi[id * (nv + 1) + lev] - Whk[id * (nv + 1) + lev];
                Whs[id * (nv + 1) + lev] = wht;
                althl = altht;
                altht = Alth[lev];
                double xi = Alt[lev - 1], xim = althl, xip = altht;
                double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
                Ws[id * nv + lev - 1] = whl * intt + wht * intl;
            }
        }
    }
}

// UpdateRK2 (thor_auxiliary.h:420-529): y(k) += deviation.
void update_rk2(Oracle &o) {
    const int nv = o.nv;
    double *Ms = o.A_("Mhs"), *Mk = o.A_("Mhk"), *Whs = o.A_("Whs"), *Whk = o.A_("Whk"), *Wk = o.A_("Wk"),
           *Rs = o.A_("Rhos"), *Rk = o.A_("Rhok"), *ps = o.A_("pressures"), *pk = o.A_("pressurek");
    const double *fr = o.g.func_r, *Alt = o.g.Altitude, *Alth = o.g.Altitudeh;
    ORACLE_PAR
    for (int id = 0; id < o.P; id++) {
        double fx = fr[id * 3 + 0], fy = fr[id * 3 + 1], fz = fr[id * 3 + 2];
        for (int lev = 0; lev < nv; lev++) {
            int e = (id * nv + lev) * 3;
            double Mx = Ms[e + 0] + Mk[e + 0], My = Ms[e + 1] + Mk[e + 1], Mz = Ms[e + 2] + Mk[e + 2];
            double aux = Mx * fx + My * fy + Mz * fz;
            Mk[e + 0] = Mx - aux * fx;
            Mk[e + 1] = My - aux * fy;
            Mk[e + 2] = Mz - aux * fz;
            Rk[id * nv + lev] += Rs[id * nv + lev];
            pk[id * nv + lev] += ps[id * nv + lev];
        }
        for (int lev = 0; lev < nv + 1; lev++) {
            if (lev == 0 || lev == nv) Whk[id * (nv + 1) + lev] = 0.0;
            else Whk[id * (nv + 1) + lev] += Whs[id * (nv + 1) + lev];
        }
        double althl = 0, altht = 0, whl = 0, wht = 0;
        for (int lev = 0; lev < nv; lev++) {
            if (lev == 0) {
                althl = 0;
                altht = Alth[1];
                whl = 0.0;
                wht = Whk[id * (nv + 1) + 1];
            }
            double xi = Alt[lev], xim = althl, xip = altht;
            double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            Wk[id * nv + lev] = whl * intt + wht * intl;
            if (lev < nv - 1) {
                althl = altht;
                altht = Alth[lev + 2];
                whl = wht;
                wht = Whk[id * (nv + 1) + lev + 2];
            }
        }
    }
}

// ------------------------------------------------------------ fast modes
// Momentum_Eq + _Poles (thor_fastmodes.h:54-253).
void momentum_eq(Oracle &o, double dt) {
    const int nv = o.nv, P = o.P;
    double *M = o.A_("Mhs"), *ps = o.A_("pressures"), *SlowMh = o.A_("SlowMh"), *DivM = o.A_("DivM");
    const double *grad = o.g.grad, *fr = o.g.func_r;
    auto body = [&](int id, int lev, int nslots, auto psv) {
        double rs = rscale_of(o, lev);
        double nf[3];
        for (int k = 0; k < 3; k++) {
            double s = grad[id * 21 + 0 * 3 + k] * psv(0);
            for (int j = 1; j < nslots; j++) s = s + grad[id * 21 + j * 3 + k] * psv(j);
            nf[k] = rs * s;
        }
        int e = (id * nv + lev) * 3;
        double Mx = (-nf[0] + SlowMh[e + 0] + DivM[e + 0]) * dt;
        double My = (-nf[1] + SlowMh[e + 1] + DivM[e + 1]) * dt;
        double Mz = (-nf[2] + SlowMh[e + 2] + DivM[e + 2]) * dt;
        double fx = fr[id * 3 + 0], fy = fr[id * 3 + 1], fz = fr[id * 3 + 2];
        double vr = Mx * fx + My * fy + Mz * fz;
        Mx += -vr * fx;
        My += -vr * fy;
        Mz += -vr * fz;
        M[e + 0] += Mx;
        M[e + 1] += My;
        M[e + 2] += Mz;
    };
    ORACLE_PAR
    for (int ib = 0; ib < o.nr; ib++) {
        double sh[NHL2];
        const int *m = o.g.maps + ib * NHL2;
        for (int lev = 0; lev < nv; lev++) {
            fill(sh, m, [&](int gi) { return ps[gi * nv + lev]; });
            for (int y = 0; y < NL; y++)
                for (int x = 0; x < NL; x++) {
                    Stencil s = stencil(m, x, y);
                    body(s.id, lev, 7, [&](int k) { return sh[s.pt[k]]; });
                }
        }
    }
    for (int id = P - 2; id < P; id++) {
        int gslot[6];
        gslot[0] = id;
        for (int i = 0; i < 5; i++) gslot[i + 1] = o.g.point_local[id * 6 + i];
        for (int lev = 0; lev < nv; lev++) body(id, lev, 6, [&](int k) { return ps[gslot[k] * nv + lev]; });
    }
}

// Prepare_Implicit_Vertical + _Poles (thor_vertical_int.h:566-771).
void prepare_implicit_vertical(Oracle &o) {
    const int nv = o.nv, P = o.P;
    double *M = o.A_("Mhs"), *h_d = o.A_("h"), *Slowp = o.A_("Slowpressure"), *SlowRho = o.A_("SlowRho"),
           *Sp = o.A_("Sp"), *Sd = o.A_("Sd"), *Cp = o.A_("Cp"), *Rd = o.A_("Rd");
    const double *div = o.g.div;
    auto body = [&](int id, int lev, int nslots, auto Mv, auto hv) {
        double RoC = Rd[id * nv + lev] / (Cp[id * nv + lev] - Rd[id * nv + lev]);
        double rs = rscale_of(o, lev);
        double nflxr = 0.0, nflxp = 0.0;
        for (int k = 0; k < 3; k++) {
            double sr = div[id * 21 + 0 * 3 + k] * Mv(0, k);
            double sp = div[id * 21 + 0 * 3 + k] * Mv(0, k) * hv(0);
            for (int j = 1; j < nslots; j++) {
                sr = sr + div[id * 21 + j * 3 + k] * Mv(j, k);
                sp = sp + div[id * 21 + j * 3 + k] * Mv(j, k) * hv(j);
            }
            nflxr -= rs * sr;
            nflxp -= (RoC)*rs * sp;
        }
        Sp[id * nv + lev] = nflxp + Slowp[id * nv + lev];
        Sd[id * nv + lev] = nflxr + SlowRho[id * nv + lev];
    };
    ORACLE_PAR
    for (int ib = 0; ib < o.nr; ib++) {
        double vs[3][NHL2], hs[NHL2];
        const int *m = o.g.maps + ib * NHL2;
        for (int lev = 0; lev < nv; lev++) {
            for (int k = 0; k < 3; k++) fill(vs[k], m, [&](int gi) { return M[(gi * nv + lev) * 3 + k]; });
            fill(hs, m, [&](int gi) { return h_d[gi * nv + lev]; });
            for (int y = 0; y < NL; y++)
                for (int x = 0; x < NL; x++) {
                    Stencil s = stencil(m, x, y);
                    body(s.id, lev, 7, [&](int j, int k) { return vs[k][s.pt[j]]; },
                         [&](int j) { return hs[s.pt[j]]; });
                }
        }
    }
    for (int id = P - 2; id < P; id++) {
        int gslot[6];
        gslot[0] = id;
        for (int i = 0; i < 5; i++) gslot[i + 1] = o.g.point_local[id * 6 + i];
        for (int lev = 0; lev < nv; lev++)
            body(id, lev, 6, [&](int j, int k) { return M[(gslot[j] * nv + lev) * 3 + k]; },
                 [&](int j) { return h_d[gslot[j] * nv + lev]; });
    }
}

// Vertical_Eq (thor_vertical_int.h:52-563): per-column tridiagonal solve for
// Whs at interfaces 1..nv-1, then Ws at layer centres.
void vertical_eq(Oracle &o, double deltat) {
    const int nv = o.nv, nvi = nv + 1;
    double *Whs = o.A_("Whs"), *Ws = o.A_("Ws"), *ps = o.A_("pressures"), *hh = o.A_("hh"),
           *Rhos = o.A_("Rhos"), *gtilh = o.A_("gtilh"), *Sp_d = o.A_("Sp"), *Sd_d = o.A_("Sd"),
           *Srh = o.A_("SlowWh"), *Cp = o.A_("Cp"), *Rd = o.A_("Rd");
    const double *Alt = o.g.Altitude, *Alth = o.g.Altitudeh;
    const double A = o.p.A, Gravit = o.p.Gravit;
    const bool deep = o.p.DeepModel, ghv = o.p.GravHeightVar, nonh = o.p.NonHydro;
    double *ta = nullptr, *tb = nullptr, *tc = nullptr, *td = nullptr;
    if (o.capture_tri) {
        for (const char *n : {"tri_a", "tri_b", "tri_c", "tri_d"}) o.arr[n].assign((size_t)o.P * nvi, 0.0);
        ta = o.A_("tri_a"), tb = o.A_("tri_b"), tc = o.A_("tri_c"), td = o.A_("tri_d");
    }
    ORACLE_PAR
    for (int id = 0; id < o.P; id++) {
        std::vector<double> cc(nvi), dd(nvi);
        double altht = 0, althl = 0, alth = 0, alt = 0, altl = 0, hp = 0, h = 0, hm = 0, gp = 0, g = 0, gm = 0;
        double Sp = 0, Spl = 0, p = 0, pl = 0, r = 0, rl = 0, Sdl = 0, Sd = 0;
        double CRddl = 0, CRddu = 0, GCoRl = 0, GCoRu = 0;
        for (int lev = 1; lev < nv; lev++) {
            if (lev == 1) {
                altht = Alth[lev + 1]; althl = Alth[lev - 1]; alth = Alth[lev];
                alt = Alt[lev]; altl = Alt[lev - 1];
                hp = hh[id * nvi + lev + 1]; h = hh[id * nvi + lev]; hm = hh[id * nvi + lev - 1];
                gp = gtilh[id * nvi + lev + 1]; g = gtilh[id * nvi + lev]; gm = gtilh[id * nvi + lev - 1];
                Sp = Sp_d[id * nv + lev]; Spl = Sp_d[id * nv + lev - 1];
                p = ps[id * nv + lev]; pl = ps[id * nv + lev - 1];
                rl = Rhos[id * nv + lev - 1]; r = Rhos[id * nv + lev];
                Sdl = Sd_d[id * nv + lev - 1]; Sd = Sd_d[id * nv + lev];
                CRddl = (Cp[id * nv + lev - 1] - Rd[id * nv + lev - 1]) / (Rd[id * nv + lev - 1] * deltat * deltat);
                CRddu = (Cp[id * nv + lev] - Rd[id * nv + lev]) / (Rd[id * nv + lev] * deltat * deltat);
                if (ghv) {
                    GCoRl = Gravit * MPOW(A / (A + altl), 2) * (Cp[id * nv + lev - 1] - Rd[id * nv + lev - 1])
                            / Rd[id * nv + lev - 1];
                    GCoRu = Gravit * MPOW(A / (A + alt), 2) * (Cp[id * nv + lev] - Rd[id * nv + lev])
                            / Rd[id * nv + lev];
                } else {
                    GCoRl = Gravit * (Cp[id * nv + lev - 1] - Rd[id * nv + lev - 1]) / Rd[id * nv + lev - 1];
                    GCoRu = Gravit * (Cp[id * nv + lev] - Rd[id * nv + lev]) / Rd[id * nv + lev];
                }
            }
            double aa, bb, C0, intt, intl, inttm, xi, xim, xip;
            double dzp = 1.0 / (altht - alth), dzh = 1.0 / (alt - altl), dzm = 1.0 / (alth - althl);
            double dzph = dzp * dzh, dzmh = dzm * dzh;
            if (deep) {
                hp = hp * MPOW(altht + A, 2);
                h = h * MPOW(alth + A, 2);
                hm = hm * MPOW(althl + A, 2);
                double or2 = 1 / MPOW(alth + A, 2);
                double tor3 = 2 / MPOW(alth + A, 3);
                xi = alt; xim = alth; xip = altht;
                intl = (xi - xim) * dzp * dzh;
                xi = altl; xim = althl; xip = alth;
                inttm = -(xi - xip) * dzm * dzh;
                double CRdd = (CRddl * (alt - alth) + CRddu * (alth - altl)) / (alt - altl);
                double GCoR = (GCoRl * (alt - alth) + GCoRu * (alth - altl)) / (alt - altl);
                cc[lev] = -dzph * or2 * hp - inttm * (gp + GCoR - tor3 * hp);
                if (nonh)
                    bb = CRdd + (dzph + dzmh) * or2 * h + (intl - inttm) * (g + GCoR - tor3 * h);
                else
                    bb = (dzph + dzmh) * or2 * h + (intl - inttm) * (g + GCoR - tor3 * h);
                aa = -dzmh * or2 * hm + intl * (gm + GCoR - tor3 * hm);
                double dSpdz = (Sp - Spl) * dzh, dPdz = (p - pl) * dzh;
                xi = alth; xim = altl; xip = alt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                double rhohs = rl * intt + r * intl;
                double Sdh = Sdl * intt + Sd * intl;
                double gh = ghv ? Gravit * MPOW(A / (A + alt), 2) : Gravit;
                if (!nonh) {
                    if (ghv)
                        C0 = (MPOW(deltat, 2.0) * dSpdz + MPOW(deltat, 2.0) * Gravit * MPOW(A / (A + alt), 2) * Sdh
                              + Gravit * MPOW(A / (A + alt), 2) * deltat * rhohs + deltat * dPdz
                              - deltat * Srh[id * nvi + lev]) * (CRdd);
                    else
                        C0 = (MPOW(deltat, 2.0) * dSpdz + MPOW(deltat, 2.0) * Gravit * Sdh
                              + Gravit * deltat * rhohs + deltat * dPdz - deltat * Srh[id * nvi + lev]) * (CRdd);
                } else {
                    if (ghv)
                        C0 = (-Whs[id * nvi + lev] + MPOW(deltat, 2.0) * dSpdz
                              + MPOW(deltat, 2.0) * Gravit * MPOW(A / (A + alt), 2) * Sdh
                              + Gravit * MPOW(A / (A + alt), 2) * deltat * rhohs + deltat * dPdz
                              - deltat * Srh[id * nvi + lev]) * (CRdd);
                    else
                        C0 = (-Whs[id * nvi + lev] + MPOW(deltat, 2.0) * dSpdz
                              + MPOW(deltat, 2.0) * Gravit * Sdh + Gravit * deltat * rhohs + deltat * dPdz
                              - deltat * Srh[id * nvi + lev]) * (CRdd);
                }
                (void)gh;
                if (lev < nv - 1) {
                    althl = alth; alth = altht; altht = Alth[lev + 2];
                    altl = alt; alt = Alt[lev + 1];
                    hm = hh[id * nvi + lev]; h = hh[id * nvi + lev + 1]; hp = hh[id * nvi + lev + 2];
                    gm = g; g = gp; gp = gtilh[id * nvi + lev + 2];
                    Spl = Sp; Sp = Sp_d[id * nv + lev + 1];
                    pl = p; p = ps[id * nv + lev + 1];
                    rl = r; r = Rhos[id * nv + lev + 1];
                    Sdl = Sd; Sd = Sd_d[id * nv + lev + 1];
                    CRddl = CRddu;
                    CRddu = (Cp[id * nv + lev + 1] - Rd[id * nv + lev + 1]) / (Rd[id * nv + lev + 1] * deltat * deltat);
                    GCoRl = GCoRu;
                    if (ghv)
                        GCoRu = Gravit * MPOW(A / (A + alt), 2) * (Cp[id * nv + lev + 1] - Rd[id * nv + lev + 1])
                                / Rd[id * nv + lev + 1];
                    else
                        GCoRu = Gravit * (Cp[id * nv + lev + 1] - Rd[id * nv + lev + 1]) / Rd[id * nv + lev + 1];
                }
            } else {
                xi = alt; xim = alth; xip = altht;
                intl = (xi - xim) * dzp * dzh;
                xi = altl; xim = althl; xip = alth;
                inttm = -(xi - xip) * dzm * dzh;
                double CRdd = (CRddl * (alt - alth) + CRddu * (alth - altl)) / (alt - altl);
                double GCoR = (GCoRl * (alt - alth) + GCoRu * (alth - altl)) / (alt - altl);
                cc[lev] = -dzph * hp - inttm * (gp + GCoR);
                if (nonh) bb = CRdd + (dzph + dzmh) * h + (intl - inttm) * (g + GCoR);
                else bb = (dzph + dzmh) * h + (intl - inttm) * (g + GCoR);
                aa = -dzmh * hm + intl * (gm + GCoR);
                double dSpdz = (Sp - Spl) * dzh, dPdz = (p - pl) * dzh;
                xi = alth; xim = altl; xip = alt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                double rhohs = rl * intt + r * intl;
                double Sdh = Sdl * intt + Sd * intl;
                if (!nonh) {
                    if (ghv)
                        C0 = (MPOW(deltat, 2.0) * dSpdz + MPOW(deltat, 2.0) * Gravit * MPOW(A / (A + alt), 2) * Sdh
                              + Gravit * MPOW(A / (A + alt), 2) * deltat * rhohs + deltat * dPdz
                              - deltat * Srh[id * nvi + lev]) * (CRdd);
                    else
                        C0 = (MPOW(deltat, 2.0) * dSpdz + MPOW(deltat, 2.0) * Gravit * Sdh
                              + Gravit * deltat * rhohs + deltat * dPdz - deltat * Srh[id * nvi + lev]) * (CRdd);
                } else {
                    if (ghv)
                        C0 = (-Whs[id * nvi + lev] + MPOW(deltat, 2.0) * dSpdz
                              + MPOW(deltat, 2.0) * Gravit * MPOW(A / (A + alt), 2) * Sdh
                              + Gravit * MPOW(A / (A + alt), 2) * deltat * rhohs + deltat * dPdz
                              - deltat * Srh[id * nvi + lev]) * (CRdd);
                    else
                        C0 = (-Whs[id * nvi + lev] + MPOW(deltat, 2.0) * dSpdz
                              + MPOW(deltat, 2.0) * Gravit * Sdh + Gravit * deltat * rhohs + deltat * dPdz
                              - deltat * Srh[id * nvi + lev]) * (CRdd);
                }
                if (lev < nv - 1) {
                    althl = alth; alth = altht; altht = Alth[lev + 2];
                    altl = alt; alt = Alt[lev + 1];
                    hm = h; h = hp; hp = hh[id * nvi + lev + 2];
                    gm = g; g = gp; gp = gtilh[id * nvi + lev + 2];
                    Spl = Sp; Sp = Sp_d[id * nv + lev + 1];
                    pl = p; p = ps[id * nv + lev + 1];
                    rl = r; r = Rhos[id * nv + lev + 1];
                    Sdl = Sd; Sd = Sd_d[id * nv + lev + 1];
                    CRddl = CRddu;
                    CRddu = (Cp[id * nv + lev + 1] - Rd[id * nv + lev + 1]) / (Rd[id * nv + lev + 1] * deltat * deltat);
                    GCoRl = GCoRu;
                    if (ghv)
                        GCoRu = Gravit * MPOW(A / (A + alt), 2) * (Cp[id * nv + lev + 1] - Rd[id * nv + lev + 1])
                                / Rd[id * nv + lev + 1];
                    else
                        GCoRu = Gravit * (Cp[id * nv + lev + 1] - Rd[id * nv + lev + 1]) / Rd[id * nv + lev + 1];
                }
            }
            dd[lev] = -C0;
            if (ta) {
                ta[id * nvi + lev] = aa;
                tb[id * nvi + lev] = bb;
                tc[id * nvi + lev] = cc[lev];
                td[id * nvi + lev] = -C0;
            }
            if (lev == 1) {
                cc[1] = cc[1] / bb;
                dd[1] = dd[1] / bb;
            } else {
                double t = 1.0 / (bb - cc[lev - 1] * aa);
                cc[lev] *= t;
                dd[lev] = (dd[lev] - dd[lev - 1] * aa) * t;
            }
        }
        Whs[id * nvi + nv] = 0.0;
        Whs[id * nvi] = 0.0;
        Whs[id * nvi + nv - 1] = dd[nv - 1];
        for (int lev = nvi - 2; lev > 0; lev--)
            Whs[id * nvi + lev] = (-cc[lev] * Whs[id * nvi + lev + 1] + dd[lev]);
        double althl2 = 0, altht2 = 0, whl = 0, wht = 0;
        for (int lev = 0; lev < nv; lev++) {
            if (lev == 0) {
                althl2 = 0; altht2 = Alth[1]; whl = 0.0; wht = Whs[id * nvi + 1];
            }
            double xi = Alt[lev], xim = althl2, xip = altht2;
            double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            Ws[id * nv + lev] = whl * intt + wht * intl;
            if (lev < nv - 1) {
                althl2 = altht2; altht2 = Alth[lev + 2]; whl = wht; wht = Whs[id * nvi + lev + 2];
            }
        }
    }
}

// Density_Pressure_Eqs + _Poles (thor_fastmodes.h:255-847), entropy and
// total-energy forms. The energy form carries h + ep + ek as the flux
// variable (centres: thor_fastmodes.h:352-372, interfaces: 459-476) and
// recovers the pressure from the total energy (501-533).
void density_pressure(Oracle &o, double dt) {
    const int nv = o.nv, P = o.P;
    double *ps = o.A_("pressures"), *pk = o.A_("pressurek"), *Rs = o.A_("Rhos"), *Rk = o.A_("Rhok"),
           *Ms = o.A_("Mhs"), *Mk = o.A_("Mhk"), *Whs = o.A_("Whs"), *Whk = o.A_("Whk"), *pt = o.A_("pt"),
           *pth = o.A_("pth"), *SlowRho = o.A_("SlowRho"), *Qh = o.A_("profx_Qheat"), *diffpr = o.A_("diffpr"),
           *diffprv = o.A_("diffprv"), *Cp = o.A_("Cp"), *Rd = o.A_("Rd");
    const bool energy = o.p.energy_equation, ghv = o.p.GravHeightVar;
    double *h = o.A_("h"), *hh = o.A_("hh"), *ep = o.A_("epotential"), *ek = o.A_("ekinetic"),
           *eph = o.A_("epotentialh"), *ekh = o.A_("ekinetich"), *Et = o.A_("Etotal");
    const double *div = o.g.div, *Alt = o.g.Altitude, *Alth = o.g.Altitudeh;
    const double A = o.p.A, P_Ref = o.p.P_Ref, Gravit = o.p.Gravit;
    auto flux_c = [&](long e) { return energy ? (h[e] + ep[e] + ek[e]) : pt[e]; };
    auto flux_h = [&](long e) { return energy ? (hh[e] + eph[e] + ekh[e]) : pth[e]; };
    auto body = [&](int id, int lev, int nslots, bool pole, auto vv, auto v1v, auto ptv) {
        double Cv = Cp[id * nv + lev] - Rd[id * nv + lev];
        double altht = Alth[lev + 1], althl = Alth[lev], dz = altht - althl;
        double r2p, r2m, r2l, rscale, dr2dz;
        if (o.p.DeepModel) {
            double alt = Alt[lev];
            r2p = MPOW(altht + A, 2.0);
            r2m = MPOW(alt + A, 2.0);
            r2l = MPOW(althl + A, 2.0);
            rscale = A / (alt + A);
            dr2dz = (MPOW(altht + A, 3.0) - MPOW(althl + A, 3.0)) / 3.0;
        } else {
            r2p = r2m = r2l = 1.0;
            rscale = 1.0;
            dr2dz = r2m * dz;
        }
        double nflxr = 0.0, nflxpt = 0.0;
        for (int k = 0; k < 3; k++) {
            double sr = div[id * 21 + 0 * 3 + k] * vv(0, k);
            double sp = div[id * 21 + 0 * 3 + k] * v1v(0, k) * ptv(0);
            for (int j = 1; j < nslots; j++) {
                sr = sr + div[id * 21 + j * 3 + k] * vv(j, k);
                sp = sp + div[id * 21 + j * 3 + k] * v1v(j, k) * ptv(j);
            }
            nflxr += rscale * sr;
            nflxpt += rscale * sp;
        }
        double whl, wht, whl2, wht2, phl, pht;
        if (lev == 0) {
            whl = 0.0;
            wht = Whs[id * (nv + 1) + lev + 1];
            whl2 = 0.0;
            wht2 = Whs[id * (nv + 1) + lev + 1] + Whk[id * (nv + 1) + lev + 1];
            phl = 0.0;
            pht = flux_h(id * (nv + 1) + lev + 1);
        } else {
            whl = Whs[id * (nv + 1) + lev];
            wht = Whs[id * (nv + 1) + lev + 1];
            whl2 = Whs[id * (nv + 1) + lev] + Whk[id * (nv + 1) + lev];
            wht2 = Whs[id * (nv + 1) + lev + 1] + Whk[id * (nv + 1) + lev + 1];
            phl = flux_h(id * (nv + 1) + lev);
            pht = flux_h(id * (nv + 1) + lev + 1);
        }
        double dwdz = pole ? (wht * r2p - whl * r2l) / (dz * r2m) : (wht * r2p - whl * r2l) / (dr2dz);
        double dwptdz = (wht2 * pht * r2p - whl2 * phl * r2l) / (dz * r2m);
        double aux = -(nflxpt + dwptdz) * dt;
        double r = Rk[id * nv + lev] + Rs[id * nv + lev];
        nflxr += dwdz;
        Rs[id * nv + lev] += (SlowRho[id * nv + lev] - nflxr) * dt;
        if (energy) {
            const long e = (long)id * nv + lev;
            aux += Et[e] + (diffpr[e] + diffprv[e]) * dt + Qh[e] * dt;
            Et[e] = aux;
            double w = (whl2 * (Alth[lev + 1] - Alt[lev]) + wht2 * (Alt[lev] - Alth[lev])) / (Alth[lev + 1] - Alth[lev]);
            double rn = Rs[e] + Rk[e];
            double epot = ghv ? rn * Gravit * MPOW(A / (A + Alt[lev]), 2) * Alt[lev] : rn * Gravit * Alt[lev];
            double v0 = v1v(0, 0), v1 = v1v(0, 1), v2 = v1v(0, 2);
            double epk = epot + 0.5 * (MPOW(v0, 2) + MPOW(v1, 2) + MPOW(v2, 2) + MPOW(w, 2)) / rn;
            ps[e] = Rd[e] / Cv * (aux - epk) - pk[e];
            return;
        }
        double ptc = (P_Ref / (Rd[id * nv + lev] * r))
                     * MPOW((ps[id * nv + lev] + pk[id * nv + lev]) / P_Ref, Cv / Cp[id * nv + lev]);
        aux += ptc * r;
        double p = P_Ref * MPOW(Rd[id * nv + lev] * aux / P_Ref, Cp[id * nv + lev] / Cv);
        ps[id * nv + lev] = p - pk[id * nv + lev] + (diffpr[id * nv + lev] + diffprv[id * nv + lev]) * dt
                            + Rd[id * nv + lev] / Cv * Qh[id * nv + lev] * dt;
    };
    ORACLE_PAR
    for (int ib = 0; ib < o.nr; ib++) {
        double vs[3][NHL2], v1s[3][NHL2], pts[NHL2];
        const int *m = o.g.maps + ib * NHL2;
        for (int lev = 0; lev < nv; lev++) {
            for (int pos : kLoad) {
                int gi = m[pos];
                for (int k = 0; k < 3; k++) {
                    if (gi >= 0) {
                        vs[k][pos] = Ms[(gi * nv + lev) * 3 + k];
                        v1s[k][pos] = vs[k][pos] + Mk[(gi * nv + lev) * 3 + k];
                    } else {
                        vs[k][pos] = v1s[k][pos] = 0.0;
                    }
                }
                pts[pos] = gi >= 0 ? flux_c((long)gi * nv + lev) : 0.0;
            }
            for (int y = 0; y < NL; y++)
                for (int x = 0; x < NL; x++) {
                    Stencil s = stencil(m, x, y);
                    body(s.id, lev, 7, false, [&](int j, int k) { return vs[k][s.pt[j]]; },
                         [&](int j, int k) { return v1s[k][s.pt[j]]; }, [&](int j) { return pts[s.pt[j]]; });
                }
        }
    }
    for (int id = P - 2; id < P; id++) {
        int gs[6];
        gs[0] = id;
        for (int i = 0; i < 5; i++) gs[i + 1] = o.g.point_local[id * 6 + i];
        for (int lev = 0; lev < nv; lev++)
            body(id, lev, 6, true, [&](int j, int k) { return Ms[(gs[j] * nv + lev) * 3 + k]; },
                 [&](int j, int k) { return Ms[(gs[j] * nv + lev) * 3 + k] + Mk[(gs[j] * nv + lev) * 3 + k]; },
                 [&](int j) { return flux_c((long)gs[j] * nv + lev); });
    }
}

// ---------------------------------------------------------------- tracers
// Tracer_Eq_Diffusion + _Poles (thor_chemistry.h:364-760) on tracerk / Rhok:
// laststep = false writes the pass into "trdiff" (THOR's diff_d), laststep =
// true the tendency into "difftr". The pole kernel computes only the first
// pass (its body sits in the !laststep branch), reads tracer 0 for the pole
// itself, indexes the ring as (k + 1 + 1) % 5 and weights x with nvecoa's z.
void tracer_diffusion(Oracle &o, bool laststep) {
    const int nv = o.nv, P = o.P, ntr = o.ntr;
    double *tr = o.A_("tracerk"), *Rho = o.A_("Rhok"), *dif = o.A_("trdiff"), *difftr = o.A_("difftr");
    const double *aTr = o.g.areasTr, *noa = o.g.nvecoa, *nti = o.g.nvecti, *nte = o.g.nvecte;
    const double o3 = 1.0 / 3.0, o6 = 1.0 / 6.0;
    double as[NHL2], rs_[NHL2];
    for (int var = 0; var < ntr; var++)
        for (int ib = 0; ib < o.nr; ib++) {
            const int *m = o.g.maps + ib * NHL2;
            for (int lev = 0; lev < nv; lev++) {
                fill(rs_, m, [&](int gi) { return Rho[gi * nv + lev]; });
                if (laststep) fill(as, m, [&](int gi) { return dif[(gi * nv + lev) * ntr + var]; });
                else fill(as, m, [&](int gi) { return tr[(gi * nv + lev) * ntr + var] / Rho[gi * nv + lev]; });
                const double rscale = rscale_of(o, lev);
                const double sdiff = laststep ? -o.Kdh4[lev] : 0.0;
                for (int y = 0; y < NL; y++)
                    for (int x = 0; x < NL; x++) {
                        Stencil st = stencil(m, x, y);
                        const int id = st.id, ir = st.ir, n_faces = st.pent ? 5 : 6;
                        double lap = 0.0, lap1 = 0, lap2 = 0, lap3 = 0, lap4, lap5, lap6, AT1 = 0, AT2;
                        double lapx1 = 0, lapx2 = 0, lapy1 = 0, lapy2 = 0, lapz1 = 0, lapz2 = 0, vdiff;
                        for (int j = 0; j < 6; j++) {
                            const int jp1 = (j + 1) % n_faces, jp2 = (j + 2) % n_faces;
                            int pt1, pt2, pt3;
                            tri_pts(x, y, j, st.pent, pt1, pt2, pt3);
                            if (laststep)
                                vdiff = 0.5 * rscale * (2.0 * rs_[ir] + rs_[pt1] + 2.0 * rs_[pt2] + rs_[pt3]) * o6 * sdiff;
                            else
                                vdiff = 0.5 * rscale;
                            if (j == 0) {
                                AT1 = rscale / aTr[id * 6 + j];
                                lap1 = o6 * (as[ir] + as[pt1]) - o3 * as[pt2];
                                lap2 = o6 * (as[pt1] + as[pt2]) - o3 * as[ir];
                                lap3 = o6 * (as[ir] + as[pt2]) - o3 * as[pt1];
                            }
                            AT2 = rscale / aTr[id * 6 + jp1];
                            lap4 = o6 * (as[ir] + as[pt2]) - o3 * as[pt3];
                            lap5 = o6 * (as[pt2] + as[pt3]) - o3 * as[ir];
                            lap6 = o6 * (as[ir] + as[pt3]) - o3 * as[pt2];
                            auto L = [&](int k, int e1, double l1, double l2, double l3, double A) {
                                return (-l1 * nti[id * 18 + e1 * 3 + k] + l2 * nte[id * 18 + e1 * 3 + k]
                                        + l3 * nti[id * 18 + ((e1 + 1) % n_faces) * 3 + k])
                                       * A;
                            };
                            if (j == 0) {
                                lapx1 = L(0, j, lap1, lap2, lap3, AT1);
                                lapx2 = L(0, jp1, lap4, lap5, lap6, AT2);
                                lapy1 = L(1, j, lap1, lap2, lap3, AT1);
                                lapy2 = L(1, jp1, lap4, lap5, lap6, AT2);
                                lapz1 = L(2, j, lap1, lap2, lap3, AT1);
                                lapz2 = L(2, jp1, lap4, lap5, lap6, AT2);
                            } else {
                                lapx1 = lapx2;
                                lapx2 = L(0, jp1, lap4, lap5, lap6, AT2);
                                lapy1 = lapy2;
                                lapy2 = L(1, jp1, lap4, lap5, lap6, AT2);
                                lapz1 = lapz2;
                                lapz2 = L(2, jp1, lap4, lap5, lap6, AT2);
                            }
                            (void)jp2;
                            lap += ((lapx1 + lapx2) * noa[id * 18 + j * 3 + 0] + (lapy1 + lapy2) * noa[id * 18 + j * 3 + 1]
                                    + (lapz1 + lapz2) * noa[id * 18 + j * 3 + 2])
                                   * vdiff;
                            if (st.pent && j == 4) break;
                        }
                        if (laststep) difftr[(id * nv + lev) * ntr + var] = lap;
                        else dif[(id * nv + lev) * ntr + var] = lap;
                    }
            }
        }
    if (laststep) return;  // the pole kernel writes nothing in its last pass
    for (int var = 0; var < ntr; var++)
        for (int id = P - 2; id < P; id++)
            for (int lev = 0; lev < nv; lev++) {
                int loc[5];
                for (int i = 0; i < 5; i++) loc[i] = o.g.point_local[id * 6 + i];
                double a_p[6];
                a_p[0] = tr[(id * nv + lev) * ntr + 0] / Rho[id * nv + lev];
                for (int i = 1; i < 6; i++) a_p[i] = tr[(loc[i - 1] * nv + lev) * ntr + var] / Rho[loc[i - 1] * nv + lev];
                const double rscale = rscale_of(o, lev);
                double lap = 0.0, lap1 = 0, lap2 = 0, lap3 = 0, lap4, lap5, lap6, AT1 = 0, AT2;
                double lapx1 = 0, lapx2 = 0, lapy1 = 0, lapy2 = 0, lapz1 = 0, lapz2 = 0;
                for (int k = 0; k < 5; k++) {
                    const int j = k + 1, jp1 = (j + 1) % 5, jp2 = (j + 2) % 5, kp1 = (k + 1) % 5, kp2 = (k + 2) % 5;
                    const double vdiff = 0.5 * rscale;
                    if (k == 0) {
                        lap1 = (o6 * (a_p[0] + a_p[j]) - o3 * a_p[jp1]);
                        lap2 = (o6 * (a_p[j] + a_p[jp1]) - o3 * a_p[0]);
                        lap3 = (o6 * (a_p[0] + a_p[jp1]) - o3 * a_p[j]);
                        AT1 = rscale / aTr[id * 6 + k];
                    }
                    AT2 = rscale / aTr[id * 6 + kp1];
                    lap4 = (o6 * (a_p[0] + a_p[jp1]) - o3 * a_p[jp2]);
                    lap5 = (o6 * (a_p[jp1] + a_p[jp2]) - o3 * a_p[0]);
                    lap6 = (o6 * (a_p[0] + a_p[jp2]) - o3 * a_p[jp1]);
                    auto L = [&](int c, int e1, int e2, double l1, double l2, double l3, double A) {
                        return (-l1 * nti[id * 18 + e1 * 3 + c] + l2 * nte[id * 18 + e1 * 3 + c]
                                + l3 * nti[id * 18 + e2 * 3 + c])
                               * A;
                    };
                    if (k == 0) {
                        lapx1 = L(0, k, kp1, lap1, lap2, lap3, AT1);
                        lapx2 = L(0, kp1, kp2, lap4, lap5, lap6, AT2);
                        lapy1 = L(1, k, kp1, lap1, lap2, lap3, AT1);
                        lapy2 = L(1, kp1, kp2, lap4, lap5, lap6, AT2);
                        lapz1 = L(2, k, kp1, lap1, lap2, lap3, AT1);
                        lapz2 = L(2, kp1, kp2, lap4, lap5, lap6, AT2);
                    } else {
                        lapx1 = lapx2;
                        lapx2 = L(0, kp1, kp2, lap4, lap5, lap6, AT2);
                        lapy1 = lapy2;
                        lapy2 = L(1, kp1, kp2, lap4, lap5, lap6, AT2);
                        lapz1 = lapz2;
                        lapz2 = L(2, kp1, kp2, lap4, lap5, lap6, AT2);
                    }
                    lap += ((lapx1 + lapx2) * noa[id * 18 + k * 3 + 2] + (lapy1 + lapy2) * noa[id * 18 + k * 3 + 1]
                            + (lapz1 + lapz2) * noa[id * 18 + k * 3 + 2])
                           * vdiff;
                }
                dif[(id * nv + lev) * ntr + var] = lap;
            }
}

// Tracer_Eq + _Poles (thor_chemistry.h:31-361): flux-form advection of the
// tracer mass by Mhs + Mhk and Whs + Whk into tracers, plus difftr * dt.
void tracer_eq(Oracle &o, double dt) {
    const int nv = o.nv, P = o.P, ntr = o.ntr;
    double *trs = o.A_("tracers"), *trk = o.A_("tracerk"), *Rho = o.A_("Rhos"), *Rhok = o.A_("Rhok"),
           *Mh = o.A_("Mhs"), *Mhk = o.A_("Mhk"), *Wh = o.A_("Whs"), *Whk = o.A_("Whk"), *difftr = o.A_("difftr");
    const double *div = o.g.div, *Alt = o.g.Altitude, *Alth = o.g.Altitudeh;
    const double A = o.p.A;
    auto mixr = [&](int gi, int l, int itr) { return trk[(gi * nv + l) * ntr + itr] / Rhok[gi * nv + l]; };
    // the column part (identical in both kernels but for the poles' lev-0 quirk)
    auto column = [&](int id, int lev, int itr, double nflx, double trc, bool pole) {
        const double altht = Alth[lev + 1], althl = Alth[lev];
        double r2p, r2m, r2l;
        if (o.p.DeepModel) {
            const double alt = Alt[lev];
            r2p = MPOW(altht + A, 2.0);
            r2m = MPOW(alt + A, 2.0);
            r2l = MPOW(althl + A, 2.0);
        } else {
            r2p = r2m = r2l = 1.0;
        }
        double trhl, trht, xi, xim1, xip1, a, b;
        if (lev == 0) {
            trhl = 0.0;
            xi = Alth[lev + 1];
            xim1 = Alt[lev];
            xip1 = Alt[lev + 1];
            a = (xi - xip1) / (xim1 - xip1);
            b = (xi - xim1) / (xip1 - xim1);
            if (pole)
                trht = trk[(id * nv + lev) * ntr + itr] / Rhok[id * nv + lev] * a
                       + trk[(id * nv + lev + 1) * ntr + itr] / Rhok[id * nv + lev] * b;
            else
                trht = mixr(id, lev, itr) * a + mixr(id, lev + 1, itr) * b;
        } else if (lev == nv - 1) {
            xi = Alth[lev];
            xim1 = Alt[lev - 1];
            xip1 = Alt[lev];
            a = (xi - xip1) / (xim1 - xip1);
            b = (xi - xim1) / (xip1 - xim1);
            trhl = mixr(id, lev - 1, itr) * a + mixr(id, lev, itr) * b;
            trht = 0.0;
        } else {
            xi = Alth[lev];
            xim1 = Alt[lev - 1];
            xip1 = Alt[lev];
            a = (xi - xip1) / (xim1 - xip1);
            b = (xi - xim1) / (xip1 - xim1);
            trhl = mixr(id, lev - 1, itr) * a + mixr(id, lev, itr) * b;
            xi = Alth[lev + 1];
            xim1 = Alt[lev];
            xip1 = Alt[lev + 1];
            a = (xi - xip1) / (xim1 - xip1);
            b = (xi - xim1) / (xip1 - xim1);
            trht = mixr(id, lev, itr) * a + mixr(id, lev + 1, itr) * b;
        }
        double whl, wht;
        if (lev == 0) {
            whl = 0.0;
            wht = Wh[id * (nv + 1) + lev + 1] + Whk[id * (nv + 1) + lev + 1];
        } else {
            whl = Wh[id * (nv + 1) + lev] + Whk[id * (nv + 1) + lev];
            wht = Wh[id * (nv + 1) + lev + 1] + Whk[id * (nv + 1) + lev + 1];
        }
        const double dz = altht - althl;
        const double dtr_dalt = (wht * trht * r2p - whl * trhl * r2l) / (dz * r2m);
        const double dtr = -(nflx + dtr_dalt) * dt;
        const double r = Rhok[id * nv + lev] + Rho[id * nv + lev];
        const double trn = trc * r + dtr;
        const long e = (id * nv + lev) * ntr + itr;
        trs[e] = trn - trc * Rhok[id * nv + lev];
        trs[e] = trs[e] + difftr[e] * dt;
    };
    double vs[3][NHL2], ts[NHL2];
    for (int itr = 0; itr < ntr; itr++)
        for (int ib = 0; ib < o.nr; ib++) {
            const int *m = o.g.maps + ib * NHL2;
            for (int lev = 0; lev < nv; lev++) {
                for (int k = 0; k < 3; k++)
                    fill(vs[k], m, [&](int gi) { return Mh[(gi * nv + lev) * 3 + k] + Mhk[(gi * nv + lev) * 3 + k]; });
                fill(ts, m, [&](int gi) { return mixr(gi, lev, itr); });
                const double rscale = rscale_of(o, lev);
                for (int y = 0; y < NL; y++)
                    for (int x = 0; x < NL; x++) {
                        Stencil st = stencil(m, x, y);
                        double nflx = 0.0;
                        for (int k = 0; k < 3; k++) {
                            const double *d = div + st.id * 21;
                            nflx += rscale * (d[0 * 3 + k] * vs[k][st.pt[0]] * ts[st.pt[0]] + d[1 * 3 + k] * vs[k][st.pt[1]] * ts[st.pt[1]]
                                              + d[2 * 3 + k] * vs[k][st.pt[2]] * ts[st.pt[2]] + d[3 * 3 + k] * vs[k][st.pt[3]] * ts[st.pt[3]]
                                              + d[4 * 3 + k] * vs[k][st.pt[4]] * ts[st.pt[4]] + d[5 * 3 + k] * vs[k][st.pt[5]] * ts[st.pt[5]]
                                              + d[6 * 3 + k] * vs[k][st.pt[6]] * ts[st.pt[6]]);
                        }
                        column(st.id, lev, itr, nflx, ts[st.ir], false);
                    }
            }
        }
    for (int itr = 0; itr < ntr; itr++)
        for (int id = P - 2; id < P; id++) {
            int loc[6];
            loc[0] = id;
            for (int i = 0; i < 5; i++) loc[i + 1] = o.g.point_local[id * 6 + i];
            for (int lev = 0; lev < nv; lev++) {
                const double rscale = rscale_of(o, lev);
                double nflx = 0.0;
                for (int k = 0; k < 3; k++) {
                    const double *d = div + id * 21;
                    auto term = [&](int i) {
                        const int q = loc[i];
                        return d[i * 3 + k] * (Mh[(q * nv + lev) * 3 + k] + Mhk[(q * nv + lev) * 3 + k]) * mixr(q, lev, itr);
                    };
                    nflx += rscale * (term(0) + term(1) + term(2) + term(3) + term(4) + term(5));
                }
                column(id, lev, itr, nflx, mixr(id, lev, itr), true);
            }
        }
}

void zero(Oracle &o, const char *n) { std::fill(o.arr[n].begin(), o.arr[n].end(), 0.0); }
void copy(Oracle &o, const char *dst, const char *src) { o.arr[dst] = o.arr[src]; }

// ESP::Thor (thor_driver.cu:70-1466) for the configuration space used here.
void thor(Oracle &o) {
    const int ns_total = o.p.n_substeps;
    copy(o, "Mhk", "Mh");
    copy(o, "Whk", "Wh");
    copy(o, "Wk", "W");
    copy(o, "Rhok", "Rho");
    copy(o, "pressurek", "pressure");
    zero(o, "Mhs");
    zero(o, "Rhos");
    zero(o, "Whs");
    zero(o, "Ws");
    zero(o, "pressures");
    if (o.ntr) {  // dyn_core_loop_init of the tracer module (chemistry.cu:333-343)
        copy(o, "tracerk", "tracer");
        zero(o, "tracers");
    }
    for (int rk = 0; rk < 3; rk++) {
        int ns_it = rk == 0 ? 1 : (rk == 1 ? ns_total / 2 : ns_total);
        double times = rk == 0 ? o.p.timestep / 3.0 : o.p.timestep / (double)ns_total;
        zero(o, "Adv");
        advec_cori1(o);
        advec_poles(o);
        advec_cori2(o);
        temperature_h_pt_geff(o);
        zero(o, "SlowMh");
        zero(o, "SlowWh");
        zero(o, "SlowRho");
        zero(o, "Slowpressure");
        if (o.p.HyDiff) {
            zero(o, "diff");
            int stepnum = 0;
            for (int ihyp = 0; ihyp < o.p.HyDiffOrder / 2 - 1; ihyp++) {
                diffusion_pass(o, stepnum, false);
                stepnum += 1;
            }
            zero(o, "diffrh");
            diffusion_pass(o, stepnum, true);
            if (o.p.VertHyDiff) vertical_diff(o);
            correct_horizontal(o);
        }
        zero(o, "DivM");
        zero(o, "divg_Mh");
        if (o.p.DivDampP) {
            divm_pass(o, "Mhk", "Whk", false);
            divm_pass(o, "Mhk", "Whk", true);
        }
        if (o.ntr && o.p.HyDiff) {  // dyn_core_loop_slow_modes (chemistry.cu:345-427)
            tracer_diffusion(o, false);
            tracer_diffusion(o, true);
        }
        if (o.sponge_on && !o.sponge_imp()) sponge_explicit(o, rk);
        slow_modes(o);
        if (rk > 0) update_rk(o);
        if (rk > 0 && o.ntr) {  // the RK arrays at the start of stages 2 and 3
            double *a = o.A_("tracers"), *ak = o.A_("tracerk"), *ai = o.A_("tracer");
            for (size_t i = 0; i < o.arr["tracers"].size(); i++) a[i] = ai[i] - ak[i];
        }
        for (int ns = 0; ns < ns_it; ns++) {
            zero(o, "DivM");
            zero(o, "divg_Mh");
            if (o.p.DivDampP) {
                divm_pass(o, "Mhs", "Whs", false);
                divm_pass(o, "Mhs", "Whs", true);
            }
            momentum_eq(o, times);
            prepare_implicit_vertical(o);
            vertical_eq(o, times);
            if (o.ntr) tracer_eq(o, times);  // dyn_core_loop_fast_modes (chemistry.cu:429-492)
            density_pressure(o, times);
        }
        update_rk2(o);
        if (o.ntr) {  // the RK arrays at the end of every stage
            double *a = o.A_("tracers"), *ak = o.A_("tracerk");
            for (size_t i = 0; i < o.arr["tracers"].size(); i++) ak[i] += a[i];
        }
    }
    if (o.ntr) copy(o, "tracer", "tracerk");  // dyn_core_loop_end (chemistry.cu:495-502)
    copy(o, "Mh", "Mhk");
    copy(o, "Wh", "Whk");
    copy(o, "W", "Wk");
    copy(o, "Rho", "Rhok");
    copy(o, "pressure", "pressurek");
    o.nstep++;
}

void held_suarez(Oracle &o);

// ESP::ProfX for the Held-Suarez benchmark (profx_driver.cu:70-474).
void profx_hs(Oracle &o) {
    const int nv = o.nv;
    double *W = o.A_("W"), *Wh = o.A_("Wh"), *T = o.A_("temperature"), *pt = o.A_("pt"), *p = o.A_("pressure"),
           *Rho = o.A_("Rho"), *Rd = o.A_("Rd"), *Cp = o.A_("Cp");
    const double *Alt = o.g.Altitude, *Alth = o.g.Altitudeh;
    zero(o, "profx_Qheat");
    // Recompute_W (profx_auxiliary.h:110-130); with the sponge it ran before it
    for (int id = 0; id < o.P && !o.sponge_imp(); id++)
        for (int lev = 0; lev < nv; lev++) {
            double whl = Wh[id * (nv + 1) + lev], wht = Wh[id * (nv + 1) + lev + 1];
            double xi = Alt[lev], xim = Alth[lev], xip = Alth[lev + 1];
            double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            W[id * nv + lev] = whl * intt + wht * intl;
        }
    // Compute_temperature (profx_auxiliary.h:64-93)
    ORACLE_PAR
    for (int id = 0; id < o.P; id++)
        for (int lev = 0; lev < nv; lev++) {
            int e = id * nv + lev;
            double Cv = Cp[e] - Rd[e], CvoCp = Cv / Cp[e];
            T[e] = p[e] / (Rd[e] * Rho[e]);
            pt[e] = (o.p.P_Ref / (Rd[e] * Rho[e])) * MPOW(p[e] / o.p.P_Ref, CvoCp);
        }
    held_suarez(o);
    // Compute_pressure (profx_auxiliary.h:95-108)
    for (int e = 0; e < o.P * nv; e++) p[e] = T[e] * Rd[e] * Rho[e];
}

// held_suarez (profx_held_suarez.h:51-117): Rayleigh drag on Mh and Newtonian
// relaxation of temperature, with dt = the large time step.
void held_suarez(Oracle &o) {
    const int nv = o.nv;
    double *T = o.A_("temperature"), *p = o.A_("pressure"), *Rho = o.A_("Rho"), *Mh = o.A_("Mh");
    const double *Alt = o.g.Altitude, *ll = o.g.lonlat;
    const double Gravit = o.p.Gravit, kappa = o.p.Rd / o.p.Cp, dt = o.p.timestep;
    ORACLE_PAR
    for (int id = 0; id < o.P; id++)
        for (int lev = 0; lev < nv; lev++) {
            double sigmab = 0.7, ka = (1.0 / 40.0) * (1.0 / 86400.0), kf = 1.0 / 86400.0,
                   ks = (1.0 / 4.0) * (1.0 / 86400.0), dTy = 60.0, dthetaz = 10.0, p0 = 100000.0;
            double lat = ll[id * 2 + 1];
            double psm1 = p[id * nv + 1] - Rho[id * nv + 0] * Gravit * (-Alt[0] - Alt[1]);
            double ps = 0.5 * (p[id * nv + 0] + psm1);
            double pre = p[id * nv + lev];
            double sigma = (pre / ps), sigma0 = (pre / p0);
            double Teq = std::max(200.0, (315.0 - dTy * MPOW(MSIN(lat), 2.0)
                                          - dthetaz * MLOG(sigma0) * MPOW(MCOS(lat), 2.0))
                                             * MPOW(sigma0, kappa));
            double kt = ka + (ks - ka) * std::max(0.0, (sigma - sigmab) / (1.0 - sigmab)) * MPOW(MCOS(lat), 4.0);
            double kv = kf * std::max(0.0, (sigma - sigmab) / (1.0 - sigmab));
            for (int k = 0; k < 3; k++) Mh[(id * nv + lev) * 3 + k] = Mh[(id * nv + lev) * 3 + k] / (1.0 + kv * dt);
            T[id * nv + lev] -= kt * dt * (T[id * nv + lev] - Teq);
        }
}

// ------------------------------------------------ boundary-layer module
// rayleighHS (boundary_layer.cu:441-490), run by phy_modules_phy_loop after
// the benchmark forcing: implicit drag on Mh below sigma = bl_sigma, with the
// surface pressure extrapolated from the two lowest layers of p (the pressure
// before ProfX's Compute_pressure).
void boundary_layer_rayleigh(Oracle &o, const double *p) {
    const int nv = o.nv;
    double *Mh = o.A_("Mh"), *Rho = o.A_("Rho");
    const double *Alt = o.g.Altitude;
    const double Gravit = o.p.Gravit, A = o.p.A, dt = o.p.timestep, kf = o.blq[0], sigmab = o.blq[1];
    ORACLE_PAR
    for (int id = 0; id < o.P; id++)
        for (int lev = 0; lev < nv; lev++) {
            double psm1;
            if (o.p.GravHeightVar)
                psm1 = p[id * nv + 1] - Rho[id * nv + 0] * Gravit * MPOW(A / (A + Alt[0]), 2) * (-Alt[0] - Alt[1]);
            else
                psm1 = p[id * nv + 1] - Rho[id * nv + 0] * Gravit * (-Alt[0] - Alt[1]);
            const double ps = 0.5 * (p[id * nv + 0] + psm1);
            const double sigma = (p[id * nv + lev] / ps);
            const double kv_hs = kf * std::max(0.0, (sigma - sigmab) / (1.0 - sigmab));
            for (int k = 0; k < 3; k++)
                Mh[(id * nv + lev) * 3 + k] = Mh[(id * nv + lev) * 3 + k] / (1.0 + kv_hs * dt);
        }
}

// --------------------------------------------------- Newtonian relaxation
// Hot-Jupiter forcing (no reference counterpart; structure of shallowHJ,
// profx_shallowHJ.h:51-117): T relaxes implicitly towards
// Teq = T_mean + dT w(p) cos(lon - lon_sub) cos(lat) with time scale tau(p);
// w = 1 above p_contrast, linear in ln p down to 0 at p_bottom; ln tau linear
// in ln p between (p_top, tau_top) and (p_bottom, tau_bottom), clamped.
void profx_nr(Oracle &o) {
    const int nv = o.nv;
    double *W = o.A_("W"), *Wh = o.A_("Wh"), *T = o.A_("temperature"), *p = o.A_("pressure"), *Rho = o.A_("Rho");
    const double *Alt = o.g.Altitude, *Alth = o.g.Altitudeh, *ll = o.g.lonlat;
    const double *q = o.nrq;  // T_mean, dT, lon_sub, p_top, p_bottom, p_contrast, tau_top, tau_bottom
    const double lnp_top = std::log(q[3]), lnp_bot = std::log(q[4]), lnp_con = std::log(q[5]);
    const double lntau_top = std::log(q[6]), lntau_bot = std::log(q[7]);
    const double Rd = o.p.Rd, dt = o.p.timestep;
    zero(o, "profx_Qheat");
    for (int id = 0; id < o.P; id++)
        for (int lev = 0; lev < nv; lev++) {
            const int e = id * nv + lev;
            double whl = Wh[id * (nv + 1) + lev], wht = Wh[id * (nv + 1) + lev + 1];
            double xi = Alt[lev], xim = Alth[lev], xip = Alth[lev + 1];
            double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            if (!o.sponge_imp()) W[e] = whl * intt + wht * intl;
            const double pre = p[e], rho = Rho[e];
            double t = pre / (Rd * rho);
            const double lnp = MLOG(pre);
            double w = (lnp_bot - lnp) / (lnp_bot - lnp_con);
            w = std::fmin(1.0, std::fmax(0.0, w));
            const double Teq = q[0] + q[1] * w * MCOS(ll[id * 2] - q[2]) * MCOS(ll[id * 2 + 1]);
            double fr = (lnp - lnp_top) / (lnp_bot - lnp_top);
            fr = std::fmin(1.0, std::fmax(0.0, fr));
            const double tau = MEXP(lntau_top + fr * (lntau_bot - lntau_top));
            t = (t + dt * Teq / tau) / (1.0 + dt / tau);
            T[e] = t;
            p[e] = t * Rd * rho;
        }
}

// ------------------------------------------------ double-gray radiation
// source_func_lin (profx_RT.h:133-160).
double rt_source(double bb, double bl, double bt, double tau, double diff_ang) {
    double e1 = 0.0, e2 = 0.0, e = 0.0;
    if (tau >= 1e-10) {
        e1 = bb - bl + (bl - (diff_ang / (tau / 2.0)) * (bb - bl)) * (1.0 - MEXP(-(tau / 2.0) / diff_ang));
        e2 = bl - bt + (bt - (diff_ang / (tau / 2.0)) * (bl - bt)) * (1.0 - MEXP(-(tau / 2.0) / diff_ang));
        e = e1 + e2 * MEXP(-(tau / 2.0) / diff_ang);
    } else {
        for (int i = 0; i < 5; i++) {
            int fac = 1;
            for (int n1 = i + 2; n1 != 0; n1--) fac *= n1;  // factorial_num (profx_RT.h:119-131)
            const double sgn = (i % 2 == 0) ? 1.0 : -1.0;   // pow(-1, i + 2)
            e1 = e1 + sgn * ((bl + (i + 1) * bl) / fac) * MPOW((tau / 2.0) / diff_ang, i + 1.0);
            e2 = e2 + sgn * ((bt + (i + 1) * bb) / fac) * MPOW((tau / 2.0) / diff_ang, i + 1.0);
        }
        e = e1 + e2 * MEXP(-(tau / 2.0) / diff_ang);
    }
    return e;
}

// ProfX with the double-gray module: Recompute_W, Compute_temperature,
// rtm_dual_band (profx_RT.h:296-587: computetau, radcsw, radclw; no surface,
// constant kappa_lw, Qheat_scaling 1), Compute_pressure. cos z for a
// synchronous rotator (insolation.cu:128-207).
void profx_rt(Oracle &o) {
    const int nv = o.nv, nvi = nv + 1;
    double *W = o.A_("W"), *Wh = o.A_("Wh"), *T = o.A_("temperature"), *p = o.A_("pressure"), *Rho = o.A_("Rho"),
           *Qh = o.A_("profx_Qheat"), *insol = o.A_("insol"), *ASR = o.A_("ASR"), *OLR = o.A_("OLR");
    const double *Alt = o.g.Altitude, *Alth = o.g.Altitudeh, *ll = o.g.lonlat, *areasT = o.g.areasT;
    // kappa_sw, kappa_lw, n_sw, n_lw, f_lw, diff_ang, Tint, albedo, incflx, ps0, lon_sub
    const double *q = o.rtq;
    const double kappa_sw = q[0], kappa_lw = q[1], n_sw = q[2], n_lw = q[3], f_lw = q[4], diff_ang = q[5],
                 tint = q[6], alb = q[7], incflx = q[8], ps0 = q[9], lon_sub = q[10];
    const double Rd = o.p.Rd, Cp = o.p.Cp, gravit = o.p.Gravit, A = o.p.A, timestep = o.p.timestep;
    const double bc = 5.677036E-8;
    zero(o, "profx_Qheat");
    for (int id = 0; id < o.P; id++)
        for (int lev = 0; lev < nv; lev++) {
            const int e = id * nv + lev;
            double whl = Wh[id * nvi + lev], wht = Wh[id * nvi + lev + 1];
            double xi = Alt[lev], xim = Alth[lev], xip = Alth[lev + 1];
            double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            if (!o.sponge_imp()) W[e] = whl * intt + wht * intl;
            T[e] = p[e] / (Rd * Rho[e]);
        }
    ORACLE_PAR
    for (int id = 0; id < o.P; id++) {
        std::vector<double> phtemp(nvi), thtemp(nvi), ttemp(nv), tau(2 * nv), dtemp(nv), fsw_up(nvi), fsw_dn(nvi),
            flw_up(nvi), flw_dn(nvi);
        const double *pr = p + (size_t)id * nv, *rh = Rho + (size_t)id * nv, *te = T + (size_t)id * nv;
        for (int lev = 0; lev < nv; lev++) dtemp[lev] = 0.0;
        for (int lev = 0; lev <= nv; lev++) {
            if (lev == 0) {
                double psm = pr[1] - rh[0] * gravit * (-Alt[0] - Alt[1]);
                phtemp[0] = 0.5 * (pr[0] + psm);
                ttemp[0] = te[0];
                thtemp[0] = ttemp[0];
            } else if (lev == nv) {
                double pp = pr[nv - 2] + (pr[nv - 1] - pr[nv - 2]) / (Alt[nv - 1] - Alt[nv - 2])
                                             * (2 * Alth[nv] - Alt[nv - 1] - Alt[nv - 2]);
                if (pp < 0) pp = 0;
                phtemp[nv] = 0.5 * (pr[nv - 1] + pp);
                thtemp[nv] = te[nv - 1];
            } else {
                ttemp[lev] = te[lev];
                double xi = Alth[lev], xim = Alt[lev - 1], xip = Alt[lev];
                double a = (xi - xip) / (xim - xip), b = (xi - xim) / (xip - xim);
                phtemp[lev] = pr[lev - 1] * a + pr[lev] * b;
                thtemp[lev] = te[lev - 1] * a + ttemp[lev] * b;
            }
        }
        double coszrs = MCOS(ll[id * 2 + 1]) * MCOS(ll[id * 2] - lon_sub);
        if (coszrs < 0.0) coszrs = 0.0;
        // computetau (profx_RT.h:267-292)
        for (int lev = 0; lev < nv; lev++) {
            tau[lev * 2] = n_sw * kappa_sw * MPOW(pr[lev] / ps0, n_sw - 1) * rh[lev] * (Alth[lev + 1] - Alth[lev]);
            tau[lev * 2 + 1] = kappa_lw * (1 + n_lw * (1 - f_lw) / f_lw * MPOW(pr[lev] / ps0, n_lw - 1)) * rh[lev]
                               * (Alth[lev + 1] - Alth[lev]);
        }
        for (int lev = 0; lev <= nv; lev++) fsw_up[lev] = fsw_dn[lev] = 0.0;
        if (coszrs > 0.0) {
            // radcsw (profx_RT.h:14-117)
            double tt = (kappa_sw / gravit) * phtemp[nv];
            insol[id] = incflx * coszrs;
            double flux_top = insol[id] * (1.0 - alb);
            fsw_dn[nv] = flux_top * MEXP(-(1.0 / coszrs) * tt);
            for (int lev = nv; lev >= 1; lev--) fsw_dn[lev - 1] = fsw_dn[lev] * MEXP(-(1.0 / coszrs) * tau[(lev - 1) * 2]);
            for (int lev = 0; lev <= nv; lev++) fsw_up[lev] = 0.0;
            for (int lev = 0; lev < nv; lev++)
                dtemp[lev] = -((fsw_up[lev] - fsw_dn[lev]) - (fsw_up[lev + 1] - fsw_dn[lev + 1]))
                             / ((Alth[lev] - Alth[lev + 1]));
        } else {
            insol[id] = 0;
        }
        const double rscale = o.p.DeepModel ? (A + Alth[nv]) / A : 1.0;
        ASR[id] = fsw_dn[nv] * areasT[id] * MPOW(rscale, 2);
        // radclw (profx_RT.h:162-265)
        for (int lev = 0; lev <= nv; lev++) flw_up[lev] = flw_dn[lev] = 0.0;
        flw_dn[nv] = 0.0;
        for (int lev = nv - 1; lev >= 0; lev--) {
            if (tau[2 * lev + 1] < 0.0) tau[2 * lev + 1] = 0.0;
            double tb = thtemp[lev], tl = ttemp[lev], tt = thtemp[lev + 1];
            double bb = bc * tb * tb * tb * tb, bl = bc * tl * tl * tl * tl, bt = bc * tt * tt * tt * tt;
            double ed = rt_source(bb, bl, bt, tau[2 * lev + 1], diff_ang);
            flw_dn[lev] = ed + flw_dn[lev + 1] * MEXP(-(1. / diff_ang) * tau[2 * lev + 1]);
        }
        flw_up[0] = bc * tint * tint * tint * tint;
        flw_up[0] += flw_dn[0];
        for (int lev = 1; lev <= nv; lev++) {
            if (tau[2 * (lev - 1) + 1] < 0.0) tau[2 * (lev - 1) + 1] = 0.0;
            double tb = thtemp[lev - 1], tl = ttemp[lev - 1], tt = thtemp[lev];
            double bb = bc * tb * tb * tb * tb, bl = bc * tl * tl * tl * tl, bt = bc * tt * tt * tt * tt;
            double eu = rt_source(bt, bl, bb, tau[2 * (lev - 1) + 1], diff_ang);
            flw_up[lev] = eu + flw_up[lev - 1] * MEXP(-(1. / diff_ang) * tau[2 * (lev - 1) + 1]);
        }
        for (int lev = 0; lev < nv; lev++)
            dtemp[lev] = dtemp[lev]
                         - ((flw_up[lev] - flw_dn[lev]) - (flw_up[lev + 1] - flw_dn[lev + 1]))
                               / ((Alth[lev] - Alth[lev + 1]));
        OLR[id] = flw_up[nv] * areasT[id] * MPOW(rscale, 2);
        for (int lev = 0; lev < nv; lev++) {
            if (pr[lev] + Rd / (Cp - Rd) * dtemp[lev] * timestep < 0) dtemp[lev] = -pr[lev] / timestep;
            Qh[id * nv + lev] += dtemp[lev];
        }
    }
    for (int e = 0; e < o.P * nv; e++) p[e] = T[e] * Rd * Rho[e];
}

// ------------------------------------------------------ Rayleigh sponge
// Latitude bins of the zonal means (grid.cu:3215-3262): nlat bins of width
// pi / nlat from -pi/2 (edges accumulated one by one); each point gets its bin,
// the number of points in that bin and its index within the bin in point
// order; points at or above the last edge go to the top bin. max_count is the
// largest bin rounded up to a power of two.
void zonal_table(Oracle &o, int nlat) {
    const int P = o.P;
    const double *ll = o.g.lonlat;
    const double des_lat = M_PI / nlat;
    std::vector<int> count(nlat, 0);
    std::vector<double> edge(nlat + 1, 0.0);
    o.ztab.assign((size_t)P * 3, 0);
    edge[0] = -M_PI / 2.0;
    for (int j = 1; j < nlat + 1; j++) edge[j] = edge[j - 1] + des_lat;
    for (int i = 0; i < P; i++) {
        const double lat = ll[i * 2 + 1];
        for (int j = 0; j < nlat; j++)
            if (lat >= edge[j] && lat < edge[j + 1]) {
                o.ztab[i * 3] = j;
                o.ztab[i * 3 + 2] = count[j]++;
                break;
            }
        if (lat >= edge[nlat]) {
            o.ztab[i * 3] = nlat - 1;
            o.ztab[i * 3 + 2] = count[nlat - 1]++;
        }
    }
    int mc = 0;
    for (int i = 0; i < P; i++) {
        o.ztab[i * 3 + 1] = count[o.ztab[i * 3]];
        mc = std::max(mc, o.ztab[i * 3 + 1]);
    }
    o.max_count = (int)std::pow(2, std::ceil(std::log(mc) / std::log(2)));
}

// RayleighSponge with raysp_calc_mode = imp at the start of ProfX
// (profx_driver.cu:92-195): Recompute_W, the zonal means of u, v, w per
// (bin, level) (zonal_uv / zonal_w, profx_sponge.cu:50-102: each point stores
// value / count in its bin's slot; gpu_sum_on_device_sponge, reduction_add.h:
// 193-282: 1024-thread tree per 2048 slots, block results added on the host in
// order), then sponge_layer (profx_sponge.cu:138-437, implicit branch): winds
// interpolated in latitude between bin centres, relaxed with
// k = R sin^2(pi/2 (n - ns) / (1 - ns)) for n = Altitude / Top >= ns, M and W
// rebuilt from the new winds, Wh interpolated from W at interior interfaces.
// With RayleighSpongeT the temperature p / Rd / rho relaxes the same way
// towards its zonal mean (zonal_temp, profx_sponge.cu:104-127, summed by
// gpu_sum_on_device at profx_driver.cu:131-149; the new pressure
// rho Rd (T / dt + kT Tbar) / (1 / dt + kT), profx_sponge.cu:388-404). THOR
// does not clear the slots of the temperature table that no point writes;
// here every unwritten slot is zero.
// Zonal means of u, v, w (and T) per (bin, level) from the given state
// arrays into o.vbar [bin][lev][3] and o.tbar [bin][lev].
void sponge_means(Oracle &o, const double *Mh, const double *Rho, const double *W, const double *pr) {
    const int nv = o.nv, P = o.P, nlat = (int)o.spq[3];
    const bool dmu = o.spq[4] != 0.0, dmw = o.spq[5] != 0.0, dmt = o.spq[6] != 0.0;
    const double *Rd = o.A_("Rd"), *ll = o.g.lonlat;
    const long L = o.max_count;
    o.vbar.assign((size_t)nlat * nv * 3, 0.0);
    o.tbar.assign((size_t)nlat * nv, 0.0);
    if (dmu || dmw) {
        std::vector<double> ut((size_t)nlat * nv * L, 0.0), vt(ut.size(), 0.0), wt(ut.size(), 0.0);
        for (int id = 0; id < P; id++)
            for (int lev = 0; lev < nv; lev++) {
                const int b = o.ztab[id * 3], cnt = o.ztab[id * 3 + 1], k = o.ztab[id * 3 + 2];
                const size_t at = ((size_t)b * nv + lev) * L + k;
                const double lon = ll[id * 2] < 0 ? ll[id * 2] + 2.0 * M_PI : ll[id * 2], lat = ll[id * 2 + 1];
                const double rho = Rho[id * nv + lev], *M = &Mh[(id * nv + lev) * 3];
                const double u = (M[0] * (-MSIN(lon)) + M[1] * (MCOS(lon)) + M[2] * 0) / rho;
                const double v = (M[0] * (-MSIN(lat) * MCOS(lon)) + M[1] * (-MSIN(lat) * MSIN(lon)) + M[2] * (MCOS(lat))) / rho;
                ut[at] = u / cnt;
                vt[at] = v / cnt;
                wt[at] = (W[id * nv + lev] / rho) / cnt;
            }
        for (int b = 0; b < nlat; b++)
            for (int lev = 0; lev < nv; lev++) {
                const size_t row = ((size_t)b * nv + lev) * L, v3 = ((size_t)b * nv + lev) * 3;
                if (dmu) {
                    o.vbar[v3 + 0] = oracle_tree_sum(&ut[row], L);
                    o.vbar[v3 + 1] = oracle_tree_sum(&vt[row], L);
                }
                if (dmw) o.vbar[v3 + 2] = oracle_tree_sum(&wt[row], L);
            }
    }
    if (dmt) {
        std::vector<double> tt((size_t)nlat * nv * L, 0.0);
        for (int id = 0; id < P; id++)
            for (int lev = 0; lev < nv; lev++) {
                const int b = o.ztab[id * 3], cnt = o.ztab[id * 3 + 1], k = o.ztab[id * 3 + 2];
                const double temperature = pr[id * nv + lev] / Rd[id * nv + lev] / Rho[id * nv + lev];
                tt[((size_t)b * nv + lev) * L + k] = temperature / cnt;
            }
        for (int b = 0; b < nlat; b++)
            for (int lev = 0; lev < nv; lev++)
                o.tbar[(size_t)b * nv + lev] = oracle_tree_sum(&tt[((size_t)b * nv + lev) * L], L);
    }
}

// The zonal mean of component c (0 u, 1 v, 2 w, 3 T) at point id, level lev:
// the polar bands' own value, else interpolated between bin centres.
double sponge_mean_at(const Oracle &o, int id, int lev, int c) {
    const int nv = o.nv, nlat = (int)o.spq[3], intlat = o.ztab[id * 3];
    const double des_lat = M_PI / nlat, lat = o.g.lonlat[id * 2 + 1];
    const double lat1 = -M_PI / 2 + 0.5 * des_lat + des_lat * (intlat - 1);
    const double lat2 = -M_PI / 2 + 0.5 * des_lat + des_lat * (intlat);
    const double lat3 = -M_PI / 2 + 0.5 * des_lat + des_lat * (intlat + 1);
    auto at = [&](int b) { return c == 3 ? o.tbar[(size_t)b * nv + lev] : o.vbar[((size_t)b * nv + lev) * 3 + c]; };
    if (lat <= -M_PI / 2 + 0.5 * des_lat) return at(intlat);
    if (lat >= -M_PI / 2 + 0.5 * des_lat + des_lat * (nlat - 1)) return at(intlat);
    if (lat >= lat1 && lat < lat2) {
        const double intt = (lat - lat2) / (lat1 - lat2), intl = (lat - lat1) / (lat2 - lat1);
        return at(intlat - 1) * intt + at(intlat) * intl;
    }
    const double intt = (lat - lat3) / (lat2 - lat3), intl = (lat - lat2) / (lat3 - lat2);
    return at(intlat) * intt + at(intlat + 1) * intl;
}

// The sponge rate of level lev for the rate R: Rv_fac R sin^2(pi/2 (n - ns) /
// (1 - ns)) above n = Altitude / Top = ns; Rv_fac is 1, or with shrink_sponge
// exp(-(step - t_shrink) dt / (1000 dt)) from step t_shrink on
// (profx_driver.cu:153-160, thor_driver.cu:822-830).
double sponge_rate(const Oracle &o, int lev, double R) {
    const double n = o.g.Altitude[lev] / o.g.Altitudeh[o.nv], ns = o.spq[0];
    double Rv_fac = 1;
    if (o.spq[9] != 0.0) {
        const long current_step = o.nstep + 1, t_shrink = (long)o.spq[10];
        const double timestep = o.p.timestep;
        if (current_step * timestep >= t_shrink * timestep) {
            const double shrink_scale = timestep * 1000;
            Rv_fac = std::exp(-(current_step * timestep - t_shrink * timestep) / shrink_scale);
        }
    }
    return n >= ns ? Rv_fac * R * MPOW(MSIN(0.5 * M_PI * (n - ns) * (1.0) / (1.0 - ns)), 2.0) : 0.0;
}

void sponge(Oracle &o) {
    const int nv = o.nv, P = o.P;
    const double Ruv = o.spq[1], Rw = o.spq[2], RT = o.spq[7];
    const bool dmu = o.spq[4] != 0.0, dmw = o.spq[5] != 0.0, dmt = o.spq[6] != 0.0;
    double *W = o.A_("W"), *Wh = o.A_("Wh"), *Mh = o.A_("Mh"), *Rho = o.A_("Rho"), *pr = o.A_("pressure"),
           *Rd = o.A_("Rd");
    const double *Alt = o.g.Altitude, *Alth = o.g.Altitudeh, *ll = o.g.lonlat;
    const double dt = o.p.timestep;
    for (int id = 0; id < P; id++)
        for (int lev = 0; lev < nv; lev++) {
            double whl = Wh[id * (nv + 1) + lev], wht = Wh[id * (nv + 1) + lev + 1];
            double xi = Alt[lev], xim = Alth[lev], xip = Alth[lev + 1];
            double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            W[id * nv + lev] = whl * intt + wht * intl;
        }
    sponge_means(o, Mh, Rho, W, pr);
    for (int id = 0; id < P; id++) {
        const double lat = ll[id * 2 + 1], lon = ll[id * 2] < 0 ? ll[id * 2] + 2.0 * M_PI : ll[id * 2];
        for (int lev = 0; lev < nv; lev++) {
            const double rho = Rho[id * nv + lev];
            const double *M = &Mh[(id * nv + lev) * 3];
            const double u = (M[0] * (-MSIN(lon)) + M[1] * (MCOS(lon)) + M[2] * 0) / rho;
            const double v = (M[0] * (-MSIN(lat) * MCOS(lon)) + M[1] * (-MSIN(lat) * MSIN(lon)) + M[2] * (MCOS(lat))) / rho;
            const double w = W[id * nv + lev] / rho;
            const double kuv = sponge_rate(o, lev, Ruv), kw = sponge_rate(o, lev, Rw);
            double unew, vnew, wnew;
            if (dmu) {
                unew = (u / dt + kuv * sponge_mean_at(o, id, lev, 0)) / (1.0 / dt + kuv);
                vnew = (v / dt + kuv * sponge_mean_at(o, id, lev, 1)) / (1.0 / dt + kuv);
            } else {
                unew = (u / dt) / (1.0 / dt + kuv);
                vnew = (v / dt) / (1.0 / dt + kuv);
            }
            if (dmw) wnew = (w / dt + kw * sponge_mean_at(o, id, lev, 2)) / (1.0 / dt + kw);
            else wnew = (w / dt) / (1.0 / dt + kw);
            const double vx = unew * (-MSIN(lon)) + vnew * (-MSIN(lat) * MCOS(lon));
            const double vy = unew * (MCOS(lon)) + vnew * (-MSIN(lat) * MSIN(lon));
            const double vz = vnew * (MCOS(lat));
            Mh[(id * nv + lev) * 3 + 0] = vx * rho;
            Mh[(id * nv + lev) * 3 + 1] = vy * rho;
            Mh[(id * nv + lev) * 3 + 2] = vz * rho;
            W[id * nv + lev] = wnew * rho;
            if (dmt) {
                const double kT = sponge_rate(o, lev, RT);
                const double temperature = pr[id * nv + lev] / Rd[id * nv + lev] / Rho[id * nv + lev];
                pr[id * nv + lev] = Rho[id * nv + lev] * Rd[id * nv + lev]
                                    * (temperature / dt + kT * sponge_mean_at(o, id, lev, 3)) / (1.0 / dt + kT);
            }
        }
        for (int lev = 1; lev < nv; lev++) {
            const double xi = Alth[lev], xim = Alt[lev - 1], xip = Alt[lev];
            const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            Wh[id * (nv + 1) + lev] = W[id * nv + lev - 1] * intt + W[id * nv + lev] * intl;
        }
    }
}

// RayleighSponge with raysp_calc_mode = exp1 / exp3, in each RK stage before
// the slow modes (thor_driver.cu:749-860): profx_dMh, profx_dWh and profx_dW
// cleared; the zonal means from the stage values Mhk, Rhok, Wk, pressurek in
// the first stage (exp1) or every stage (exp3); then sponge_layer's explicit
// branch (profx_sponge.cu:255-292, 388-404, 417-434): the drag -k (u - ubar)
// etc. times rho into profx_dMh and profx_dW, interpolated to profx_dWh at the
// interior interfaces, and the heating Rd rho (-k_T (T - Tbar)) added to
// profx_Qheat (which ProfX cleared: it adds up over the stages of a step).
void sponge_explicit(Oracle &o, int rk) {
    const int nv = o.nv, P = o.P;
    const double Ruv = o.spq[1], Rw = o.spq[2], RT = o.spq[7];
    const bool dmu = o.spq[4] != 0.0, dmw = o.spq[5] != 0.0, dmt = o.spq[6] != 0.0;
    double *Mh = o.A_("Mhk"), *Rho = o.A_("Rhok"), *W = o.A_("Wk"), *pr = o.A_("pressurek"), *Rd = o.A_("Rd");
    double *dMh = o.A_("profx_dMh"), *dWh = o.A_("profx_dWh"), *Qh = o.A_("profx_Qheat");
    const double *Alt = o.g.Altitude, *Alth = o.g.Altitudeh, *ll = o.g.lonlat;
    zero(o, "profx_dMh");
    zero(o, "profx_dWh");
    std::vector<double> dW((size_t)P * nv, 0.0);
    if (rk == 0 || (int)o.spq[8] == 3) sponge_means(o, Mh, Rho, W, pr);
    for (int id = 0; id < P; id++) {
        const double lat = ll[id * 2 + 1], lon = ll[id * 2] < 0 ? ll[id * 2] + 2.0 * M_PI : ll[id * 2];
        for (int lev = 0; lev < nv; lev++) {
            const long e = (long)id * nv + lev;
            const double rho = Rho[e];
            const double *M = &Mh[e * 3];
            const double u = (M[0] * (-MSIN(lon)) + M[1] * (MCOS(lon)) + M[2] * 0) / rho;
            const double v = (M[0] * (-MSIN(lat) * MCOS(lon)) + M[1] * (-MSIN(lat) * MSIN(lon)) + M[2] * (MCOS(lat))) / rho;
            const double w = W[e] / rho;
            const double kuv = sponge_rate(o, lev, Ruv), kw = sponge_rate(o, lev, Rw);
            double du, dv, dw;
            if (dmu) {
                du = -kuv * (u - sponge_mean_at(o, id, lev, 0));
                dv = -kuv * (v - sponge_mean_at(o, id, lev, 1));
            } else {
                du = -kuv * u;
                dv = -kuv * v;
            }
            const double vx = du * (-MSIN(lon)) + dv * (-MSIN(lat) * MCOS(lon));
            const double vy = du * (MCOS(lon)) + dv * (-MSIN(lat) * MSIN(lon));
            const double vz = dv * (MCOS(lat));
            dMh[e * 3 + 0] += vx * rho;
            dMh[e * 3 + 1] += vy * rho;
            dMh[e * 3 + 2] += vz * rho;
            dw = dmw ? -kw * (w - sponge_mean_at(o, id, lev, 2)) : -kw * (w);
            dW[e] += dw * rho;
            if (dmt) {
                const double kT = sponge_rate(o, lev, RT);
                const double temperature = pr[e] / Rd[e] / Rho[e];
                const double dT = -kT * (temperature - sponge_mean_at(o, id, lev, 3));
                Qh[e] += Rd[e] * Rho[e] * dT;
            }
        }
        for (int lev = 1; lev < nv; lev++) {
            const double xi = Alth[lev], xim = Alt[lev - 1], xip = Alt[lev];
            const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            dWh[id * (nv + 1) + lev] = dW[(long)id * nv + lev - 1] * intt + dW[(long)id * nv + lev] * intl;
        }
    }
}

void make_kd(Oracle &o) {
    // Horizontal diffusion / divergence-damping strengths (esp_initial.cu:1292-1321).
    // With the diffusive sponge (esp_initial.cu:1297-1321), its strength adds
    // to Kdh4 when its order is HyDiffOrder and fills Kdh2 otherwise.
    o.Kdh4.assign(o.nv, 0.0);
    o.Kdhz.assign(o.nv, 0.0);
    o.Kdh2.assign(o.nv, 0.0);
    const double top = o.g.Altitudeh[o.nv];
    for (int lev = 0; lev < o.nv; lev++) {
        double dbar = std::sqrt(2 * M_PI / 5) * o.p.A / (std::pow(2, o.p.glevel));
        o.Kdh4[lev] = (o.p.Diffc) * std::pow(dbar, 1.0 * o.p.HyDiffOrder) / o.p.timestep;
        o.Kdhz[lev] = (o.p.DivDampc) * std::pow(dbar, 4.) / o.p.timestep;
        if (o.p.DiffSponge) {
            const double n = o.g.Altitude[lev] / top, ns = o.p.ns_diff_sponge;
            const double ks = n > ns ? o.p.Dv_sponge * std::pow(std::sin(0.5 * M_PI * (n - ns) / (1.0 - ns)), 2) : 0.0;
            if (o.p.order_diff_sponge == o.p.HyDiffOrder)
                o.Kdh4[lev] += ks * std::pow(dbar, 1.0 * o.p.order_diff_sponge) / o.p.timestep;
            else
                o.Kdh2[lev] = ks * std::pow(dbar, 1.0 * o.p.order_diff_sponge) / o.p.timestep;
        }
    }
    // Vertical hyperdiffusion strength (esp_initial.cu:1327-1334).
    o.Kdv6.assign(o.nv, 0.0);
    for (int lev = 0; lev < o.nv; lev++) {
        const double dz = o.g.Altitudeh[lev + 1] - o.g.Altitudeh[lev];
        o.Kdv6[lev] = o.p.Diffc_v * std::pow(dz, 1.0 * o.p.VertHyDiffOrder) / o.p.timestep;
    }
}

} // namespace

double oracle_tree_sum(const double *d, long n) {
    const long B = 1024;
    long nb = (long)std::ceil(double(n) / double(2 * B));
    double out = 0.0;
    std::vector<double> buf(2 * B);
    for (long b = 0; b < nb; b++) {
        for (long j = 0; j < 2 * B; j++) {
            long idx = b * 2 * B + j;
            buf[j] = idx < n ? d[idx] : 0.0;
        }
        for (long stride = B; stride > 0; stride /= 2)
            for (long t = 0; t < stride; t++) buf[t] += buf[t + stride];
        out += buf[0];
    }
    return out;
}

struct oracle : Oracle {};

// Elementary functions as the restatement calls them (portable in the parity
// build, glibc otherwise): fn 0 = pow(x, y), 1 = log(x), 2 = sin(x), 3 = cos(x),
// 4 = exp(x).
extern "C" double oracle_math(int fn, double x, double y) {
    switch (fn) {
    case 0: return MPOW(x, y);
    case 1: return MLOG(x);
    case 2: return MSIN(x);
    case 3: return MCOS(x);
    default: return MEXP(x);
    }
}

extern "C" {

oracle *oracle_create(const oracle_grid *g, const oracle_params *p) {
    oracle *o = new oracle();
    o->g = *g;
    o->p = *p;
    o->P = g->point_num;
    o->nv = g->nv;
    o->nvi = g->nv + 1;
    o->nr = g->nr;
    const long N = (long)o->P * o->nv, Ni = (long)o->P * o->nvi;
    const char *sc[] = {"W", "Rho", "pressure", "temperature", "Rd", "Cp", "profx_Qheat", "Wk", "Rhok",
                        "pressurek", "Ws", "Rhos", "pressures", "h", "pt", "gtil", "SlowRho", "Slowpressure",
                        "diffw", "diffrh", "diffpr", "Sp", "Sd", "diffwv", "diffrv", "diffprv",
                        "epotential", "ekinetic", "Etotal"};
    const char *it[] = {"Wh", "profx_dWh", "Whk", "Whs", "hh", "pth", "gtilh", "SlowWh", "epotentialh", "ekinetich"};
    const char *vc[] = {"Mh", "profx_dMh", "Mhk", "Mhs", "Adv", "v", "SlowMh", "diffmh", "diffmv", "DivM", "divg_Mh"};
    for (auto n : sc) o->arr[n].assign(N, 0.0);
    for (auto n : it) o->arr[n].assign(Ni, 0.0);
    for (auto n : vc) o->arr[n].assign(3 * N, 0.0);
    o->arr["diff"].assign(6 * N, 0.0);
    for (long e = 0; e < N; e++) {
        o->arr["Rd"][e] = p->Rd;
        o->arr["Cp"][e] = p->Cp;
    }
    make_kd(*o);
    o->arr["insol"].assign(o->P, 0.0);
    o->arr["ASR"].assign(o->P, 0.0);
    o->arr["OLR"].assign(o->P, 0.0);
    o->arr["Kdh4"] = o->Kdh4;  // read-only copies for callers
    o->arr["Kdhz"] = o->Kdhz;
    o->arr["Kdv6"] = o->Kdv6;
    o->arr["Kdh2"] = o->Kdh2;
    o->arr["diff_sponge"].assign(6 * N, 0.0);
    return o;
}

void oracle_destroy(oracle *o) { delete o; }

long oracle_array_size(oracle *o, const char *name) {
    auto it = o->arr.find(name);
    return it == o->arr.end() ? -1 : (long)it->second.size();
}
int oracle_set(oracle *o, const char *name, const double *src, long n) {
    auto it = o->arr.find(name);
    if (it == o->arr.end() || (long)it->second.size() != n) return -1;
    std::memcpy(it->second.data(), src, n * sizeof(double));
    return 0;
}
int oracle_get(oracle *o, const char *name, double *dst, long n) {
    auto it = o->arr.find(name);
    if (it == o->arr.end() || (long)it->second.size() != n) return -1;
    std::memcpy(dst, it->second.data(), n * sizeof(double));
    return 0;
}
void oracle_thor(oracle *o) { thor(*o); }
void oracle_profx_held_suarez(oracle *o) { profx_hs(*o); }
int oracle_kernel(oracle *o, const char *name, int a0, int a1) {
    const std::string n(name);
    if (n == "advec_cori1") {
        advec_cori1(*o);
        advec_poles(*o);
    } else if (n == "advec_cori2") {
        advec_cori2(*o);
    } else if (n == "diffusion") {
        diffusion_pass(*o, a0, a1 != 0);
    } else if (n == "sponge_explicit") {  // a0: the RK stage
        if (!o->sponge_on) return -1;
        sponge_explicit(*o, a0);
    } else if (n == "vertical_diff") {
        vertical_diff(*o);
    } else if (n == "correct_horizontal") {
        correct_horizontal(*o);
    } else if (n == "divm") {
        divm_pass(*o, "Mhk", "Whk", a0 != 0);
    } else if (n == "slow_modes") {
        slow_modes(*o);
    } else if (n == "held_suarez") {
        held_suarez(*o);
    } else if (n == "boundary_layer") {  // on the current pressure
        if (!o->bl_on) return -1;
        std::vector<double> p0 = o->arr["pressure"];
        boundary_layer_rayleigh(*o, p0.data());
    } else if (n == "vertical_eq") {  // a0: the substep length in milliseconds
        o->capture_tri = a1 != 0;
        vertical_eq(*o, a0 * 1e-3);
        o->capture_tri = false;
    } else if (n == "tracer_diffusion") {  // a0: laststep
        if (!o->ntr) return -1;
        tracer_diffusion(*o, a0 != 0);
    } else if (n == "tracer_eq") {  // a0: the substep length in milliseconds
        if (!o->ntr) return -1;
        tracer_eq(*o, a0 * 1e-3);
    } else if (n == "temperature_h_pt_geff") {
        temperature_h_pt_geff(*o);
    } else if (n == "update_rk") {
        update_rk(*o);
    } else if (n == "update_rk2") {
        update_rk2(*o);
    } else if (n == "momentum_eq") {  // a0: the substep length in milliseconds
        momentum_eq(*o, a0 * 1e-3);
    } else if (n == "prepare_implicit_vertical") {
        prepare_implicit_vertical(*o);
    } else if (n == "density_pressure") {  // a0: the substep length in milliseconds
        density_pressure(*o, a0 * 1e-3);
    } else {
        return -1;
    }
    return 0;
}
void oracle_set_tracers(oracle *o, int ntr) {
    o->ntr = ntr;
    const size_t n = (size_t)o->P * o->nv * (ntr > 0 ? ntr : 0);
    for (const char *name : {"tracer", "tracerk", "tracers", "difftr", "trdiff"}) o->arr[name].assign(n, 0.0);
}
void oracle_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}
int oracle_get_threads(void) { return omp_get_max_threads(); }
void oracle_steps(oracle *o, int n) {
    for (int i = 0; i < n; i++) {
        profx_hs(*o);
        thor(*o);
    }
}
void oracle_set_newtonian(oracle *o, const double q[8]) { std::memcpy(o->nrq, q, sizeof(o->nrq)); }
void oracle_set_rt(oracle *o, const double q[11]) { std::memcpy(o->rtq, q, sizeof(o->rtq)); }
void oracle_set_sponge(oracle *o, const double q[11]) {
    o->sponge_on = q != nullptr;
    if (!q) return;
    std::memcpy(o->spq, q, sizeof(o->spq));
    zonal_table(*o, (int)q[3]);
}
int oracle_zonal_table(oracle *o, int *dst, long n, int *max_count) {
    if (n != (long)o->ztab.size()) return -1;
    std::memcpy(dst, o->ztab.data(), n * sizeof(int));
    *max_count = o->max_count;
    return 0;
}
void oracle_set_boundary_layer(oracle *o, const double q[2]) {
    o->bl_on = q != nullptr;
    if (q) std::memcpy(o->blq, q, sizeof(o->blq));
}
int oracle_profx(oracle *o, int forcing) {
    if (forcing == 0) return 0;
    if (o->sponge_imp()) sponge(*o);
    // the physics modules see the pressure ProfX started from (THOR computes
    // the new one after them, profx_driver.cu:403-445)
    std::vector<double> p0;
    if (o->bl_on) p0 = o->arr["pressure"];
    if (forcing == 1) profx_hs(*o);
    else if (forcing == 2) profx_nr(*o);
    else if (forcing == 3) profx_rt(*o);
    else return -1;
    if (o->bl_on) boundary_layer_rayleigh(*o, p0.data());
    return 0;
}
int oracle_run(oracle *o, int n, int forcing) {
    for (int i = 0; i < n; i++) {
        if (oracle_profx(o, forcing) != 0) return -1;
        thor(*o);
    }
    return 0;
}

// CalcMass / CalcTotEnergy / CalcAngMom (profx_globdiag.cu:45-212) + tree sums.
void oracle_globdiag(oracle *o, double out[5]) {
    const int nv = o->nv, P = o->P;
    const long N = (long)P * nv;
    std::vector<double> mass(N), en(N), amx(N), amy(N), amz(N);
    double *Rho = o->A_("Rho"), *Mh = o->A_("Mh"), *W = o->A_("W"), *T = o->A_("temperature"), *Cp = o->A_("Cp"),
           *Rd = o->A_("Rd");
    const double *Alt = o->g.Altitude, *Alth = o->g.Altitudeh, *ll = o->g.lonlat, *aT = o->g.areasT,
                 *fr = o->g.func_r;
    const double A = o->p.A, Omega = o->p.Omega, Gravit = o->p.Gravit;
    for (int id = 0; id < P; id++)
        for (int lev = 0; lev < nv; lev++) {
            long e = (long)id * nv + lev;
            double zup = Alth[lev + 1] + A, zlow = Alth[lev] + A, Vol;
            if (o->p.DeepModel) Vol = aT[id] / MPOW(A, 2) * (MPOW(zup, 3) - MPOW(zlow, 3)) / 3;
            else Vol = aT[id] * (zup - zlow);
            mass[e] = Rho[e] * Vol;
            double lat = ll[id * 2 + 1], lon = ll[id * 2];
            double wx = W[e] * MCOS(lat) * MCOS(lon), wy = W[e] * MCOS(lat) * MSIN(lon),
                   wz = W[e] * MSIN(lat);
            double Cv = Cp[e] - Rd[e];
            double Ek = 0.5
                        * ((Mh[e * 3 + 0] + wx) * (Mh[e * 3 + 0] + wx) + (Mh[e * 3 + 1] + wy) * (Mh[e * 3 + 1] + wy)
                           + (Mh[e * 3 + 2] + wz) * (Mh[e * 3 + 2] + wz))
                        / Rho[e];
            double Eint = Cv * T[e] * Rho[e];
            double Eg = Rho[e] * Gravit * Alt[lev];
            en[e] = (Ek + Eint + Eg) * Vol;
            double r = (A + Alt[lev]);
            double rx = r * fr[id * 3 + 0], ry = r * fr[id * 3 + 1], rz = r * fr[id * 3 + 2];
            double AMx = ry * Mh[e * 3 + 2] - rz * Mh[e * 3 + 1]
                         - Rho[e] * Omega * r * rz * MCOS(lat) * MCOS(lon);
            double AMy = -rx * Mh[e * 3 + 2] + rz * Mh[e * 3 + 0]
                         - Rho[e] * Omega * r * rz * MCOS(lat) * MSIN(lon);
            double AMz = rx * Mh[e * 3 + 1] - ry * Mh[e * 3 + 0]
                         + Rho[e] * Omega * r * r * MCOS(lat) * MCOS(lat);
            amx[e] = AMx * Vol;
            amy[e] = AMy * Vol;
            amz[e] = AMz * Vol;
        }
    out[0] = oracle_tree_sum(mass.data(), N);
    out[1] = oracle_tree_sum(en.data(), N);
    out[2] = oracle_tree_sum(amx.data(), N);
    out[3] = oracle_tree_sum(amy.data(), N);
    out[4] = oracle_tree_sum(amz.data(), N);
}

} // extern "C"
