// HIP kernels of the HE-VI dynamical-core step for gfx950 (FP64 path).
//
// Each kernel fuses one or more of THOR's kernels; inside a fused kernel the
// per-cell arithmetic is the reference's, operation for operation (compiled
// with -ffp-contract=off), so the FP64 path reproduces THOR's own results.
// Horizontal stencils read the six ring neighbours through a per-point index
// table instead of THOR's shared-memory halo tiles; stencil kernels map one
// thread to one point and loop over a chunk of LC levels so the per-point
// geometry stays in L1/L2 across the levels of a chunk.
#pragma once
#include "dyn_common.h"
#include "portable_math.h"

namespace thor {

#define LV(arr, lev) (arr)[(size_t)(lev) * S + p]
#define LVP(arr, lev, q) (arr)[(size_t)(lev) * S + (q)]
#define VEC(arr, c, lev, q) (arr)[((size_t)(c) * nv + (lev)) * S + (q)]

__device__ __forceinline__ double rscale_at(const Geo &g, int lev) {
    return g.deep ? g.A / (g.Alt[lev] + g.A) : 1.0;
}

// Global indices of the 7 stencil slots: 0 = self, 1..6 = ring pt1..pt6.
__device__ __forceinline__ void load_slots(const Geo &g, int p, int gi[7]) {
    const size_t S = g.S;
    gi[0] = p;
#pragma unroll
    for (int m = 0; m < 6; m++) gi[1 + m] = g.ring[(size_t)m * S + p];
}

// Laplacian of a[] around point p (Diffusion_Op, thor_diff.h:359-494):
// a[0] is the centre, a[1 + m] the value at ring neighbour m.
template <int NF>
__device__ __forceinline__ double laplacian(const Geo &g, int p, double rscale, const double a[7]) {
    const size_t S = g.S;
    double lap = 0.0, lap1 = 0, lap2 = 0, lap3 = 0, lap4, lap5, lap6;
    double lapx1 = 0, lapx2 = 0, lapy1 = 0, lapy2 = 0, lapz1 = 0, lapz2 = 0, AT1 = 0, AT2, astar;
    const double a0 = a[0];
#pragma unroll
    for (int j = 0; j < NF; j++) {
        const int jp1 = (j + 1) % NF, jp2 = (j + 2) % NF;
        const double a1 = a[1 + j], a2 = a[1 + jp1], a3 = a[1 + jp2];
        const double vdiff = 0.5 * rscale;
        if (j == 0) {
            const double aTr = g.areasTr[(size_t)j * S + p];
            AT1 = rscale / aTr;
            astar = (g.areas[(size_t)(j * 3 + 0) * S + p] * a0 + g.areas[(size_t)(j * 3 + 1) * S + p] * a1
                     + g.areas[(size_t)(j * 3 + 2) * S + p] * a2)
                    / aTr;
            lap1 = (0.5 * (a1 + a0) - astar);
            lap2 = (0.5 * (a1 + a2) - astar);
            lap3 = (0.5 * (a0 + a2) - astar);
        }
        const double aTr2 = g.areasTr[(size_t)jp1 * S + p];
        AT2 = rscale / aTr2;
        astar = (g.areas[(size_t)(jp1 * 3 + 0) * S + p] * a0 + g.areas[(size_t)(jp1 * 3 + 1) * S + p] * a2
                 + g.areas[(size_t)(jp1 * 3 + 2) * S + p] * a3)
                / aTr2;
        lap4 = (0.5 * (a0 + a2) - astar);
        lap5 = (0.5 * (a2 + a3) - astar);
        lap6 = (0.5 * (a0 + a3) - astar);
        const double *ti = g.nti, *te = g.nte;
        if (j == 0) {
            lapx1 = (-lap1 * ti[(size_t)(j * 3 + 0) * S + p] + lap2 * te[(size_t)(j * 3 + 0) * S + p]
                     + lap3 * ti[(size_t)(jp1 * 3 + 0) * S + p])
                    * AT1;
            lapy1 = (-lap1 * ti[(size_t)(j * 3 + 1) * S + p] + lap2 * te[(size_t)(j * 3 + 1) * S + p]
                     + lap3 * ti[(size_t)(jp1 * 3 + 1) * S + p])
                    * AT1;
            lapz1 = (-lap1 * ti[(size_t)(j * 3 + 2) * S + p] + lap2 * te[(size_t)(j * 3 + 2) * S + p]
                     + lap3 * ti[(size_t)(jp1 * 3 + 2) * S + p])
                    * AT1;
        } else {
            lapx1 = lapx2;
            lapy1 = lapy2;
            lapz1 = lapz2;
        }
        lapx2 = (-lap4 * ti[(size_t)(jp1 * 3 + 0) * S + p] + lap5 * te[(size_t)(jp1 * 3 + 0) * S + p]
                 + lap6 * ti[(size_t)(jp2 * 3 + 0) * S + p])
                * AT2;
        lapy2 = (-lap4 * ti[(size_t)(jp1 * 3 + 1) * S + p] + lap5 * te[(size_t)(jp1 * 3 + 1) * S + p]
                 + lap6 * ti[(size_t)(jp2 * 3 + 1) * S + p])
                * AT2;
        lapz2 = (-lap4 * ti[(size_t)(jp1 * 3 + 2) * S + p] + lap5 * te[(size_t)(jp1 * 3 + 2) * S + p]
                 + lap6 * ti[(size_t)(jp2 * 3 + 2) * S + p])
                * AT2;
        lap += ((lapx1 + lapx2) * g.noa[(size_t)(j * 3 + 0) * S + p]
                + (lapy1 + lapy2) * g.noa[(size_t)(j * 3 + 1) * S + p]
                + (lapz1 + lapz2) * g.noa[(size_t)(j * 3 + 2) * S + p])
               * vdiff;
    }
    return lap;
}

__device__ __forceinline__ double laplacian_nf(const Geo &g, int p, int nf, double rscale, const double a[7]) {
    return nf == 6 ? laplacian<6>(g, p, rscale, a) : laplacian<5>(g, p, rscale, a);
}

// Divergence-damping operator at point p (DivM_Op, thor_div.h:49-315 and the
// pole variant :317-535, whose vertical means weight the third corner with
// areas[.][1]). a[s][k]: momentum (pass 0) or divg (pass 1) at slot s;
// wl/wt: Wh at the layer's lower/upper interface (pass 0 only).
template <int NF>
__device__ __forceinline__ void divm_op(const Geo &g, int p, int lev, bool pole, bool laststep,
                                        const double a[7][3], const double wl[7], const double wt[7],
                                        double out[3]) {
    const size_t S = g.S;
    const double rs = rscale_at(g, lev);
    const double AT0 = 0.5 * rs;
    double lapx = 0.0, lapy = 0.0, lapz = 0.0;
    const double dz = laststep ? 0.0 : 1.0 / (g.Alth[lev] - g.Alth[lev + 1]);
    double a01x = 0, a01y = 0, a01z = 0, a12x = 0, a12y = 0, a12z = 0, AT1 = 0, lap1 = 0, lap2 = 0;
#define AR(j, k) g.areas[(size_t)((j)*3 + (k)) * S + p]
#define TI(j, k) g.nti[(size_t)((j)*3 + (k)) * S + p]
#define TE(j, k) g.nte[(size_t)((j)*3 + (k)) * S + p]
#pragma unroll
    for (int j = 0; j < NF; j++) {
        const int jp1 = (j + 1) % NF, jp2 = (j + 2) % NF;
        const int s1 = 1 + j, s2 = 1 + jp1, s3 = 1 + jp2;
        const double aT1 = g.areasTr[(size_t)j * S + p], aT2 = g.areasTr[(size_t)jp1 * S + p];
        double dwdz, dwdz2;
        if (!laststep) {
            const double w1 = pole ? AR(j, 1) : AR(j, 2);
            const double w2 = pole ? AR(jp1, 1) : AR(jp1, 2);
            const double meanwl = (AR(j, 0) * wl[0] + AR(j, 1) * wl[s1] + w1 * wl[s2]) / aT1;
            const double meanwt = (AR(j, 0) * wt[0] + AR(j, 1) * wt[s1] + w1 * wt[s2]) / aT1;
            const double meanwl2 = (AR(jp1, 0) * wl[0] + AR(jp1, 1) * wl[s2] + w2 * wl[s3]) / aT2;
            const double meanwt2 = (AR(jp1, 0) * wt[0] + AR(jp1, 1) * wt[s2] + w2 * wt[s3]) / aT2;
            dwdz = (meanwl - meanwt) * dz;
            dwdz2 = (meanwl2 - meanwt2) * dz;
        } else {
            dwdz = 0.0;
            dwdz2 = 0.0;
        }
        if (j == 0) {
            a01x = a[0][0] + a[s1][0];
            a01y = a[0][1] + a[s1][1];
            a01z = a[0][2] + a[s1][2];
            a12x = a[s1][0] + a[s2][0];
            a12y = a[s1][1] + a[s2][1];
            a12z = a[s1][2] + a[s2][2];
            AT1 = 0.5 * rs / aT1;
        }
        const double AT2 = 0.5 * rs / aT2;
        const double a02x = a[0][0] + a[s2][0], a02y = a[0][1] + a[s2][1], a02z = a[0][2] + a[s2][2];
        const double a23x = a[s2][0] + a[s3][0], a23y = a[s2][1] + a[s3][1], a23z = a[s2][2] + a[s3][2];
        const double a03x = a[0][0] + a[s3][0], a03y = a[0][1] + a[s3][1], a03z = a[0][2] + a[s3][2];
        if (j == 0) {
            lap1 = (-a01x * TI(j, 0) - a01y * TI(j, 1) - a01z * TI(j, 2) + a12x * TE(j, 0) + a12y * TE(j, 1)
                    + a12z * TE(j, 2) + a02x * TI(jp1, 0) + a02y * TI(jp1, 1) + a02z * TI(jp1, 2))
                   * AT1;
        } else {
            lap1 = lap2;
        }
        lap2 = (-a02x * TI(jp1, 0) - a02y * TI(jp1, 1) - a02z * TI(jp1, 2) + a23x * TE(jp1, 0)
                + a23y * TE(jp1, 1) + a23z * TE(jp1, 2) + a03x * TI(jp2, 0) + a03y * TI(jp2, 1)
                + a03z * TI(jp2, 2))
               * AT2;
        lapx += (lap1 + dwdz + lap2 + dwdz2) * g.noa[(size_t)(j * 3 + 0) * S + p] * AT0;
        lapy += (lap1 + dwdz + lap2 + dwdz2) * g.noa[(size_t)(j * 3 + 1) * S + p] * AT0;
        lapz += (lap1 + dwdz + lap2 + dwdz2) * g.noa[(size_t)(j * 3 + 2) * S + p] * AT0;
    }
#undef AR
#undef TI
#undef TE
    if (laststep) {
        const double sdiff = g.Kdhz[lev];
        const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
        const double lapr = fx * lapx + fy * lapy + fz * lapz;
        lapx = lapx - fx * lapr;
        lapy = lapy - fy * lapr;
        lapz = lapz - fz * lapr;
        out[0] = -lapx * sdiff;
        out[1] = -lapy * sdiff;
        out[2] = -lapz * sdiff;
    } else {
        out[0] = lapx;
        out[1] = lapy;
        out[2] = lapz;
    }
}

// DivM pass 0 (divergence) on momentum M and vertical momentum Wh at one cell.
__device__ __forceinline__ void divm_pass0(const Geo &g, int p, int lev, int nf, const int gi[7],
                                           const double *M, const double *Wh, double out[3]) {
    const size_t S = g.S;
    const int nv = g.nv;
    double a[7][3], wl[7], wt[7];
#pragma unroll
    for (int s = 0; s < 7; s++) {
        const int q = gi[s];
#pragma unroll
        for (int k = 0; k < 3; k++) a[s][k] = VEC(M, k, lev, q);
        wl[s] = LVP(Wh, lev, q);
        wt[s] = LVP(Wh, lev + 1, q);
    }
    const bool pole = p >= g.P - 2;
    if (nf == 6) divm_op<6>(g, p, lev, pole, false, a, wl, wt, out);
    else divm_op<5>(g, p, lev, pole, false, a, wl, wt, out);
}

// DivM pass 1 (gradient of the divergence, tangential, times Kdhz) at one cell.
__device__ __forceinline__ void divm_pass1(const Geo &g, int p, int lev, int nf, const int gi[7],
                                           const double *divg, double out[3]) {
    const size_t S = g.S;
    const int nv = g.nv;
    double a[7][3], wl[7], wt[7];
#pragma unroll
    for (int s = 0; s < 7; s++) {
        const int q = gi[s];
#pragma unroll
        for (int k = 0; k < 3; k++) a[s][k] = VEC(divg, k, lev, q);
        wl[s] = 0.0;
        wt[s] = 0.0;
    }
    const bool pole = p >= g.P - 2;
    if (nf == 6) divm_op<6>(g, p, lev, pole, true, a, wl, wt, out);
    else divm_op<5>(g, p, lev, pole, true, a, wl, wt, out);
}

// ===================================================================== K_A
// Per column: Compute_Temperature_H_Pt_Geff (thor_auxiliary.h:55-309, entropy
// form) and the vertical advection + Coriolis part of Compute_Advec_Cori2
// (thor_adv_cor.h:366-536), written to Adv; the horizontal part is added by
// k_adv_diff1.
__global__ __launch_bounds__(256) void k_column_prep(Geo g, Fields f) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const double Rd = g.Rd, Cp = g.Cp, P_Ref = g.P_Ref, Gravit = g.Gravit, A = g.A;
    const double *Alt = g.Alt, *Alth = g.Alth;
    const double *pk = f.pk, *Rho = f.Rhok, *Wh = f.Whk;
    {
        double rho = 0, rhoh, rhol = 0, rhop, pressure = 0, h = 0, hl = 0, pt = 0, ptl = 0, pl = 0, pp, dpdz;
        double gtilh = 0, gtilht, dz, dp_dz, Cv, xi, xim, xip, intt, intl, extr;
        double alt = 0, altl = 0, alth = 0, altht = 0;
        for (int lev = 0; lev < nv + 1; lev++) {
            if (lev < nv) {
                Cv = Cp - Rd;
                const double CvoCp = Cv / Cp;
                if (lev > 0) {
                    pl = pressure;
                    rhol = rho;
                    hl = h;
                    ptl = pt;
                    altl = alt;
                }
                pressure = LV(pk, lev);
                rho = LV(Rho, lev);
                const double temperature = pressure / (rho * Rd);
                LV(f.T, lev) = temperature;
                alt = Alt[lev];
                alth = Alth[lev];
                altht = Alth[lev + 1];
                h = Cv * temperature + pressure / rho;
                LV(f.h, lev) = h;
                pt = (P_Ref / (Rd * rho)) * MPOW(pressure / P_Ref, CvoCp);
                LV(f.pt, lev) = pt;
            }
            if (lev == 0) {
                extr = (-alt - Alt[lev + 1]) / (Alt[lev + 1] - alt);
                LV(f.hh, 0) = h;
                if (g.ghv)
                    pl = LV(pk, 1) - rho * Gravit * MPOW(A / (A + Alt[0]), 2) * (-alt - Alt[1]);
                else
                    pl = LV(pk, 1) - rho * Gravit * (-alt - Alt[1]);
                rhol = LV(Rho, 1) + (LV(Rho, 1) - rho) * extr;
                rhoh = 0.5 * (rho + rhol);
                dz = 2.0 * Alt[0];
                dpdz = (LV(pk, 0) - pl) / dz;
                LV(f.gtilh, 0) = -(1.0 / rhoh) * dpdz;
                gtilh = 0.0;
            } else if (lev == nv) {
                extr = (2 * Alth[nv] - alt - Alt[nv - 2]) / (Alt[nv - 1] - Alt[nv - 2]);
                LV(f.hh, nv) = h;
                rhop = LV(Rho, nv - 2) + (LV(Rho, nv - 1) - LV(Rho, nv - 2)) * extr;
                rhoh = 0.5 * (rho + rhop);
                LV(f.pth, nv) = LV(f.pt, nv - 2) + (LV(f.pt, nv - 1) - LV(f.pt, nv - 2)) * extr;
                LV(f.pth, 0) = LV(f.pt, 1) + (LV(f.pt, 1) - LV(f.pt, 0)) * (Alth[0] - Alt[1]) / (Alt[1] - Alt[0]);
                dz = 2.0 * (Alth[nv] - Alt[nv - 1]);
                if (g.ghv)
                    pp = LV(pk, nv - 2)
                         - LV(Rho, nv - 1) * Gravit * MPOW(A / (A + Alt[nv - 1]), 2)
                               * (2 * Alth[nv] - Alt[nv - 1] - Alt[nv - 2]);
                else
                    pp = LV(pk, nv - 2) - LV(Rho, nv - 1) * Gravit * (2 * Alth[nv] - Alt[nv - 1] - Alt[nv - 2]);
                dpdz = (pp - LV(pk, nv - 1)) / dz;
                LV(f.gtilh, nv) = -(1.0 / rhoh) * dpdz;
                gtilht = 0.0;
                xi = alt;
                xim = alth;
                xip = altht;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                LV(f.gtil, nv - 1) = gtilh * intt + gtilht * intl;
            } else {
                dz = alt - altl;
                dp_dz = (pressure - pl) / dz;
                xi = alth;
                xim = altl;
                xip = alt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                LV(f.hh, lev) = hl * intt + h * intl;
                LV(f.pth, lev) = ptl * intt + pt * intl;
                rhoh = rhol * intt + rho * intl;
                gtilht = -(LV(Wh, lev) / rhoh) * dp_dz;
                xi = alt;
                xim = alth;
                xip = altht;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                LV(f.gtil, lev - 1) = gtilh * intt + gtilht * intl;
                gtilh = gtilht;
                LV(f.gtilh, lev) = -(1.0 / rhoh) * dp_dz;
            }
        }
    }
    // Vertical advection of momentum and Coriolis (Compute_Advec_Cori2).
    {
        const double fr0 = g.fr[p], fr1 = g.fr[S + p], fr2 = g.fr[2 * S + p];
        const double Omega = g.Omega;
        auto vel = [&](int lev, double v[3]) {
            const double rho = LV(Rho, lev), w = LV(f.Wk, lev);
            const double fr[3] = {fr0, fr1, fr2};
#pragma unroll
            for (int c = 0; c < 3; c++) v[c] = VEC(f.Mhk, c, lev, p) / rho + (w / rho) * fr[c];
        };
        double altht = 0, althl = 0, altt = 0, alt = 0, altl = 0, wht = 0, whl = 0;
        double vl[3] = {0, 0, 0}, vc[3] = {0, 0, 0}, vt[3] = {0, 0, 0};
        double dvl[3], dvt[3], intt, intl, xi, xim, xip;
        for (int lev = 0; lev < nv; lev++) {
            if (lev == 0) {
                dvl[0] = dvl[1] = dvl[2] = 0.0;
                altht = Alth[lev + 1];
                althl = Alth[lev];
                alt = Alt[lev];
                altt = Alt[lev + 1];
                vel(lev, vc);
                vel(lev + 1, vt);
                wht = LV(Wh, lev + 1);
                xi = altht;
                xim = alt;
                xip = altt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                for (int c = 0; c < 3; c++) dvt[c] = (vc[c] * intt + vt[c] * intl) * wht;
            } else if (lev == nv - 1) {
                dvt[0] = dvt[1] = dvt[2] = 0.0;
                xi = althl;
                xim = altl;
                xip = alt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                for (int c = 0; c < 3; c++) dvl[c] = (vl[c] * intt + vc[c] * intl) * whl;
            } else {
                xi = althl;
                xim = altl;
                xip = alt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                for (int c = 0; c < 3; c++) dvl[c] = (vl[c] * intt + vc[c] * intl) * whl;
                xi = altht;
                xim = alt;
                xip = altt;
                intt = (xi - xip) / (xim - xip);
                intl = (xi - xim) / (xip - xim);
                for (int c = 0; c < 3; c++) dvt[c] = (vc[c] * intt + vt[c] * intl) * wht;
            }
            double r2p, r2m, r2l;
            if (g.deep) {
                r2p = MPOW(altht + A, 2.0);
                r2m = MPOW(alt + A, 2.0);
                r2l = MPOW(althl + A, 2.0);
            } else {
                r2p = r2m = r2l = 1.0;
            }
            const double dzi = 1.0 / ((althl - altht) * r2m);
            double dav[3];
            for (int c = 0; c < 3; c++) dav[c] = (dvl[c] * r2l - dvt[c] * r2p) * dzi;
            const double rho = LV(Rho, lev);
            double Cx, Cy, Cz;
            if (g.deep) {
                Cx = -2.0 * Omega * vc[1] * rho;
                Cy = 2.0 * Omega * vc[0] * rho;
                Cz = 0.0;
            } else {
                Cx = 2 * Omega * (vc[2] * fr2 * fr1 - vc[1] * MPOW(fr2, 2)) * rho;
                Cy = -2 * Omega * (vc[2] * fr2 * fr0 - vc[0] * MPOW(fr2, 2)) * rho;
                Cz = 2 * Omega * (vc[1] * fr2 * fr0 - vc[0] * fr2 * fr1) * rho;
            }
            VEC(f.Adv, 0, lev, p) = dav[0] + Cx;
            VEC(f.Adv, 1, lev, p) = dav[1] + Cy;
            VEC(f.Adv, 2, lev, p) = dav[2] + Cz;
            if (lev < nv - 1) {
                althl = altht;
                altht = Alth[lev + 2];
                altl = alt;
                alt = altt;
                if (lev < nv - 2) altt = Alt[lev + 2];
                whl = wht;
                if (lev < nv - 2) wht = LV(Wh, lev + 2);
                for (int c = 0; c < 3; c++) {
                    vl[c] = vc[c];
                    vc[c] = vt[c];
                }
                if (lev < nv - 2) vel(lev + 2, vt);
            }
        }
    }
}

// ===================================================================== K_B
// Horizontal advection (Compute_Advec_Cori1 / _Poles, thor_adv_cor.h:51-362)
// added to the vertical part, and the first Laplacian pass of the 4th-order
// hyperdiffusion (Diffusion_Op stepnum 0, thor_diff.h:54-522) of rho, u, v,
// w = M/rho, W/rho and Rd*T.
template <int LC>
__global__ __launch_bounds__(256) void k_adv_diff1(Geo g, Fields f, int hydiff) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const int l0 = blockIdx.y * LC, l1 = min(l0 + LC, nv);
    const int nf = g.nfc[p];
    int gi[7];
    load_slots(g, p, gi);
    for (int lev = l0; lev < l1; lev++) {
        const double rs = rscale_at(g, lev);
        double M[7][3], v[7][3], rho[7], W[7];
#pragma unroll
        for (int s = 0; s < 7; s++) {
            const int q = gi[s];
            rho[s] = LVP(f.Rhok, lev, q);
            W[s] = LVP(f.Wk, lev, q);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                M[s][k] = VEC(f.Mhk, k, lev, q);
                v[s][k] = M[s][k] / rho[s] + (W[s] / rho[s]) * g.fr[(size_t)k * S + q];
            }
        }
        double nf3[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 3; kk++) {
            double d[7];
#pragma unroll
            for (int s = 0; s < 7; s++) d[s] = g.div[(size_t)(s * 3 + kk) * S + p];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double sum = d[0] * v[0][c] * M[0][kk];
#pragma unroll
                for (int s = 1; s < 7; s++) sum = sum + d[s] * v[s][c] * M[s][kk];
                nf3[c] += rs * sum;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) VEC(f.Adv, c, lev, p) = nf3[c] + VEC(f.Adv, c, lev, p);
        if (hydiff) {
            double a[7];
#pragma unroll
            for (int var = 0; var < 6; var++) {
#pragma unroll
                for (int s = 0; s < 7; s++) {
                    if (var == 0) a[s] = rho[s];
                    else if (var <= 3) a[s] = M[s][var - 1] / rho[s];
                    else if (var == 4) a[s] = W[s] / rho[s];
                    else a[s] = g.Rd * LVP(f.T, lev, gi[s]);
                }
                VEC(f.diff, var, lev, p) = laplacian_nf(g, p, nf, rs, a);
            }
        }
    }
}

// ===================================================================== K_C
// Last hyperdiffusion pass (Diffusion_Op laststep) + Correct_Horizontal
// (thor_diff.h:882-915), and divergence-damping pass 0 on Mhk/Whk
// (DivM_Op laststep = false).
template <int LC>
__global__ __launch_bounds__(256) void k_diff2_divm0(Geo g, Fields f, int hydiff, int divdamp, double sign) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const int l0 = blockIdx.y * LC, l1 = min(l0 + LC, nv);
    const int nf = g.nfc[p];
    const bool pole = p >= g.P - 2;
    int gi[7];
    load_slots(g, p, gi);
    for (int lev = l0; lev < l1; lev++) {
        if (hydiff) {
            const double rs = rscale_at(g, lev);
            const double sdiff = -g.Kdh4[lev] * sign;
            double rho[7];
#pragma unroll
            for (int s = 0; s < 7; s++) rho[s] = LVP(f.Rhok, lev, gi[s]);
            double out[6];
#pragma unroll
            for (int var = 0; var < 6; var++) {
                double a[7];
#pragma unroll
                for (int s = 0; s < 7; s++) {
                    const double dv = VEC(f.diff, var, lev, gi[s]);
                    a[s] = var == 0 ? dv : dv * rho[s];
                }
                double lap = laplacian_nf(g, p, nf, rs, a);
                lap = pole ? sdiff * lap : lap * sdiff;
                out[var] = lap;
            }
            LV(f.diffrh, lev) = out[0];
            LV(f.diffw, lev) = out[4];
            LV(f.diffpr, lev) = out[5];
            const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
            const double dmhr = fx * out[1] + fy * out[2] + fz * out[3];
            VEC(f.diffmh, 0, lev, p) = out[1] + -fx * dmhr;
            VEC(f.diffmh, 1, lev, p) = out[2] + -fy * dmhr;
            VEC(f.diffmh, 2, lev, p) = out[3] + -fz * dmhr;
        }
        if (divdamp) {
            double o3[3];
            divm_pass0(g, p, lev, nf, gi, f.Mhk, f.Whk, o3);
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(f.divg, k, lev, p) = o3[k];
        }
    }
}

// ===================================================================== K_D
// Divergence-damping pass 1 on the stage values + Compute_Slow_Modes / _Poles
// (thor_slowmodes.h:48-683, entropy form) + UpdateRK (thor_auxiliary.h:311-418)
// for rk > 0, or the zero start of the deviations for rk = 0
// (thor_driver.cu:104-108).
template <int LC>
__global__ __launch_bounds__(256) void k_slow(Geo g, Fields f, int divdamp, int rk) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const int l0 = blockIdx.y * LC, l1 = min(l0 + LC, nv);
    const int nf = g.nfc[p];
    const bool pole = p >= g.P - 2;
    int gi[7];
    load_slots(g, p, gi);
    const double A = g.A, Gravit = g.Gravit, Rd = g.Rd, Cv = g.Cp - g.Rd;
    const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
    const double *Alt = g.Alt, *Alth = g.Alth;
    for (int lev = l0; lev < l1; lev++) {
        double DivM[3] = {0.0, 0.0, 0.0};
        if (divdamp) divm_pass1(g, p, lev, nf, gi, f.divg, DivM);
        double advx = VEC(f.Adv, 0, lev, p), advy = VEC(f.Adv, 1, lev, p), advz = VEC(f.Adv, 2, lev, p);
        const double rho = LV(f.Rhok, lev);
        double advrl = 0, rhol = 0, altl = 0, pressurel = 0;
        if (lev != 0) {
            const double ax = VEC(f.Adv, 0, lev - 1, p), ay = VEC(f.Adv, 1, lev - 1, p),
                         az = VEC(f.Adv, 2, lev - 1, p);
            advrl = ax * fx + ay * fy + az * fz;
            rhol = LV(f.Rhok, lev - 1);
            altl = Alt[lev - 1];
            pressurel = LV(f.pk, lev - 1);
        }
        double advr = advx * fx + advy * fy + advz * fz;
        advx += -advr * fx;
        advy += -advr * fy;
        advz += -advr * fz;
        const double altht = Alth[lev + 1], althl = Alth[lev], alt = Alt[lev];
        const double M0x = VEC(f.Mhk, 0, lev, p), M0y = VEC(f.Mhk, 1, lev, p), M0z = VEC(f.Mhk, 2, lev, p);
        if (!g.nonhydro) {
            advr = (MPOW(M0x, 2) + MPOW(M0y, 2) + MPOW(M0z, 2)) / (rho * (alt + A));
            if (lev != 0) {
                const double m1x = VEC(f.Mhk, 0, lev - 1, p), m1y = VEC(f.Mhk, 1, lev - 1, p),
                             m1z = VEC(f.Mhk, 2, lev - 1, p);
                advrl = (MPOW(m1x, 2) + MPOW(m1y, 2) + MPOW(m1z, 2)) / (rhol * (altl + A));
            }
        }
        const double dz = altht - althl;
        double r2p, r2m, r2l, rscale, dr2dz;
        if (g.deep) {
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
        double pv[7], hv[7];
#pragma unroll
        for (int s = 0; s < 7; s++) {
            pv[s] = LVP(f.pk, lev, gi[s]);
            hv[s] = LVP(f.h, lev, gi[s]);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double mv[7];
#pragma unroll
            for (int s = 0; s < 7; s++) mv[s] = VEC(f.Mhk, k, lev, gi[s]);
            double sg = g.grad[(size_t)(0 * 3 + k) * S + p] * pv[0];
            double sr = g.div[(size_t)(0 * 3 + k) * S + p] * mv[0];
            double sp = g.div[(size_t)(0 * 3 + k) * S + p] * mv[0] * hv[0];
#pragma unroll
            for (int s = 1; s < 7; s++) {
                const double gd = g.grad[(size_t)(s * 3 + k) * S + p], dd = g.div[(size_t)(s * 3 + k) * S + p];
                sg = sg + gd * pv[s];
                sr = sr + dd * mv[s];
                sp = sp + dd * mv[s] * hv[s];
            }
            nflxv[k] = rscale * sg;
            nflxr += rscale * sr;
            nflxp += rscale * sp;
        }
        double whl, wht, hhl, hht;
        if (lev == 0) {
            whl = 0.0;
            wht = LV(f.Whk, lev + 1);
            hhl = 0.0;
            hht = LV(f.hh, lev + 1);
        } else {
            whl = LV(f.Whk, lev);
            wht = LV(f.Whk, lev + 1);
            hhl = LV(f.hh, lev);
            hht = LV(f.hh, lev + 1);
        }
        const double dwdz = pole ? (wht * r2p - whl * r2l) / (dz * r2m) : (wht * r2p - whl * r2l) / (dr2dz);
        const double dwhdz = (wht * r2p * hht - whl * r2l * hhl) / (dz * r2m);
        const double dpr = nflxv[0] * fx + nflxv[1] * fy + nflxv[2] * fz;
        const double fv[3] = {fx, fy, fz}, ad[3] = {advx, advy, advz};
#pragma unroll
        for (int k = 0; k < 3; k++)
            VEC(f.SlowMh, k, lev, p) = -(nflxv[k] - dpr * fv[k]) - ad[k] + DivM[k] + VEC(f.diffmh, k, lev, p) + 0.0
                                       + VEC(f.dMh, k, lev, p);
        if (lev == 0) {
            LV(f.SlowWh, 0) = 0.0;
            LV(f.SlowWh, nv) = 0.0;
        } else {
            const double xi = althl, xim = altl, xip = alt;
            const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            double swr;
            if (g.deep) {
                if (g.ghv)
                    swr = (advrl + rhol * Gravit * MPOW(A / (A + altl), 2)) * intt
                          + (advr + rho * Gravit * MPOW(A / (A + alt), 2)) * intl;
                else
                    swr = (advrl + rhol * Gravit) * intt + (advr + rho * Gravit) * intl;
            } else {
                if (g.ghv)
                    swr = (rhol * Gravit * MPOW(A / (A + altl), 2)) * intt + (rho * Gravit * MPOW(A / (A + alt), 2)) * intl;
                else
                    swr = (rhol * Gravit) * intt + (rho * Gravit) * intl;
            }
            if (g.nonhydro) swr += (-LV(f.diffw, lev - 1) - 0.0) * intt + (-LV(f.diffw, lev) - 0.0) * intl;
            const double dzl = alt - altl;
            const double dpdz = (pv[0] - pressurel) / dzl;
            swr += dpdz;
            swr -= LV(f.dWh, lev);
            LV(f.SlowWh, lev) = -swr;
        }
        nflxr += dwdz;
        LV(f.SlowRho, lev) = -nflxr + LV(f.diffrh, lev) + 0.0;
        const double r = 1.0 / rho;
        const double vgp = (nflxv[0] - dpr * fx) * M0x * r + (nflxv[1] - dpr * fy) * M0y * r
                           + (nflxv[2] - dpr * fz) * M0z * r;
        double sp = (Rd / Cv) * (-nflxp - LV(f.gtil, lev) - dwhdz + vgp + LV(f.Qheat, lev));
        sp += LV(f.diffpr, lev) + 0.0;
        LV(f.Slowp, lev) = sp;

        // Deviations at the start of the stage.
        if (rk == 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(f.Mhs, k, lev, p) = 0.0;
            LV(f.Rhos, lev) = 0.0;
            LV(f.ps, lev) = 0.0;
            LV(f.Whs, lev) = 0.0;
            if (lev == nv - 1) LV(f.Whs, nv) = 0.0;
        } else {
            const double Mx = VEC(f.Mh, 0, lev, p) - M0x, My = VEC(f.Mh, 1, lev, p) - M0y,
                         Mz = VEC(f.Mh, 2, lev, p) - M0z;
            const double aux = Mx * fx + My * fy + Mz * fz;
            VEC(f.Mhs, 0, lev, p) = Mx - aux * fx;
            VEC(f.Mhs, 1, lev, p) = My - aux * fy;
            VEC(f.Mhs, 2, lev, p) = Mz - aux * fz;
            LV(f.Rhos, lev) = LV(f.Rho, lev) - rho;
            LV(f.ps, lev) = LV(f.p, lev) - pv[0];
            LV(f.Whs, lev) = lev == 0 ? 0.0 : LV(f.Wh, lev) - LV(f.Whk, lev);
            if (lev == nv - 1) LV(f.Whs, nv) = 0.0;
        }
    }
}

// ===================================================================== K_E
// Divergence-damping pass 0 on the deviations Mhs / Whs.
template <int LC>
__global__ __launch_bounds__(256) void k_divm0(Geo g, Fields f) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const int l0 = blockIdx.y * LC, l1 = min(l0 + LC, nv);
    const int nf = g.nfc[p];
    int gi[7];
    load_slots(g, p, gi);
    for (int lev = l0; lev < l1; lev++) {
        double o3[3];
        divm_pass0(g, p, lev, nf, gi, f.Mhs, f.Whs, o3);
#pragma unroll
        for (int k = 0; k < 3; k++) VEC(f.divg, k, lev, p) = o3[k];
    }
}

// ===================================================================== K_F
// Divergence-damping pass 1 on the deviations + Momentum_Eq / _Poles
// (thor_fastmodes.h:54-253).
template <int LC>
__global__ __launch_bounds__(256) void k_momentum(Geo g, Fields f, int divdamp, double dt) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const int l0 = blockIdx.y * LC, l1 = min(l0 + LC, nv);
    const int nf = g.nfc[p];
    int gi[7];
    load_slots(g, p, gi);
    const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
    for (int lev = l0; lev < l1; lev++) {
        double DivM[3] = {0.0, 0.0, 0.0};
        if (divdamp) divm_pass1(g, p, lev, nf, gi, f.divg, DivM);
        const double rs = rscale_at(g, lev);
        double pv[7];
#pragma unroll
        for (int s = 0; s < 7; s++) pv[s] = LVP(f.ps, lev, gi[s]);
        double nf3[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double s0 = g.grad[(size_t)(0 * 3 + k) * S + p] * pv[0];
#pragma unroll
            for (int s = 1; s < 7; s++) s0 = s0 + g.grad[(size_t)(s * 3 + k) * S + p] * pv[s];
            nf3[k] = rs * s0;
        }
        double Mx = (-nf3[0] + VEC(f.SlowMh, 0, lev, p) + DivM[0]) * dt;
        double My = (-nf3[1] + VEC(f.SlowMh, 1, lev, p) + DivM[1]) * dt;
        double Mz = (-nf3[2] + VEC(f.SlowMh, 2, lev, p) + DivM[2]) * dt;
        const double vr = Mx * fx + My * fy + Mz * fz;
        Mx += -vr * fx;
        My += -vr * fy;
        Mz += -vr * fz;
        VEC(f.Mhs, 0, lev, p) += Mx;
        VEC(f.Mhs, 1, lev, p) += My;
        VEC(f.Mhs, 2, lev, p) += Mz;
    }
}

// ===================================================================== K_G1
// Per column: Prepare_Implicit_Vertical (thor_vertical_int.h:566-771) for the
// source terms Sp, Sd and Vertical_Eq (thor_vertical_int.h:52-563): the
// tridiagonal solve for Whs at interfaces 1..nv-1 with the modify-in-place
// Thomas algorithm; cc/dd live in LDS.
constexpr int kThomasBlock = 64;

__device__ __forceinline__ void sources_at(const Geo &g, const Fields &f, int p, int lev, const int gi[7],
                                           double &Sp, double &Sd) {
    const size_t S = g.S;
    const int nv = g.nv;
    const double RoC = g.Rd / (g.Cp - g.Rd);
    const double rs = rscale_at(g, lev);
    double nflxr = 0.0, nflxp = 0.0, hv[7];
#pragma unroll
    for (int s = 0; s < 7; s++) hv[s] = LVP(f.h, lev, gi[s]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double mv[7];
#pragma unroll
        for (int s = 0; s < 7; s++) mv[s] = VEC(f.Mhs, k, lev, gi[s]);
        double sr = g.div[(size_t)(0 * 3 + k) * S + p] * mv[0];
        double sp = g.div[(size_t)(0 * 3 + k) * S + p] * mv[0] * hv[0];
#pragma unroll
        for (int s = 1; s < 7; s++) {
            const double dd = g.div[(size_t)(s * 3 + k) * S + p];
            sr = sr + dd * mv[s];
            sp = sp + dd * mv[s] * hv[s];
        }
        nflxr -= rs * sr;
        nflxp -= (RoC)*rs * sp;
    }
    Sp = nflxp + LV(f.Slowp, lev);
    Sd = nflxr + LV(f.SlowRho, lev);
}

__global__ __launch_bounds__(kThomasBlock) void k_thomas(Geo g, Fields f, double deltat) {
    extern __shared__ double lds[];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv, nvi = nv + 1;
    double *cc = lds + (size_t)threadIdx.x * nvi;
    double *dd = lds + (size_t)kThomasBlock * nvi + (size_t)threadIdx.x * nvi;
    int gi[7];
    load_slots(g, p, gi);
    const double A = g.A, Gravit = g.Gravit, Rd = g.Rd, Cp = g.Cp;
    const double *Alt = g.Alt, *Alth = g.Alth;
    const bool ghv = g.ghv, nonh = g.nonhydro;
    double altht = 0, althl = 0, alth = 0, alt = 0, altl = 0, hp = 0, h = 0, hm = 0, gp = 0, gg = 0, gm = 0;
    double Sp = 0, Spl = 0, pr = 0, pl = 0, r = 0, rl = 0, Sdl = 0, Sd = 0;
    double CRddl = 0, CRddu = 0, GCoRl = 0, GCoRu = 0;
    const double CRdd_c = (Cp - Rd) / (Rd * deltat * deltat);
    auto gcor = [&](double altx) {
        return ghv ? Gravit * MPOW(A / (A + altx), 2) * (Cp - Rd) / Rd : Gravit * (Cp - Rd) / Rd;
    };
    for (int lev = 1; lev < nv; lev++) {
        if (lev == 1) {
            altht = Alth[lev + 1];
            althl = Alth[lev - 1];
            alth = Alth[lev];
            alt = Alt[lev];
            altl = Alt[lev - 1];
            hp = LV(f.hh, lev + 1);
            h = LV(f.hh, lev);
            hm = LV(f.hh, lev - 1);
            gp = LV(f.gtilh, lev + 1);
            gg = LV(f.gtilh, lev);
            gm = LV(f.gtilh, lev - 1);
            sources_at(g, f, p, lev - 1, gi, Spl, Sdl);
            sources_at(g, f, p, lev, gi, Sp, Sd);
            pr = LV(f.ps, lev);
            pl = LV(f.ps, lev - 1);
            rl = LV(f.Rhos, lev - 1);
            r = LV(f.Rhos, lev);
            CRddl = CRdd_c;
            CRddu = CRdd_c;
            GCoRl = gcor(altl);
            GCoRu = gcor(alt);
        }
        double aa, bb, C0, intt, intl, inttm, xi, xim, xip;
        const double dzp = 1.0 / (altht - alth), dzh = 1.0 / (alt - altl), dzm = 1.0 / (alth - althl);
        const double dzph = dzp * dzh, dzmh = dzm * dzh;
        const double CRdd = (CRddl * (alt - alth) + CRddu * (alth - altl)) / (alt - altl);
        const double GCoR = (GCoRl * (alt - alth) + GCoRu * (alth - altl)) / (alt - altl);
        if (g.deep) {
            hp = hp * MPOW(altht + A, 2);
            h = h * MPOW(alth + A, 2);
            hm = hm * MPOW(althl + A, 2);
            const double or2 = 1 / MPOW(alth + A, 2);
            const double tor3 = 2 / MPOW(alth + A, 3);
            xi = alt;
            xim = alth;
            intl = (xi - xim) * dzp * dzh;
            xi = altl;
            xip = alth;
            inttm = -(xi - xip) * dzm * dzh;
            cc[lev] = -dzph * or2 * hp - inttm * (gp + GCoR - tor3 * hp);
            if (nonh) bb = CRdd + (dzph + dzmh) * or2 * h + (intl - inttm) * (gg + GCoR - tor3 * h);
            else bb = (dzph + dzmh) * or2 * h + (intl - inttm) * (gg + GCoR - tor3 * h);
            aa = -dzmh * or2 * hm + intl * (gm + GCoR - tor3 * hm);
        } else {
            xi = alt;
            xim = alth;
            intl = (xi - xim) * dzp * dzh;
            xi = altl;
            xip = alth;
            inttm = -(xi - xip) * dzm * dzh;
            cc[lev] = -dzph * hp - inttm * (gp + GCoR);
            if (nonh) bb = CRdd + (dzph + dzmh) * h + (intl - inttm) * (gg + GCoR);
            else bb = (dzph + dzmh) * h + (intl - inttm) * (gg + GCoR);
            aa = -dzmh * hm + intl * (gm + GCoR);
        }
        const double dSpdz = (Sp - Spl) * dzh, dPdz = (pr - pl) * dzh;
        xi = alth;
        xim = altl;
        xip = alt;
        intt = (xi - xip) / (xim - xip);
        intl = (xi - xim) / (xip - xim);
        const double rhohs = rl * intt + r * intl;
        const double Sdh = Sdl * intt + Sd * intl;
        const double srh = LV(f.SlowWh, lev);
        if (!nonh) {
            if (ghv)
                C0 = (MPOW(deltat, 2.0) * dSpdz + MPOW(deltat, 2.0) * Gravit * MPOW(A / (A + alt), 2) * Sdh
                      + Gravit * MPOW(A / (A + alt), 2) * deltat * rhohs + deltat * dPdz - deltat * srh)
                     * (CRdd);
            else
                C0 = (MPOW(deltat, 2.0) * dSpdz + MPOW(deltat, 2.0) * Gravit * Sdh + Gravit * deltat * rhohs
                      + deltat * dPdz - deltat * srh)
                     * (CRdd);
        } else {
            const double whs = LV(f.Whs, lev);
            if (ghv)
                C0 = (-whs + MPOW(deltat, 2.0) * dSpdz + MPOW(deltat, 2.0) * Gravit * MPOW(A / (A + alt), 2) * Sdh
                      + Gravit * MPOW(A / (A + alt), 2) * deltat * rhohs + deltat * dPdz - deltat * srh)
                     * (CRdd);
            else
                C0 = (-whs + MPOW(deltat, 2.0) * dSpdz + MPOW(deltat, 2.0) * Gravit * Sdh
                      + Gravit * deltat * rhohs + deltat * dPdz - deltat * srh)
                     * (CRdd);
        }
        if (lev < nv - 1) {
            althl = alth;
            alth = altht;
            altht = Alth[lev + 2];
            altl = alt;
            alt = Alt[lev + 1];
            if (g.deep) {
                hm = LV(f.hh, lev);
                h = LV(f.hh, lev + 1);
            } else {
                hm = h;
                h = hp;
            }
            hp = LV(f.hh, lev + 2);
            gm = gg;
            gg = gp;
            gp = LV(f.gtilh, lev + 2);
            Spl = Sp;
            Sdl = Sd;
            sources_at(g, f, p, lev + 1, gi, Sp, Sd);
            pl = pr;
            pr = LV(f.ps, lev + 1);
            rl = r;
            r = LV(f.Rhos, lev + 1);
            CRddl = CRddu;
            CRddu = CRdd_c;
            GCoRl = GCoRu;
            GCoRu = gcor(alt);
        }
        dd[lev] = -C0;
        if (lev == 1) {
            cc[1] = cc[1] / bb;
            dd[1] = dd[1] / bb;
        } else {
            const double t = 1.0 / (bb - cc[lev - 1] * aa);
            cc[lev] *= t;
            dd[lev] = (dd[lev] - dd[lev - 1] * aa) * t;
        }
    }
    LV(f.Whs, nv) = 0.0;
    LV(f.Whs, 0) = 0.0;
    double wn = dd[nv - 1];
    LV(f.Whs, nv - 1) = wn;
    double wup = 0.0;
    for (int lev = nvi - 2; lev > 0; lev--) {
        const double w = (-cc[lev] * wup + dd[lev]);
        LV(f.Whs, lev) = w;
        wup = w;
    }
}

// ===================================================================== K_G2
// Density_Pressure_Eqs / _Poles (thor_fastmodes.h:255-847, entropy form) and,
// unless this is the stage's last substep, divergence-damping pass 0 on the
// updated deviations for the next substep.
template <int LC>
__global__ __launch_bounds__(256) void k_density_pressure(Geo g, Fields f, double dt, int next_divm) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const int l0 = blockIdx.y * LC, l1 = min(l0 + LC, nv);
    const int nf = g.nfc[p];
    const bool pole = p >= g.P - 2;
    int gi[7];
    load_slots(g, p, gi);
    const double A = g.A, P_Ref = g.P_Ref, Rd = g.Rd, Cp = g.Cp, Cv = Cp - Rd;
    for (int lev = l0; lev < l1; lev++) {
        const double altht = g.Alth[lev + 1], althl = g.Alth[lev], dz = altht - althl;
        double r2p, r2m, r2l, rscale, dr2dz;
        if (g.deep) {
            const double alt = g.Alt[lev];
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
        double nflxr = 0.0, nflxpt = 0.0, ptv[7];
#pragma unroll
        for (int s = 0; s < 7; s++) ptv[s] = LVP(f.pt, lev, gi[s]);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double vv[7], v1[7];
#pragma unroll
            for (int s = 0; s < 7; s++) {
                vv[s] = VEC(f.Mhs, k, lev, gi[s]);
                v1[s] = vv[s] + VEC(f.Mhk, k, lev, gi[s]);
            }
            double sr = g.div[(size_t)(0 * 3 + k) * S + p] * vv[0];
            double sp = g.div[(size_t)(0 * 3 + k) * S + p] * v1[0] * ptv[0];
#pragma unroll
            for (int s = 1; s < 7; s++) {
                const double dd = g.div[(size_t)(s * 3 + k) * S + p];
                sr = sr + dd * vv[s];
                sp = sp + dd * v1[s] * ptv[s];
            }
            nflxr += rscale * sr;
            nflxpt += rscale * sp;
        }
        double whl, wht, whl2, wht2, phl, pht;
        if (lev == 0) {
            whl = 0.0;
            wht = LV(f.Whs, lev + 1);
            whl2 = 0.0;
            wht2 = LV(f.Whs, lev + 1) + LV(f.Whk, lev + 1);
            phl = 0.0;
            pht = LV(f.pth, lev + 1);
        } else {
            whl = LV(f.Whs, lev);
            wht = LV(f.Whs, lev + 1);
            whl2 = LV(f.Whs, lev) + LV(f.Whk, lev);
            wht2 = LV(f.Whs, lev + 1) + LV(f.Whk, lev + 1);
            phl = LV(f.pth, lev);
            pht = LV(f.pth, lev + 1);
        }
        const double dwdz = pole ? (wht * r2p - whl * r2l) / (dz * r2m) : (wht * r2p - whl * r2l) / (dr2dz);
        const double dwptdz = (wht2 * pht * r2p - whl2 * phl * r2l) / (dz * r2m);
        double aux = -(nflxpt + dwptdz) * dt;
        const double rhos = LV(f.Rhos, lev), pss = LV(f.ps, lev), pkv = LV(f.pk, lev);
        const double r = LV(f.Rhok, lev) + rhos;
        nflxr += dwdz;
        LV(f.Rhos, lev) = rhos + (LV(f.SlowRho, lev) - nflxr) * dt;
        const double ptc = (P_Ref / (Rd * r)) * MPOW((pss + pkv) / P_Ref, Cv / Cp);
        aux += ptc * r;
        const double pnew = P_Ref * MPOW(Rd * aux / P_Ref, Cp / Cv);
        const double psn = pnew - pkv + (LV(f.diffpr, lev) + 0.0) * dt + Rd / Cv * LV(f.Qheat, lev) * dt;
        LV(f.ps, lev) = psn;
        if (isnan(psn)) atomicOr(f.flags, FLAG_NAN_P);
        if (next_divm) {
            double o3[3];
            divm_pass0(g, p, lev, nf, gi, f.Mhs, f.Whs, o3);
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(f.divg, k, lev, p) = o3[k];
        }
    }
}

// ===================================================================== K_H
// UpdateRK2 (thor_auxiliary.h:420-529): y(k) += deviation. Whk is written to
// a separate buffer so each level can form W at the layer centre from both of
// its interfaces without a race.
template <int LC>
__global__ __launch_bounds__(256) void k_update_rk2(Geo g, Fields f) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const int l0 = blockIdx.y * LC, l1 = min(l0 + LC, nv);
    const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
    auto whk_new = [&](int lev) {
        return (lev == 0 || lev == nv) ? 0.0 : LV(f.Whk, lev) + LV(f.Whs, lev);
    };
    for (int lev = l0; lev < l1; lev++) {
        const double Mx = VEC(f.Mhs, 0, lev, p) + VEC(f.Mhk, 0, lev, p);
        const double My = VEC(f.Mhs, 1, lev, p) + VEC(f.Mhk, 1, lev, p);
        const double Mz = VEC(f.Mhs, 2, lev, p) + VEC(f.Mhk, 2, lev, p);
        const double aux = Mx * fx + My * fy + Mz * fz;
        VEC(f.Mhk_out, 0, lev, p) = Mx - aux * fx;
        VEC(f.Mhk_out, 1, lev, p) = My - aux * fy;
        VEC(f.Mhk_out, 2, lev, p) = Mz - aux * fz;
        LV(f.Rhok_out, lev) = LV(f.Rhok, lev) + LV(f.Rhos, lev);
        LV(f.pk_out, lev) = LV(f.pk, lev) + LV(f.ps, lev);
        const double wl = whk_new(lev), wt = whk_new(lev + 1);
        LV(f.Whk_out, lev) = wl;
        if (lev == nv - 1) LV(f.Whk_out, nv) = wt;
        const double althl = lev == 0 ? 0 : g.Alth[lev];
        const double altht = g.Alth[lev + 1];
        const double xi = g.Alt[lev], xim = althl, xip = altht;
        const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
        LV(f.Wk_out, lev) = wl * intt + wt * intl;
    }
}

// ===================================================================== K_P
// ESP::ProfX for the Held-Suarez benchmark, per column: Recompute_W
// (profx_auxiliary.h:110-130), Compute_temperature (:64-93), held_suarez
// (profx_held_suarez.h:51-117) and Compute_pressure (:95-108).
__global__ __launch_bounds__(256) void k_profx_hs(Geo g, Fields f, double time_step, int zero_q) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const double *Alt = g.Alt, *Alth = g.Alth;
    const double Rd = g.Rd, Cp = g.Cp, Gravit = g.Gravit;
    const double kappa = Rd / Cp;
    const double sigmab = 0.7, ka = (1.0 / 40.0) * (1.0 / 86400.0), kf = 1.0 / 86400.0,
                 ks = (1.0 / 4.0) * (1.0 / 86400.0), dTy = 60.0, dthetaz = 10.0, p0 = 100000.0;
    const double lat = g.lonlat[S + p];
    const double psm1 = LV(f.p, 1) - LV(f.Rho, 0) * Gravit * (-Alt[0] - Alt[1]);
    const double ps = 0.5 * (LV(f.p, 0) + psm1);
    bool nan_t = false;
    for (int lev = 0; lev < nv; lev++) {
        {
            const double whl = LV(f.Wh, lev), wht = LV(f.Wh, lev + 1);
            const double xi = Alt[lev], xim = Alth[lev], xip = Alth[lev + 1];
            const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            LV(f.W, lev) = whl * intt + wht * intl;
        }
        const double pre = LV(f.p, lev), rho = LV(f.Rho, lev);
        double T = pre / (Rd * rho);
        const double sigma = (pre / ps), sigma0 = (pre / p0);
        const double Teq = fmax(200.0, (315.0 - dTy * MPOW(MSIN(lat), 2.0) - dthetaz * MLOG(sigma0) * MPOW(MCOS(lat), 2.0))
                                           * MPOW(sigma0, kappa));
        const double kt = ka + (ks - ka) * fmax(0.0, (sigma - sigmab) / (1.0 - sigmab)) * MPOW(MCOS(lat), 4.0);
        const double kv = kf * fmax(0.0, (sigma - sigmab) / (1.0 - sigmab));
#pragma unroll
        for (int k = 0; k < 3; k++) VEC(f.Mh, k, lev, p) = VEC(f.Mh, k, lev, p) / (1.0 + kv * time_step);
        T -= kt * time_step * (T - Teq);
        nan_t |= isnan(T);
        LV(f.p, lev) = T * Rd * rho;
        if (zero_q) LV(f.Qheat, lev) = 0.0;
    }
    if (nan_t) atomicOr(f.flags, FLAG_NAN_T);
}

// ================================================================== globdiag
// Per-cell mass, total energy and angular momentum (CalcMass, CalcTotEnergy,
// CalcAngMom, profx_globdiag.cu:45-212), enumerated in THOR's cell order
// e = id * nv + lev and summed with the 1024-thread tree of
// gpu_reduction_sum (reduction_add.h:59-99): one partial per 2048 cells.
__device__ __forceinline__ void cell_diag(const Geo &g, const Fields &f, long e, double out[5]) {
    const size_t S = g.S;
    const int nv = g.nv;
    const int p = (int)(e / nv), lev = (int)(e % nv);
    const double A = g.A, Omega = g.Omega, Gravit = g.Gravit;
    const double zup = g.Alth[lev + 1] + A, zlow = g.Alth[lev] + A;
    double Vol;
    if (g.deep) Vol = g.areasT[p] / MPOW(A, 2) * (MPOW(zup, 3) - MPOW(zlow, 3)) / 3;
    else Vol = g.areasT[p] * (zup - zlow);
    const double rho = LV(f.Rho, lev);
    const double M0 = VEC(f.Mh, 0, lev, p), M1 = VEC(f.Mh, 1, lev, p), M2 = VEC(f.Mh, 2, lev, p);
    out[0] = rho * Vol;
    const double lon = g.lonlat[p], lat = g.lonlat[S + p];
    const double Wv = LV(f.W, lev);
    const double wx = Wv * MCOS(lat) * MCOS(lon), wy = Wv * MCOS(lat) * MSIN(lon), wz = Wv * MSIN(lat);
    const double Cv = g.Cp - g.Rd;
    const double Ek = 0.5 * ((M0 + wx) * (M0 + wx) + (M1 + wy) * (M1 + wy) + (M2 + wz) * (M2 + wz)) / rho;
    const double Eint = Cv * LV(f.T, lev) * rho;
    const double Eg = rho * Gravit * g.Alt[lev];
    out[1] = (Ek + Eint + Eg) * Vol;
    const double r = (A + g.Alt[lev]);
    const double rx = r * g.fr[p], ry = r * g.fr[S + p], rz = r * g.fr[2 * S + p];
    const double AMx = ry * M2 - rz * M1 - rho * Omega * r * rz * MCOS(lat) * MCOS(lon);
    const double AMy = -rx * M2 + rz * M0 - rho * Omega * r * rz * MCOS(lat) * MSIN(lon);
    const double AMz = rx * M1 - ry * M0 + rho * Omega * r * r * MCOS(lat) * MCOS(lat);
    out[2] = AMx * Vol;
    out[3] = AMy * Vol;
    out[4] = AMz * Vol;
}

__global__ __launch_bounds__(1024) void k_globdiag(Geo g, Fields f, double *partials, long n) {
    __shared__ double buf[5][2048];
    const long base = (long)blockIdx.x * 2048;
    for (int h = 0; h < 2; h++) {
        const long e = base + 2 * threadIdx.x + h;
        double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (e < n) cell_diag(g, f, e, v);
#pragma unroll
        for (int q = 0; q < 5; q++) buf[q][2 * threadIdx.x + h] = v[q];
    }
    for (int stride = 1024; stride > 0; stride /= 2) {
        __syncthreads();
        if ((int)threadIdx.x < stride) {
#pragma unroll
            for (int q = 0; q < 5; q++) buf[q][threadIdx.x] += buf[q][threadIdx.x + stride];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 5; q++) partials[(size_t)q * gridDim.x + blockIdx.x] = buf[q][0];
    }
}

// ============================================================ layout changes
// THOR layout <-> level-planar layout, on the device.
__global__ void k_aos_to_soa(const double *src, double *dst, int P, int S, int nlev, int ncomp) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;  // index into src
    const long n = (long)P * nlev * ncomp;
    if (i >= n) return;
    const int c = (int)(i % ncomp);
    const long r = i / ncomp;
    const int lev = (int)(r % nlev);
    const int p = (int)(r / nlev);
    dst[((size_t)c * nlev + lev) * S + p] = src[i];
}
__global__ void k_soa_to_aos(const double *src, double *dst, int P, int S, int nlev, int ncomp) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)P * nlev * ncomp;
    if (i >= n) return;
    const int c = (int)(i % ncomp);
    const long r = i / ncomp;
    const int lev = (int)(r % nlev);
    const int p = (int)(r / nlev);
    dst[i] = src[((size_t)c * nlev + lev) * S + p];
}

} // namespace thor
