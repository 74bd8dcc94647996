// Passive tracers of THOR's chemistry module, its dynamical part
// (src/physics/modules/inc/thor_chemistry.h, driven from chemistry.cu:333-502
// through the four physics-module hooks): the tracer mass rho q advected in
// flux form with the stage momentum plus its deviation in every acoustic
// substep (Tracer_Eq, :31-213, poles :216-361) and a fourth-order diffusion of
// the mixing ratio formed once per RK stage (Tracer_Eq_Diffusion, :364-597,
// poles :600-760). Arrays are level-planar [ntr][nv][S]; the arithmetic of
// each value is the reference's, including three things its pole kernels do:
// the first diffusion pass reads tracer 0 for the pole itself and weights the
// x part of each triangle with the z component of nvecoa, indexing the ring
// as (k + 1 + 1) % 5 (which wraps to the pole at one corner); the last
// diffusion pass writes nothing at the poles (THOR leaves its uninitialised
// values there; here they stay zero); the bottom layer's upper-interface
// value divides the tracer above by the density of the layer itself.
#pragma once

namespace thor {

struct TracerArgs {
    int ntr;
    const double *trk;   // tracer mass of the stage (tracerk_d)
    double *trs;         // its deviation (tracers_d)
    const double *dif;   // diffusion tendency (difftr_d)
    double *scr;         // first diffusion pass (THOR's diff_d)
    double *dif_out;     // last pass output (difftr_d)
    const double *Rhok, *Rhos, *Mhk, *Whs, *Whk;
    const void *Mhs;     // FP32 in the mixed form
};

#define TR(arr, itr, lev, q) (arr)[((size_t)(itr) * nv + (lev)) * S + (q)]

// Tracer_Eq_Diffusion's Laplacian around a hexagon or pentagon (1/6, 1/3
// weights; triangle j spans the ring neighbours j, j + 1 and j + 2 modulo the
// ring size). a(s) / rho(s): slot s = 0 the point, 1 + m ring neighbour m.
template <bool LAST, class GetA, class GetR>
__device__ __forceinline__ double tracer_lap(const Geo &g, int p, int nf, double rscale, double sdiff, GetA a,
                                             GetR rho) {
    const size_t S = g.S;
    const double o3 = 1.0 / 3.0, o6 = 1.0 / 6.0;
    const double a0 = a(0);
    double lap = 0.0, lapx1 = 0, lapy1 = 0, lapz1 = 0, lapx2 = 0, lapy2 = 0, lapz2 = 0, AT1 = 0;
    double lap1 = 0, lap2 = 0, lap3 = 0;
#define NI(j, k) g.nti[(size_t)((j)*3 + (k)) * S + p]
#define NE(j, k) g.nte[(size_t)((j)*3 + (k)) * S + p]
#define NO(j, k) g.noa[(size_t)((j)*3 + (k)) * S + p]
#pragma unroll 1
    for (int j = 0; j < nf; j++) {
        const int jp1 = (j + 1) % nf, jp2 = (j + 2) % nf;
        const double a1 = a(1 + j), a2 = a(1 + jp1), a3 = a(1 + jp2);
        double vdiff;
        if (LAST) vdiff = 0.5 * rscale * (2.0 * rho(0) + rho(1 + j) + 2.0 * rho(1 + jp1) + rho(1 + jp2)) * o6 * sdiff;
        else vdiff = 0.5 * rscale;
        if (j == 0) {
            AT1 = rscale / g.areasTr[(size_t)j * S + p];
            lap1 = o6 * (a0 + a1) - o3 * a2;
            lap2 = o6 * (a1 + a2) - o3 * a0;
            lap3 = o6 * (a0 + a2) - o3 * a1;
        }
        const double AT2 = rscale / g.areasTr[(size_t)jp1 * S + p];
        const double lap4 = o6 * (a0 + a2) - o3 * a3;
        const double lap5 = o6 * (a2 + a3) - o3 * a0;
        const double lap6 = o6 * (a0 + a3) - o3 * a2;
        if (j == 0) {
            lapx1 = (-lap1 * NI(j, 0) + lap2 * NE(j, 0) + lap3 * NI(jp1, 0)) * AT1;
            lapx2 = (-lap4 * NI(jp1, 0) + lap5 * NE(jp1, 0) + lap6 * NI(jp2, 0)) * AT2;
            lapy1 = (-lap1 * NI(j, 1) + lap2 * NE(j, 1) + lap3 * NI(jp1, 1)) * AT1;
            lapy2 = (-lap4 * NI(jp1, 1) + lap5 * NE(jp1, 1) + lap6 * NI(jp2, 1)) * AT2;
            lapz1 = (-lap1 * NI(j, 2) + lap2 * NE(j, 2) + lap3 * NI(jp1, 2)) * AT1;
            lapz2 = (-lap4 * NI(jp1, 2) + lap5 * NE(jp1, 2) + lap6 * NI(jp2, 2)) * AT2;
        } else {
            lapx1 = lapx2;
            lapx2 = (-lap4 * NI(jp1, 0) + lap5 * NE(jp1, 0) + lap6 * NI(jp2, 0)) * AT2;
            lapy1 = lapy2;
            lapy2 = (-lap4 * NI(jp1, 1) + lap5 * NE(jp1, 1) + lap6 * NI(jp2, 1)) * AT2;
            lapz1 = lapz2;
            lapz2 = (-lap4 * NI(jp1, 2) + lap5 * NE(jp1, 2) + lap6 * NI(jp2, 2)) * AT2;
        }
        lap += ((lapx1 + lapx2) * NO(j, 0) + (lapy1 + lapy2) * NO(j, 1) + (lapz1 + lapz2) * NO(j, 2)) * vdiff;
    }
    return lap;
}

// First pass at a pole (Tracer_Eq_Diffusion_Poles with laststep = 0), as the
// reference evaluates it (see the header comment).
template <class GetA>
__device__ __forceinline__ double tracer_lap_pole(const Geo &g, int p, double rscale, GetA a) {
    const size_t S = g.S;
    const double o3 = 1.0 / 3.0, o6 = 1.0 / 6.0;
    double lap = 0.0, lapx1 = 0, lapy1 = 0, lapz1 = 0, lapx2 = 0, lapy2 = 0, lapz2 = 0, AT1 = 0;
    double lap1 = 0, lap2 = 0, lap3 = 0;
#pragma unroll 1
    for (int k = 0; k < 5; k++) {
        const int j = k + 1, jp1 = (j + 1) % 5, jp2 = (j + 2) % 5, kp1 = (k + 1) % 5, kp2 = (k + 2) % 5;
        const double vdiff = 0.5 * rscale;
        if (k == 0) {
            lap1 = (o6 * (a(0) + a(j)) - o3 * a(jp1));
            lap2 = (o6 * (a(j) + a(jp1)) - o3 * a(0));
            lap3 = (o6 * (a(0) + a(jp1)) - o3 * a(j));
            AT1 = rscale / g.areasTr[(size_t)k * S + p];
        }
        const double AT2 = rscale / g.areasTr[(size_t)kp1 * S + p];
        const double lap4 = (o6 * (a(0) + a(jp1)) - o3 * a(jp2));
        const double lap5 = (o6 * (a(jp1) + a(jp2)) - o3 * a(0));
        const double lap6 = (o6 * (a(0) + a(jp2)) - o3 * a(jp1));
        if (k == 0) {
            lapx1 = (-lap1 * NI(k, 0) + lap2 * NE(k, 0) + lap3 * NI(kp1, 0)) * AT1;
            lapx2 = (-lap4 * NI(kp1, 0) + lap5 * NE(kp1, 0) + lap6 * NI(kp2, 0)) * AT2;
            lapy1 = (-lap1 * NI(k, 1) + lap2 * NE(k, 1) + lap3 * NI(kp1, 1)) * AT1;
            lapy2 = (-lap4 * NI(kp1, 1) + lap5 * NE(kp1, 1) + lap6 * NI(kp2, 1)) * AT2;
            lapz1 = (-lap1 * NI(k, 2) + lap2 * NE(k, 2) + lap3 * NI(kp1, 2)) * AT1;
            lapz2 = (-lap4 * NI(kp1, 2) + lap5 * NE(kp1, 2) + lap6 * NI(kp2, 2)) * AT2;
        } else {
            lapx1 = lapx2;
            lapx2 = (-lap4 * NI(kp1, 0) + lap5 * NE(kp1, 0) + lap6 * NI(kp2, 0)) * AT2;
            lapy1 = lapy2;
            lapy2 = (-lap4 * NI(kp1, 1) + lap5 * NE(kp1, 1) + lap6 * NI(kp2, 1)) * AT2;
            lapz1 = lapz2;
            lapz2 = (-lap4 * NI(kp1, 2) + lap5 * NE(kp1, 2) + lap6 * NI(kp2, 2)) * AT2;
        }
        lap += ((lapx1 + lapx2) * NO(k, 2) + (lapy1 + lapy2) * NO(k, 1) + (lapz1 + lapz2) * NO(k, 2)) * vdiff;
    }
    return lap;
}
#undef NI
#undef NE
#undef NO

// Diffusion pass (LAST = false: scr = lap(trk / Rhok); LAST = true:
// dif_out = -Kdh4 rho-weighted lap(scr)), one tile x one level x one tracer
// per block; the pole block runs the poles' first pass only.
template <bool LAST> __global__ STENCIL_BOUNDS void k_tracer_diff(Geo g, TracerArgs t) {
    int r, lev;
    if (!block_tile(g, r, lev)) return;
    const size_t S = g.S;
    const int nv = g.nv, itr = blockIdx.y;
    const double rs = rscale_at(g, lev);
    if (r == g.nr) {
        if (LAST) return;
        PoleCtx c;
        if (!pole_ctx(g, c)) return;
        // the pole itself reads tracer 0
        auto a = [&](int s) {
            const int q = c.gi(s);
            return TR(t.trk, s == 0 ? 0 : itr, lev, q) / LVP(t.Rhok, lev, q);
        };
        TR(t.scr, itr, lev, c.p) = tracer_lap_pole(g, c.p, rs, a);
        return;
    }
    __shared__ double sA[NHL2], sR[NHL2];
    const TileThread T = tile_thread(g, r);
    stage_tile(g, r, T, [&](int pos, int q) {
        if (q < 0) {
            sA[pos] = sR[pos] = 0.0;
            return;
        }
        const double rho = LVP(t.Rhok, lev, q);
        sR[pos] = rho;
        sA[pos] = LAST ? TR(t.scr, itr, lev, q) : TR(t.trk, itr, lev, q) / rho;
    });
    __syncthreads();
    // slot 1 + m: ring neighbour m (a pentagon's ring has five members; its
    // slot 5 is staged at THOR's pt6 position)
    const double lap = tracer_lap<LAST>(
        g, T.p, T.nf, rs, LAST ? -g.Kdh4[lev] : 0.0, [&](int s) { return sA[T.pos(s)]; },
        [&](int s) { return sR[T.pos(s)]; });
    if (LAST) TR(t.dif_out, itr, lev, T.p) = lap;
    else TR(t.scr, itr, lev, T.p) = lap;
}

// Tracer_Eq: flux-form advection of the tracer mass by (Mhs + Mhk) and
// (Whs + Whk), plus the diffusion tendency, into the tracer deviation.
template <bool MX> __global__ STENCIL_BOUNDS void k_tracer_eq(Geo g, TracerArgs t, double dt) {
    int r, lev;
    if (!block_tile(g, r, lev)) return;
    const size_t S = g.S;
    const int nv = g.nv, itr = blockIdx.y;
    using R = typename FS<MX>::R;
    const R *Mhs = (const R *)t.Mhs;
    const double A = g.A;
    const double altht = g.Alth[lev + 1], althl = g.Alth[lev];
    double r2p, r2m, r2l, rscale;
    if (g.deep) {
        const double alt = g.Alt[lev];
        r2p = MPOW(altht + A, 2.0);
        r2m = MPOW(alt + A, 2.0);
        r2l = MPOW(althl + A, 2.0);
        rscale = A / (alt + A);
    } else {
        r2p = r2m = r2l = rscale = 1.0;
    }
    auto mixr = [&](int l, int q) { return TR(t.trk, itr, l, q) / LVP(t.Rhok, l, q); };
    // the column part, after the horizontal flux divergence nflx
    auto finish = [&](int p, double nflx, double trc, bool pole) {
        double trhl, trht, xi, xim1, xip1, a, b;
        if (lev == 0) {
            trhl = 0.0;
            xi = g.Alth[lev + 1];
            xim1 = g.Alt[lev];
            xip1 = g.Alt[lev + 1];
            a = (xi - xip1) / (xim1 - xip1);
            b = (xi - xim1) / (xip1 - xim1);
            trht = pole ? TR(t.trk, itr, lev, p) / LVP(t.Rhok, lev, p) * a
                              + TR(t.trk, itr, lev + 1, p) / LVP(t.Rhok, lev, p) * b
                        : mixr(lev, p) * a + mixr(lev + 1, p) * b;
        } else if (lev == nv - 1) {
            xi = g.Alth[lev];
            xim1 = g.Alt[lev - 1];
            xip1 = g.Alt[lev];
            a = (xi - xip1) / (xim1 - xip1);
            b = (xi - xim1) / (xip1 - xim1);
            trhl = mixr(lev - 1, p) * a + mixr(lev, p) * b;
            trht = 0.0;
        } else {
            xi = g.Alth[lev];
            xim1 = g.Alt[lev - 1];
            xip1 = g.Alt[lev];
            a = (xi - xip1) / (xim1 - xip1);
            b = (xi - xim1) / (xip1 - xim1);
            trhl = mixr(lev - 1, p) * a + mixr(lev, p) * b;
            xi = g.Alth[lev + 1];
            xim1 = g.Alt[lev];
            xip1 = g.Alt[lev + 1];
            a = (xi - xip1) / (xim1 - xip1);
            b = (xi - xim1) / (xip1 - xim1);
            trht = mixr(lev, p) * a + mixr(lev + 1, p) * b;
        }
        const double whl = lev == 0 ? 0.0 : LVP(t.Whs, lev, p) + LVP(t.Whk, lev, p);
        const double wht = LVP(t.Whs, lev + 1, p) + LVP(t.Whk, lev + 1, p);
        const double dz = altht - althl;
        const double dtr_dalt = (wht * trht * r2p - whl * trhl * r2l) / (dz * r2m);
        const double dtr = -(nflx + dtr_dalt) * dt;
        const double rho = LVP(t.Rhok, lev, p) + LVP(t.Rhos, lev, p);
        const double tr = trc * rho + dtr;
        double v = tr - trc * LVP(t.Rhok, lev, p);
        v = v + TR(t.dif, itr, lev, p) * dt;
        TR(t.trs, itr, lev, p) = v;
    };
    auto v1 = [&](int q, int k) { return (double)VEC(Mhs, k, lev, q) + VEC(t.Mhk, k, lev, q); };
    if (r == g.nr) {
        PoleCtx c;
        if (!pole_ctx(g, c)) return;
        const int p = c.p;
        double nflx = 0.0;
        for (int k = 0; k < 3; k++) {
            double s = g.div[(size_t)(0 * 3 + k) * S + p] * v1(p, k) * mixr(lev, p);
            for (int i = 1; i < 6; i++) {
                const int q = c.gi(i);
                s = s + g.div[(size_t)(i * 3 + k) * S + p] * v1(q, k) * mixr(lev, q);
            }
            nflx += rscale * s;
        }
        finish(p, nflx, mixr(lev, p), true);
        return;
    }
    __shared__ double sV[3][NHL2], sT[NHL2];
    const TileThread T = tile_thread(g, r);
    stage_tile(g, r, T, [&](int pos, int q) {
        if (q < 0) {
            sV[0][pos] = sV[1][pos] = sV[2][pos] = sT[pos] = 0.0;
            return;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) sV[k][pos] = v1(q, k);
        sT[pos] = mixr(lev, q);
    });
    __syncthreads();
    const int p = T.p;
    double nflx = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double s = g.div[(size_t)(0 * 3 + k) * S + p] * sV[k][T.pos(0)] * sT[T.pos(0)];
#pragma unroll
        for (int i = 1; i < 7; i++) s = s + g.div[(size_t)(i * 3 + k) * S + p] * sV[k][T.pos(i)] * sT[T.pos(i)];
        nflx += rscale * s;
    }
    finish(p, nflx, sT[T.ir], false);
}

#undef TR
} // namespace thor
