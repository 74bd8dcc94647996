// FP64 horizontal stencils over NL consecutive levels per block.
//
// The per-level FP64 stencil kernels (one block per tile and level) read each
// point's operator geometry (div, grad, areas, nvecoa, nvecti, nvecte,
// areasTr: up to ~100 doubles) once per level, so at 48 levels the geometry
// dominates their memory traffic and their dependent loads. Here a block takes
// one 16x16 tile at NL consecutive levels: the tile and its halo are staged in
// LDS for all NL levels, and each thread walks its point's triangles once,
// loading a triangle's geometry once and applying it to all NL levels.
// Every cell's arithmetic is the per-level kernel's (the same expressions in
// the same order, compiled without contraction); only the interleaving of
// independent cells changes, so the results are bit for bit those of the
// per-level kernels and of the reference's kernels.
//
// Levels past nv in the last chunk read level nv - 1 (in bounds) and write
// nothing; interface fields are staged at interfaces l0 .. l0 + NL clamped to
// nv, so the chunk's last real level still sees its own upper interface. The two poles are handled by the extra tile block with the per-level
// point functions (thread t: pole t & 1 at level l0 + t / 2).
#pragma once
#include "dyn_kernels.h"

namespace thor {

// Minimum resident blocks per CU requested for the substep kernels (momentum,
// divm0, prepare) and for the once-per-stage kernels (adv_diff1, diff2_divm0):
// bounds their register use against their LDS use. (The slow modes stay one
// level per block: over two levels their register need spills, 463 against
// 314 us at g6 L48.)
#ifndef THOR_ML_MINB_SUB
#define THOR_ML_MINB_SUB 2
#endif
#ifndef THOR_ML_MINB_STAGE
#define THOR_ML_MINB_STAGE 2
#endif
#define ML_SUB_BOUNDS __launch_bounds__(256, THOR_ML_MINB_SUB)
#define ML_STAGE_BOUNDS __launch_bounds__(256, THOR_ML_MINB_STAGE)

// Block -> (tile, first level): XCD-major like block_tile_lc.
template <int NL> __device__ __forceinline__ bool block_tile_ml(const Geo &g, int &r, int &l0) {
    const int i = blockIdx.x, xcd = i & 7, idx = i >> 3;
    const int nch = (g.nv + NL - 1) / NL;
    r = xcd * g.tiles_per_xcd + idx / nch;
    l0 = (idx % nch) * NL;
    return r <= g.nr;
}

// put(l, pos, q, lev): the values of global point q (q < 0: zeros) at level
// lev (clamped to nv - 1) into LDS position pos of level slot l; for the
// thread's own point and, for threads 0..65, one halo point.
template <int NL, class Put>
__device__ __forceinline__ void stage_tile_ml(const Geo &g, int r, const TileThread &T, int l0, Put put) {
    const int t = threadIdx.x;
    const bool halo = t < NHALO;
    const int hq = halo ? g.halo[r * NHALO + t] : -1;
    const int hpos = halo ? halo_pos(t) : 0;
#pragma unroll
    for (int l = 0; l < NL; l++) {
        const int lev = min(l0 + l, g.nv - 1);
        put(l, T.ir, T.p, lev);
        if (halo) put(l, hpos, hq, lev);
    }
}

// Pole threads of an NL-level block: pole P - 2 + (t & 1) at level l0 + t / 2.
template <int NL> __device__ __forceinline__ bool pole_ctx_ml(const Geo &g, int l0, PoleCtx &c, int &lev) {
    const int t = threadIdx.x;
    if (t >= 2 * NL) return false;
    lev = l0 + (t >> 1);
    if (lev >= g.nv) return false;
    c.p = g.P - 2 + (t & 1);
    c.ring = g.ring;
    c.S = g.S;
    return true;
}

// laplacians<NV> at NL levels: a(v, l, s) is field v at level slot l, stencil
// slot s (an LDS read; called again wherever laplacians<NV> reuses a value, so
// it must return the same number each time); rs[l] the level's rscale;
// out[v][l].
template <int NV, int NL, class GetA>
__device__ __forceinline__ void laplacians_ml(const Geo &g, int p, int nf, const double rs[NL], GetA a,
                                              double out[NV][NL]) {
    const size_t S = g.S;
    double lap[NV][NL], lx[NV][NL], ly[NV][NL], lz[NV][NL];
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int l = 0; l < NL; l++) lap[v][l] = 0.0;
#pragma unroll 1
    for (int j = 0; j < nf; j++) {
        const int jp1 = j + 1 == nf ? 0 : j + 1;
        const int jp2 = jp1 + 1 == nf ? 0 : jp1 + 1;
        const double aTr2 = g.areasTr[(size_t)jp1 * S + p];
        const double b0 = g.areas[(size_t)(jp1 * 3 + 0) * S + p], b1 = g.areas[(size_t)(jp1 * 3 + 1) * S + p],
                     b2 = g.areas[(size_t)(jp1 * 3 + 2) * S + p];
        const double ti1x = g.nti[(size_t)(jp1 * 3 + 0) * S + p], ti1y = g.nti[(size_t)(jp1 * 3 + 1) * S + p],
                     ti1z = g.nti[(size_t)(jp1 * 3 + 2) * S + p];
        const double te1x = g.nte[(size_t)(jp1 * 3 + 0) * S + p], te1y = g.nte[(size_t)(jp1 * 3 + 1) * S + p],
                     te1z = g.nte[(size_t)(jp1 * 3 + 2) * S + p];
        const double ti2x = g.nti[(size_t)(jp2 * 3 + 0) * S + p], ti2y = g.nti[(size_t)(jp2 * 3 + 1) * S + p],
                     ti2z = g.nti[(size_t)(jp2 * 3 + 2) * S + p];
        const double oax = g.noa[(size_t)(j * 3 + 0) * S + p], oay = g.noa[(size_t)(j * 3 + 1) * S + p],
                     oaz = g.noa[(size_t)(j * 3 + 2) * S + p];
        if (j == 0) {
            const double aTr = g.areasTr[p];
            const double c0 = g.areas[p], c1 = g.areas[S + p], c2 = g.areas[2 * S + p];
            const double ti0x = g.nti[p], ti0y = g.nti[S + p], ti0z = g.nti[2 * S + p];
            const double te0x = g.nte[p], te0y = g.nte[S + p], te0z = g.nte[2 * S + p];
#pragma unroll
            for (int l = 0; l < NL; l++) {
                const double AT1 = rs[l] / aTr;
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    const double a0 = a(v, l, 0), a1 = a(v, l, 1), a2 = a(v, l, 2);
                    const double astar = (c0 * a0 + c1 * a1 + c2 * a2) / aTr;
                    const double lap1 = (0.5 * (a1 + a0) - astar);
                    const double lap2 = (0.5 * (a1 + a2) - astar);
                    const double lap3 = (0.5 * (a0 + a2) - astar);
                    lx[v][l] = (-lap1 * ti0x + lap2 * te0x + lap3 * ti1x) * AT1;
                    ly[v][l] = (-lap1 * ti0y + lap2 * te0y + lap3 * ti1y) * AT1;
                    lz[v][l] = (-lap1 * ti0z + lap2 * te0z + lap3 * ti1z) * AT1;
                }
            }
        }
#pragma unroll
        for (int l = 0; l < NL; l++) {
            const double AT2 = rs[l] / aTr2, vdiff = 0.5 * rs[l];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                // a2: slot 1 + jp1 (the previous triangle's third corner)
                const double a0 = a(v, l, 0), a2 = a(v, l, 1 + jp1), a3 = a(v, l, 1 + jp2);
                const double astar = (b0 * a0 + b1 * a2 + b2 * a3) / aTr2;
                const double lap4 = (0.5 * (a0 + a2) - astar);
                const double lap5 = (0.5 * (a2 + a3) - astar);
                const double lap6 = (0.5 * (a0 + a3) - astar);
                const double x2 = (-lap4 * ti1x + lap5 * te1x + lap6 * ti2x) * AT2;
                const double y2 = (-lap4 * ti1y + lap5 * te1y + lap6 * ti2y) * AT2;
                const double z2 = (-lap4 * ti1z + lap5 * te1z + lap6 * ti2z) * AT2;
                lap[v][l] += ((lx[v][l] + x2) * oax + (ly[v][l] + y2) * oay + (lz[v][l] + z2) * oaz) * vdiff;
                lx[v][l] = x2;
                ly[v][l] = y2;
                lz[v][l] = z2;
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int l = 0; l < NL; l++) out[v][l] = lap[v][l];
}

// The per-level factors of divm_nf for the block's NL levels, formed once per
// block into LDS (sLv[0][l] = rscale, sLv[1][l] = 1 / (Alth[lev] -
// Alth[lev + 1])) by threads 0 .. NL - 1, before the block's barrier: the
// triangle loop reads them instead of dividing again for every triangle.
template <int NL> __device__ __forceinline__ void level_factors_ml(const Geo &g, int l0, double (*sLv)[NL]) {
    const int t = threadIdx.x;
    if (t < NL) {
        const int lev = min(l0 + t, g.nv - 1);
        sLv[0][t] = rscale_at(g, lev);
        sLv[1][t] = 1.0 / (g.Alth[lev] - g.Alth[lev + 1]);
    }
}

// divm_nf at NL levels l0 .. l0 + NL - 1 (clamped to nv - 1 like the staging):
// a(l, s, k), wl(l, s), wt(l, s) (LDS reads, called again wherever divm_nf
// reuses a value); sLv from level_factors_ml; out[l][3]. Tile points only
// (pole = false).
template <bool LAST, int NL, class GetA, class GetWl, class GetWt>
__device__ __forceinline__ void divm_ml(const Geo &g, int p, int l0, int nf, GetA a, GetWl wl, GetWt wt,
                                        const double (*sLv)[NL], double out[NL][3]) {
    const size_t S = g.S;
    double lapx[NL], lapy[NL], lapz[NL], lap2[NL], dwdz2[NL];
#pragma unroll
    for (int l = 0; l < NL; l++) lapx[l] = lapy[l] = lapz[l] = lap2[l] = dwdz2[l] = 0.0;
#define AR(j, k) g.areas[(size_t)((j)*3 + (k)) * S + p]
#define TI(j, k) g.nti[(size_t)((j)*3 + (k)) * S + p]
#define TE(j, k) g.nte[(size_t)((j)*3 + (k)) * S + p]
    double ti1x = TI(0, 0), ti1y = TI(0, 1), ti1z = TI(0, 2);
    double aT1 = g.areasTr[p];
#pragma unroll 1
    for (int j = 0; j < nf; j++) {
        const int jp1 = j + 1 == nf ? 0 : j + 1;
        const int jp2 = jp1 + 1 == nf ? 0 : jp1 + 1;
        const int s1 = 1 + j, s2 = 1 + jp1, s3 = 1 + jp2;
        const double aT2 = g.areasTr[(size_t)jp1 * S + p];
        const double ti2x = TI(jp1, 0), ti2y = TI(jp1, 1), ti2z = TI(jp1, 2);
        const double te2x = TE(jp1, 0), te2y = TE(jp1, 1), te2z = TE(jp1, 2);
        const double ti3x = TI(jp2, 0), ti3y = TI(jp2, 1), ti3z = TI(jp2, 2);
        const double nox = g.noa[(size_t)(j * 3 + 0) * S + p], noy = g.noa[(size_t)(j * 3 + 1) * S + p],
                     noz = g.noa[(size_t)(j * 3 + 2) * S + p];
        double r0 = 0.0, r1 = 0.0, w2 = 0.0;
        if (!LAST) {
            r0 = AR(jp1, 0);
            r1 = AR(jp1, 1);
            w2 = AR(jp1, 2);
        }
#pragma unroll
        for (int l = 0; l < NL; l++) {
            const double rs = sLv[0][l];
            const double AT0 = 0.5 * rs;
            const double dz = LAST ? 0.0 : sLv[1][l];
            const double a0x = a(l, 0, 0), a0y = a(l, 0, 1), a0z = a(l, 0, 2);
            // slot 1 + j: the previous triangle's third corner
            const double a1x = a(l, s1, 0), a1y = a(l, s1, 1), a1z = a(l, s1, 2);
            const double a2x = a(l, s2, 0), a2y = a(l, s2, 1), a2z = a(l, s2, 2);
            const double a3x = a(l, s3, 0), a3y = a(l, s3, 1), a3z = a(l, s3, 2);
            double dwdz = 0.0;
            if (!LAST) {
                const double wl0 = wl(l, 0), wt0 = wt(l, 0), wl2 = wl(l, s2), wt2 = wt(l, s2);
                if (j == 0) {
                    const double wl1 = wl(l, s1), wt1 = wt(l, s1);
                    const double q0 = AR(0, 0), q1 = AR(0, 1), w1 = AR(0, 2);
                    const double meanwl = (q0 * wl0 + q1 * wl1 + w1 * wl2) / aT1;
                    const double meanwt = (q0 * wt0 + q1 * wt1 + w1 * wt2) / aT1;
                    dwdz = (meanwl - meanwt) * dz;
                } else {
                    dwdz = dwdz2[l];
                }
                const double wl3 = wl(l, s3), wt3 = wt(l, s3);
                const double meanwl2 = (r0 * wl0 + r1 * wl2 + w2 * wl3) / aT2;
                const double meanwt2 = (r0 * wt0 + r1 * wt2 + w2 * wt3) / aT2;
                dwdz2[l] = (meanwl2 - meanwt2) * dz;
            }
            const double AT2 = 0.5 * rs / aT2;
            const double a02x = a0x + a2x, a02y = a0y + a2y, a02z = a0z + a2z;
            double lap1;
            if (j == 0) {
                const double a01x = a0x + a1x, a01y = a0y + a1y, a01z = a0z + a1z;
                const double a12x = a1x + a2x, a12y = a1y + a2y, a12z = a1z + a2z;
                const double AT1 = 0.5 * rs / aT1;
                lap1 = (-a01x * ti1x - a01y * ti1y - a01z * ti1z + a12x * TE(0, 0) + a12y * TE(0, 1)
                        + a12z * TE(0, 2) + a02x * ti2x + a02y * ti2y + a02z * ti2z)
                       * AT1;
            } else {
                lap1 = lap2[l];
            }
            const double a23x = a2x + a3x, a23y = a2y + a3y, a23z = a2z + a3z;
            const double a03x = a0x + a3x, a03y = a0y + a3y, a03z = a0z + a3z;
            lap2[l] = (-a02x * ti2x - a02y * ti2y - a02z * ti2z + a23x * te2x + a23y * te2y + a23z * te2z
                       + a03x * ti3x + a03y * ti3y + a03z * ti3z)
                      * AT2;
            lapx[l] += (lap1 + dwdz + lap2[l] + dwdz2[l]) * nox * AT0;
            lapy[l] += (lap1 + dwdz + lap2[l] + dwdz2[l]) * noy * AT0;
            lapz[l] += (lap1 + dwdz + lap2[l] + dwdz2[l]) * noz * AT0;
        }
        ti1x = ti2x;
        ti1y = ti2y;
        ti1z = ti2z;
        aT1 = aT2;
    }
#undef AR
#undef TI
#undef TE
#pragma unroll
    for (int l = 0; l < NL; l++) {
        if (LAST) {
            const int lev = min(l0 + l, g.nv - 1);
            const double sdiff = g.Kdhz[lev];
            const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
            const double lapr = fx * lapx[l] + fy * lapy[l] + fz * lapz[l];
            const double x = lapx[l] - fx * lapr, y = lapy[l] - fy * lapr, z = lapz[l] - fz * lapr;
            out[l][0] = -x * sdiff;
            out[l][1] = -y * sdiff;
            out[l][2] = -z * sdiff;
        } else {
            out[l][0] = lapx[l];
            out[l][1] = lapy[l];
            out[l][2] = lapz[l];
        }
    }
}

// ------------------------------------------------------------------ kernels
// Grid: 8 * tiles_per_xcd * ceil(nv / NL) blocks of 256 threads.

// Divergence-damping pass 0 on the deviations (k_divm0, FP64).
template <int NL> __global__ ML_SUB_BOUNDS void k_divm0_ml(Geo g, Fields f) {
    int r, l0;
    if (!block_tile_ml<NL>(g, r, l0)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const double *Mhs = f.Mhs;
    double *divg = f.divg;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (!pole_ctx_ml<NL>(g, l0, c, lev)) return;
        double o3[3];
        divm_nf<false>(
            g, c.p, lev, 5, true, [&](int s, int k) { return VEC(Mhs, k, lev, c.gi(s)); },
            [&](int s) { return LVP(f.Whs, lev, c.gi(s)); }, [&](int s) { return LVP(f.Whs, lev + 1, c.gi(s)); }, o3);
#pragma unroll
        for (int k = 0; k < 3; k++) VEC(divg, k, lev, c.p) = o3[k];
        return;
    }
    __shared__ double sM[NL][3][NHL2], sW[NL + 1][NHL2], sLv[2][NL];
    const TileThread T = tile_thread(g, r);
    level_factors_ml<NL>(g, l0, sLv);
    stage_tile_ml<NL>(g, r, T, l0, [&](int l, int pos, int q, int lev) {
#pragma unroll
        for (int k = 0; k < 3; k++) sM[l][k][pos] = q < 0 ? 0.0 : VEC(Mhs, k, lev, q);
        // interfaces l0 .. l0 + NL, clamped to nv (not nv - 1): the last real
        // level of a partial chunk reads its upper interface from slot l + 1
        sW[l][pos] = q < 0 ? 0.0 : LVP(f.Whs, min(l0 + l, nv), q);
        if (l == NL - 1) sW[NL][pos] = q < 0 ? 0.0 : LVP(f.Whs, min(l0 + NL, nv), q);
    });
    __syncthreads();
    double o[NL][3];
    divm_ml<false, NL>(
        g, T.p, l0, T.nf, [&](int l, int s, int k) { return sM[l][k][T.pos(s)]; },
        [&](int l, int s) { return sW[l][T.pos(s)]; }, [&](int l, int s) { return sW[l + 1][T.pos(s)]; }, sLv,
        o);
#pragma unroll
    for (int l = 0; l < NL; l++) {
        const int lev = l0 + l;
        if (lev < nv) {
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(divg, k, lev, T.p) = o[l][k];
        }
    }
}

// Divergence-damping pass 1 + Momentum_Eq (k_momentum, FP64).
template <int NL> __global__ ML_SUB_BOUNDS void k_momentum_ml(Geo g, Fields f, int divdamp, double dt) {
    int r, l0;
    if (!block_tile_ml<NL>(g, r, l0)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const double *divg = f.divg;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (!pole_ctx_ml<NL>(g, l0, c, lev)) return;
        momentum_point<false>(
            g, f, c.p, lev, 5, true, divdamp, dt, [&](int s, int k) { return VEC(divg, k, lev, c.gi(s)); },
            [&](int s) { return LVP(f.ps, lev, c.gi(s)); });
        return;
    }
    __shared__ double sG[NL][3][NHL2], sP[NL][NHL2], sLv[2][NL];
    const TileThread T = tile_thread(g, r);
    level_factors_ml<NL>(g, l0, sLv);
    stage_tile_ml<NL>(g, r, T, l0, [&](int l, int pos, int q, int lev) {
        if (divdamp) {
#pragma unroll
            for (int k = 0; k < 3; k++) sG[l][k][pos] = q < 0 ? 0.0 : VEC(divg, k, lev, q);
        }
        sP[l][pos] = q < 0 ? 0.0 : LVP(f.ps, lev, q);
    });
    __syncthreads();
    const int p = T.p;
    double DivM[NL][3];
#pragma unroll
    for (int l = 0; l < NL; l++) DivM[l][0] = DivM[l][1] = DivM[l][2] = 0.0;
    if (divdamp) {
        auto none = [&](int, int) { return 0.0; };
        divm_ml<true, NL>(
            g, p, l0, T.nf, [&](int l, int s, int k) { return sG[l][k][T.pos(s)]; }, none, none, sLv, DivM);
    }
    double rs[NL], nf3[NL][3];
#pragma unroll
    for (int l = 0; l < NL; l++) rs[l] = rscale_at(g, min(l0 + l, nv - 1));
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double s0[NL];
        const double g0 = g.grad[(size_t)(0 * 3 + k) * S + p];
#pragma unroll
        for (int l = 0; l < NL; l++) s0[l] = g0 * sP[l][T.pos(0)];
#pragma unroll
        for (int s = 1; s < 7; s++) {
            const double gs = g.grad[(size_t)(s * 3 + k) * S + p];
#pragma unroll
            for (int l = 0; l < NL; l++) s0[l] = s0[l] + gs * sP[l][T.pos(s)];
        }
#pragma unroll
        for (int l = 0; l < NL; l++) nf3[l][k] = rs[l] * s0[l];
    }
#pragma unroll
    for (int l = 0; l < NL; l++)
        if (l0 + l < nv) momentum_tail<false>(g, f, p, l0 + l, nf3[l], DivM[l], dt);
}

// Divergence-damping pass 1 of the slow modes (DivM of slow_point, FP64) at
// the tile points, into dm [3][nv][S], for k_slow_lc2<false> (whose pole
// blocks form the poles' DivM themselves): the same sums as momentum's pass 1.
template <int NL> __global__ ML_SUB_BOUNDS void k_divm1_ml(Geo g, Fields f, double *dm) {
    int r, l0;
    if (!block_tile_ml<NL>(g, r, l0) || r == g.nr) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const double *divg = f.divg;
    __shared__ double sG[NL][3][NHL2], sLv[2][NL];
    const TileThread T = tile_thread(g, r);
    level_factors_ml<NL>(g, l0, sLv);
    stage_tile_ml<NL>(g, r, T, l0, [&](int l, int pos, int q, int lev) {
#pragma unroll
        for (int k = 0; k < 3; k++) sG[l][k][pos] = q < 0 ? 0.0 : VEC(divg, k, lev, q);
    });
    __syncthreads();
    double DivM[NL][3];
    auto none = [&](int, int) { return 0.0; };
    divm_ml<true, NL>(
        g, T.p, l0, T.nf, [&](int l, int s, int k) { return sG[l][k][T.pos(s)]; }, none, none, sLv, DivM);
#pragma unroll
    for (int l = 0; l < NL; l++)
        if (l0 + l < nv) {
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(dm, k, l0 + l, T.p) = DivM[l][k];
        }
}

// Prepare_Implicit_Vertical + the horizontal flux divergences of
// Density_Pressure_Eqs (k_prepare, FP64): prepare_point_w's sums at NL levels.
template <int NL> __global__ ML_SUB_BOUNDS void k_prepare_ml(Geo g, Fields f) {
    int r, l0;
    if (!block_tile_ml<NL>(g, r, l0)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const double *Mhs = f.Mhs;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (!pole_ctx_ml<NL>(g, l0, c, lev)) return;
        prepare_point(
            g, f, c.p, lev, [&](int s, int k) { return VEC(Mhs, k, lev, c.gi(s)); },
            [&](int s) { return LVP(f.h, lev, c.gi(s)); },
            [&](int s, int k) { return VEC(Mhs, k, lev, c.gi(s)) + VEC(f.Mhk, k, lev, c.gi(s)); },
            [&](int s) { return LVP(f.pt, lev, c.gi(s)); });
        return;
    }
    __shared__ double sM[NL][3][NHL2], sV1[NL][3][NHL2], sH[NL][NHL2], sPt[NL][NHL2];
    const TileThread T = tile_thread(g, r);
    stage_tile_ml<NL>(g, r, T, l0, [&](int l, int pos, int q, int lev) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double m = q < 0 ? 0.0 : VEC(Mhs, k, lev, q);
            sM[l][k][pos] = m;
            sV1[l][k][pos] = q < 0 ? 0.0 : m + VEC(f.Mhk, k, lev, q);
        }
        sH[l][pos] = q < 0 ? 0.0 : LVP(f.h, lev, q);
        sPt[l][pos] = q < 0 ? 0.0 : LVP(f.pt, lev, q);
    });
    __syncthreads();
    const int p = T.p;
    const double RoC = g.Rd / (g.Cp - g.Rd);
    double rs[NL], nflxr[NL], nflxp[NL], dr[NL], dpt[NL];
#pragma unroll
    for (int l = 0; l < NL; l++) {
        rs[l] = rscale_at(g, min(l0 + l, nv - 1));
        nflxr[l] = nflxp[l] = dr[l] = dpt[l] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double sr[NL], sp[NL], st[NL];
        const double d0 = g.div[(size_t)(0 * 3 + k) * S + p];
#pragma unroll
        for (int l = 0; l < NL; l++) {
            const int q0 = T.pos(0);
            sr[l] = d0 * sM[l][k][q0];
            sp[l] = d0 * sM[l][k][q0] * sH[l][q0];
            st[l] = d0 * sV1[l][k][q0] * sPt[l][q0];
        }
#pragma unroll
        for (int s = 1; s < 7; s++) {
            const double dd = g.div[(size_t)(s * 3 + k) * S + p];
            const int qs = T.pos(s);
#pragma unroll
            for (int l = 0; l < NL; l++) {
                sr[l] = sr[l] + dd * sM[l][k][qs];
                sp[l] = sp[l] + dd * sM[l][k][qs] * sH[l][qs];
                st[l] = st[l] + dd * sV1[l][k][qs] * sPt[l][qs];
            }
        }
#pragma unroll
        for (int l = 0; l < NL; l++) {
            nflxr[l] -= rs[l] * sr[l];
            nflxp[l] -= (RoC)*rs[l] * sp[l];
            dr[l] += rs[l] * sr[l];
            dpt[l] += rs[l] * st[l];
        }
    }
#pragma unroll
    for (int l = 0; l < NL; l++) {
        const int lev = l0 + l;
        if (lev < nv) {
            LV(f.Sp, lev) = nflxp[l] + LV(f.Slowp, lev);
            LV(f.Sd, lev) = nflxr[l] + LV(f.SlowRho, lev);
            LV(f.hflx_r, lev) = dr[l];
            LV(f.hflx_pt, lev) = dpt[l];
        }
    }
}

// Horizontal advection (Compute_Advec_Cori1) + the first hyperdiffusion pass
// (k_adv_diff1, FP64).
template <int NL> __global__ ML_STAGE_BOUNDS void k_adv_diff1_ml(Geo g, Fields f, int hydiff) {
    int r, l0;
    if (!block_tile_ml<NL>(g, r, l0)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (!pole_ctx_ml<NL>(g, l0, c, lev)) return;
        auto vel = [&](int q, int k) {
            const double rho = LVP(f.Rhok, lev, q), w = LVP(f.Wk, lev, q);
            return VEC(f.Mhk, k, lev, q) / rho + (w / rho) * g.fr[(size_t)k * S + q];
        };
        adv_diff1_point<false>(
            g, f, c.p, lev, 5, hydiff, [&](int s, int k) { return VEC(f.Mhk, k, lev, c.gi(s)); },
            [&](int s, int k) { return vel(c.gi(s), k); }, [&](int s) { return LVP(f.Rhok, lev, c.gi(s)); },
            [&](int s) { return LVP(f.Wk, lev, c.gi(s)); }, [&](int s) { return LVP(f.T, lev, c.gi(s)); });
        return;
    }
    // sA: the hyperdiffused variables rho, M / rho, W / rho and Rd T (internal
    // energy (Cp - Rd) T with the energy equation, thor_diff.h:196-202),
    // formed once per staged value
    __shared__ double sM[NL][3][NHL2], sV[NL][3][NHL2], sA[NL][6][NHL2];
    const TileThread T = tile_thread(g, r);
    const double Rd = g.Rd;
    stage_tile_ml<NL>(g, r, T, l0, [&](int l, int pos, int q, int lev) {
        if (q < 0) {
            for (int k = 0; k < 3; k++) sM[l][k][pos] = sV[l][k][pos] = 0.0;
            for (int k = 0; k < 6; k++) sA[l][k][pos] = 0.0;
            return;
        }
        const double rho = LVP(f.Rhok, lev, q), w = LVP(f.Wk, lev, q);
        const double wr = w / rho;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double m = VEC(f.Mhk, k, lev, q);
            sM[l][k][pos] = m;
            sV[l][k][pos] = m / rho + wr * g.fr[(size_t)k * S + q];
            sA[l][1 + k][pos] = m / rho;
        }
        sA[l][0][pos] = rho;
        sA[l][4][pos] = wr;
        sA[l][5][pos] = g.energy ? (g.Cp - Rd) * LVP(f.T, lev, q) : Rd * LVP(f.T, lev, q);
    });
    __syncthreads();
    const int p = T.p;
    double rs[NL], nf3[NL][3];
#pragma unroll
    for (int l = 0; l < NL; l++) {
        rs[l] = rscale_at(g, min(l0 + l, nv - 1));
        nf3[l][0] = nf3[l][1] = nf3[l][2] = 0.0;
    }
#pragma unroll
    for (int kk = 0; kk < 3; kk++) {
        double d[7];
#pragma unroll
        for (int s = 0; s < 7; s++) d[s] = g.div[(size_t)(s * 3 + kk) * S + p];
#pragma unroll
        for (int l = 0; l < NL; l++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double sum = d[0] * sV[l][c][T.pos(0)] * sM[l][kk][T.pos(0)];
#pragma unroll
                for (int s = 1; s < 7; s++) sum = sum + d[s] * sV[l][c][T.pos(s)] * sM[l][kk][T.pos(s)];
                nf3[l][c] += rs[l] * sum;
            }
    }
    double *Adv = f.Adv;
#pragma unroll
    for (int l = 0; l < NL; l++) {
        const int lev = l0 + l;
        if (lev < nv) {
#pragma unroll
            for (int c = 0; c < 3; c++) VEC(Adv, c, lev, p) = nf3[l][c] + VEC(Adv, c, lev, p);
        }
    }
    if (!hydiff) return;
    auto a = [&](int var, int l, int s) -> double { return sA[l][var][T.pos(s)]; };
    double lap[6][NL];
    laplacians_ml<6, NL>(g, p, T.nf, rs, a, lap);
    double *diff = f.diff;
#pragma unroll
    for (int l = 0; l < NL; l++) {
        const int lev = l0 + l;
        if (lev < nv) {
#pragma unroll
            for (int var = 0; var < 6; var++) VEC(diff, var, lev, p) = lap[var][l];
        }
    }
}

// Last hyperdiffusion pass + Correct_Horizontal, and divergence-damping pass 0
// on the stage values (k_diff2_divm0, FP64). DD = false: the hyperdiffusion
// pass alone (divergence-damping pass 0 then runs as k_divm0_ml on the stage
// values). 323 -> 218 + 104 us at g6 L48, and the following divergence-damping
// pass 1 runs faster after it (76 -> 69 us).
template <int NL, bool DD = true>
__global__ ML_STAGE_BOUNDS void k_diff2_divm0_ml(Geo g, Fields f, int hydiff, int divdamp, double sign) {
    int r, l0;
    if (!block_tile_ml<NL>(g, r, l0)) return;
    if (!DD) divdamp = 0;
    const size_t S = g.S;
    const int nv = g.nv;
    const double *diff = f.diff;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (!pole_ctx_ml<NL>(g, l0, c, lev)) return;
        diff2_divm0_point<false>(
            g, f, c.p, lev, 5, true, hydiff, divdamp, sign,
            [&](int s, int var) { return VEC(diff, var, lev, c.gi(s)); },
            [&](int s) { return LVP(f.Rhok, lev, c.gi(s)); },
            [&](int s, int k) { return VEC(f.Mhk, k, lev, c.gi(s)); },
            [&](int s) { return LVP(f.Whk, lev, c.gi(s)); }, [&](int s) { return LVP(f.Whk, lev + 1, c.gi(s)); });
        return;
    }
    __shared__ double sD[NL][6][NHL2], sR[NL][NHL2], sM[NL][3][NHL2], sW[NL + 1][NHL2], sLv[2][NL];
    const TileThread T = tile_thread(g, r);
    if (divdamp) level_factors_ml<NL>(g, l0, sLv);
    stage_tile_ml<NL>(g, r, T, l0, [&](int l, int pos, int q, int lev) {
        if (hydiff) {
#pragma unroll
            for (int k = 0; k < 6; k++) sD[l][k][pos] = q < 0 ? 0.0 : VEC(diff, k, lev, q);
            sR[l][pos] = q < 0 ? 0.0 : LVP(f.Rhok, lev, q);
        }
        if (divdamp) {
#pragma unroll
            for (int k = 0; k < 3; k++) sM[l][k][pos] = q < 0 ? 0.0 : VEC(f.Mhk, k, lev, q);
            // interfaces l0 .. l0 + NL, clamped to nv (see k_divm0_ml)
            sW[l][pos] = q < 0 ? 0.0 : LVP(f.Whk, min(l0 + l, nv), q);
            if (l == NL - 1) sW[NL][pos] = q < 0 ? 0.0 : LVP(f.Whk, min(l0 + NL, nv), q);
        }
    });
    __syncthreads();
    const int p = T.p;
    if (hydiff) {
        double rs[NL];
#pragma unroll
        for (int l = 0; l < NL; l++) rs[l] = rscale_at(g, min(l0 + l, nv - 1));
        auto a = [&](int var, int l, int s) -> double {
            const int q = T.pos(s);
            return var == 0 ? sD[l][0][q] : sD[l][var][q] * sR[l][q];
        };
        double lap[6][NL];
        laplacians_ml<6, NL>(g, p, T.nf, rs, a, lap);
#pragma unroll
        for (int l = 0; l < NL; l++) {
            if (l0 + l < nv) {
                double out[6];
#pragma unroll
                for (int var = 0; var < 6; var++) out[var] = lap[var][l];
                diff2_tail<false>(g, f, p, l0 + l, false, sign, out, sR[l][T.pos(0)]);
            }
        }
    }
    if (divdamp) {
        double o[NL][3];
        divm_ml<false, NL>(
            g, p, l0, T.nf, [&](int l, int s, int k) { return sM[l][k][T.pos(s)]; },
            [&](int l, int s) { return sW[l][T.pos(s)]; }, [&](int l, int s) { return sW[l + 1][T.pos(s)]; }, sLv,
            o);
        double *divg = f.divg;
#pragma unroll
        for (int l = 0; l < NL; l++) {
            const int lev = l0 + l;
            if (lev < nv) {
#pragma unroll
                for (int k = 0; k < 3; k++) VEC(divg, k, lev, p) = o[l][k];
            }
        }
    }
}

} // namespace thor
