// HIP kernels of the HE-VI dynamical-core step for gfx950 (FP64 path).
//
// Each kernel fuses one or more of THOR's kernels; inside a fused kernel the
// per-cell arithmetic is the reference's, operation for operation (compiled
// with -ffp-contract=off), so the FP64 path reproduces THOR's own results.
// Horizontal stencils run one 256-thread block per (rhombus tile, level): the
// tile and its one-point halo are staged in LDS, one thread per point; the two
// poles are one extra block that gathers its neighbours directly. Column
// kernels (vertical sweeps, the Thomas solve) map one thread to one column.
#pragma once
#include <type_traits>

#include "dyn_common.h"
#include "bf16.h"
#include "mlp_layer.h"
#include "portable_math.h"

namespace thor {

// Minimum resident 256-thread blocks per CU requested from the compiler for
// the LDS-tiled stencil kernels (bounds their register use).
#ifndef THOR_STENCIL_MIN_BLOCKS
#define THOR_STENCIL_MIN_BLOCKS 2
#endif
#define STENCIL_BOUNDS __launch_bounds__(256, THOR_STENCIL_MIN_BLOCKS)
// Unroll factor of the 7-slot operator sums (register use vs load latency).
#ifndef THOR_SLOT_UNROLL
#define THOR_SLOT_UNROLL 1
#endif
#define PRAGMA_STR(x) _Pragma(#x)
#define SLOT_UNROLL(n) PRAGMA_STR(unroll n)

// pow inside the dynamical-core step (the equation of state in the column
// preparation and the density/pressure update, r^2 and r^3 of the deep
// model): the portable one in both precision forms and both builds. It is
// several times cheaper than the device libm's and within 2.5 (1 + |y log x|)
// ulp of the exact value (tests/test_oracle.py). The forcings keep MPOW.
#define XPOW(x, y) pm::pow((x), (y))

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

// ---------------------------------------------------------------- LDS tiles
// A stencil block is one 16x16 rhombus tile at one level. Its values and the
// one-point halo are staged in LDS in THOR's 18x18 map order (point (x, y) at
// (y + 1) * 18 + x + 1), so every field value is read from HBM once per block.
// The two poles are handled by one extra block with direct gathers.
constexpr int NHL = 18, NHL2 = NHL * NHL, NHALO = 66;

// LDS position of halo entry t (left, right, bottom, top sides, then the
// (-1,-1) and (16,16) corners; the other two corners are never read).
__host__ __device__ __forceinline__ int halo_pos(int t) {
    if (t < 16) return (t + 1) * NHL;
    if (t < 32) return (t - 16 + 1) * NHL + 17;
    if (t < 48) return t - 32 + 1;
    if (t < 64) return 17 * NHL + (t - 48) + 1;
    if (t == 64) return 0;
    return 17 * NHL + 17;
}

// Block -> (tile, level) mapping. Workgroups are dispatched round-robin over
// the 8 XCDs (block i runs on XCD i % 8); each XCD is given a contiguous range
// of tiles and walks all levels of a tile consecutively, so a tile's geometry
// and its neighbours' halo lines stay in that XCD's L2 across levels.
__device__ __forceinline__ bool block_tile(const Geo &g, int &r, int &lev) {
    const int i = blockIdx.x, xcd = i & 7, idx = i >> 3;
    r = xcd * g.tiles_per_xcd + idx / g.nv;
    lev = idx % g.nv;
    return r <= g.nr;
}

struct TileThread {
    int p, nf, ir;
    // LDS position of stencil slot s: 0 = self, 1..6 = pt1..pt6 (a pentagon's pt5
    // is its pt6 position, thor_fastmodes.h:126).
    __device__ __forceinline__ int pos(int s) const {
        // byte s of kOff: 0, +18, +19, +1, -18, -19, -1 (two's complement)
        constexpr unsigned long long kOff = 0x00FFEDEE01131200ull;
        int o = (int)(signed char)((kOff >> (8 * s)) & 0xff);
        if (s == 5 && nf == 5) o = -1;
        return ir + o;
    }
};

__device__ __forceinline__ TileThread tile_thread(const Geo &g, int r) {
    TileThread T;
    const int t = threadIdx.x, x = t & 15, y = t >> 4;
    T.p = r * kTile + t;
    T.nf = g.nfc[T.p];
    T.ir = (y + 1) * NHL + x + 1;
    return T;
}

// put(pos, q) stores the values of global point q (q < 0: zeros, the missing
// pentagon corner) at LDS position pos; called for the thread's own point and,
// for threads 0..65, one halo point.
template <class Put> __device__ __forceinline__ void stage_tile(const Geo &g, int r, const TileThread &T, Put put) {
    put(T.ir, T.p);
    const int t = threadIdx.x;
    if (t < NHALO) put(halo_pos(t), g.halo[r * NHALO + t]);
}

// Laplacians of NV fields around point p, all in one pass over the triangles
// so each geometry coefficient is read once (Diffusion_Op, thor_diff.h:359-494;
// poles :524-880). a(v, 0) is field v at the centre, a(v, 1 + m) at ring
// neighbour m; nf = 6 (hexagon) or 5 (pentagon, pole). The arithmetic of each
// field is exactly the reference's.
template <int NV, class GetA>
__device__ __forceinline__ void laplacians(const Geo &g, int p, int nf, double rscale, GetA a, double out[NV]) {
    const size_t S = g.S;
    double a0[NV], a1[NV], a2[NV], lap[NV], lx[NV], ly[NV], lz[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        a0[v] = a(v, 0);
        a1[v] = a(v, 1);
        a2[v] = a(v, 2);
        lap[v] = 0.0;
    }
    const double vdiff = 0.5 * rscale;
#pragma unroll 1
    for (int j = 0; j < nf; j++) {
        const int jp1 = j + 1 == nf ? 0 : j + 1;
        const int jp2 = jp1 + 1 == nf ? 0 : jp1 + 1;
        const double aTr2 = g.areasTr[(size_t)jp1 * S + p];
        const double AT2 = rscale / aTr2;
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
            const double AT1 = rscale / aTr;
            const double c0 = g.areas[p], c1 = g.areas[S + p], c2 = g.areas[2 * S + p];
            const double ti0x = g.nti[p], ti0y = g.nti[S + p], ti0z = g.nti[2 * S + p];
            const double te0x = g.nte[p], te0y = g.nte[S + p], te0z = g.nte[2 * S + p];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                const double astar = (c0 * a0[v] + c1 * a1[v] + c2 * a2[v]) / aTr;
                const double lap1 = (0.5 * (a1[v] + a0[v]) - astar);
                const double lap2 = (0.5 * (a1[v] + a2[v]) - astar);
                const double lap3 = (0.5 * (a0[v] + a2[v]) - astar);
                lx[v] = (-lap1 * ti0x + lap2 * te0x + lap3 * ti1x) * AT1;
                ly[v] = (-lap1 * ti0y + lap2 * te0y + lap3 * ti1y) * AT1;
                lz[v] = (-lap1 * ti0z + lap2 * te0z + lap3 * ti1z) * AT1;
            }
        }
#pragma unroll
        for (int v = 0; v < NV; v++) {
            const double a3 = a(v, 1 + jp2);
            const double astar = (b0 * a0[v] + b1 * a2[v] + b2 * a3) / aTr2;
            const double lap4 = (0.5 * (a0[v] + a2[v]) - astar);
            const double lap5 = (0.5 * (a2[v] + a3) - astar);
            const double lap6 = (0.5 * (a0[v] + a3) - astar);
            const double x2 = (-lap4 * ti1x + lap5 * te1x + lap6 * ti2x) * AT2;
            const double y2 = (-lap4 * ti1y + lap5 * te1y + lap6 * ti2y) * AT2;
            const double z2 = (-lap4 * ti1z + lap5 * te1z + lap6 * ti2z) * AT2;
            lap[v] += ((lx[v] + x2) * oax + (ly[v] + y2) * oay + (lz[v] + z2) * oaz) * vdiff;
            lx[v] = x2;
            ly[v] = y2;
            lz[v] = z2;
            a2[v] = a3;
        }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) out[v] = lap[v];
}

// Divergence-damping operator at point p (DivM_Op, thor_div.h:49-315 and the
// pole variant :317-535, whose vertical means weight the third corner with
// areas[.][1]). a(s, k): momentum (pass 0) or divg (pass 1) at slot s;
// wl(s)/wt(s): Wh at the layer's lower/upper interface (pass 0 only).
template <bool LAST, class GetA, class GetWl, class GetWt>
__device__ __forceinline__ void divm_nf(const Geo &g, int p, int lev, int nf, bool pole, GetA a, GetWl wl, GetWt wt,
                                        double out[3]) {
    // Triangle j's quantities (vertical mean, dw/dz, divergence lap1) are the
    // same expressions as the previous iteration's second-triangle ones
    // (thor_div.h:207-262), so they are carried instead of recomputed, along
    // with the geometry of the shared corner.
    const size_t S = g.S;
    const double rs = rscale_at(g, lev);
    const double AT0 = 0.5 * rs;
    double lapx = 0.0, lapy = 0.0, lapz = 0.0;
    const double dz = LAST ? 0.0 : 1.0 / (g.Alth[lev] - g.Alth[lev + 1]);
    const double a0x = a(0, 0), a0y = a(0, 1), a0z = a(0, 2);
    const double wl0 = LAST ? 0.0 : wl(0), wt0 = LAST ? 0.0 : wt(0);
#define AR(j, k) g.areas[(size_t)((j)*3 + (k)) * S + p]
#define TI(j, k) g.nti[(size_t)((j)*3 + (k)) * S + p]
#define TE(j, k) g.nte[(size_t)((j)*3 + (k)) * S + p]
    // corner slot 1 + jp1 values, carried to the next triangle
    double a2x = a(1, 0), a2y = a(1, 1), a2z = a(1, 2), wl2 = 0.0, wt2 = 0.0;
    double lap2 = 0.0, dwdz2 = 0.0;
    double ti1x = TI(0, 0), ti1y = TI(0, 1), ti1z = TI(0, 2);  // nvecti of the current first edge
    double aT1 = g.areasTr[p];
    if (!LAST) {
        wl2 = wl(1);
        wt2 = wt(1);
    }
#pragma unroll 1
    for (int j = 0; j < nf; j++) {
        const int jp1 = j + 1 == nf ? 0 : j + 1;
        const int jp2 = jp1 + 1 == nf ? 0 : jp1 + 1;
        const int s2 = 1 + jp1, s3 = 1 + jp2;
        const double a1x = a2x, a1y = a2y, a1z = a2z, wl1 = wl2, wt1 = wt2;
        a2x = a(s2, 0);
        a2y = a(s2, 1);
        a2z = a(s2, 2);
        if (!LAST) {
            wl2 = wl(s2);
            wt2 = wt(s2);
        }
        const double aT2 = g.areasTr[(size_t)jp1 * S + p];
        const double a3x = a(s3, 0), a3y = a(s3, 1), a3z = a(s3, 2);
        double dwdz = 0.0;
        if (!LAST) {
            if (j == 0) {
                const double w1 = pole ? AR(j, 1) : AR(j, 2);
                const double meanwl = (AR(j, 0) * wl0 + AR(j, 1) * wl1 + w1 * wl2) / aT1;
                const double meanwt = (AR(j, 0) * wt0 + AR(j, 1) * wt1 + w1 * wt2) / aT1;
                dwdz = (meanwl - meanwt) * dz;
            } else {
                dwdz = dwdz2;
            }
            const double wl3 = wl(s3), wt3 = wt(s3);
            const double r0 = AR(jp1, 0), r1 = AR(jp1, 1);
            const double w2 = pole ? r1 : AR(jp1, 2);
            const double meanwl2 = (r0 * wl0 + r1 * wl2 + w2 * wl3) / aT2;
            const double meanwt2 = (r0 * wt0 + r1 * wt2 + w2 * wt3) / aT2;
            dwdz2 = (meanwl2 - meanwt2) * dz;
        }
        const double AT2 = 0.5 * rs / aT2;
        const double ti2x = TI(jp1, 0), ti2y = TI(jp1, 1), ti2z = TI(jp1, 2);
        const double a02x = a0x + a2x, a02y = a0y + a2y, a02z = a0z + a2z;
        double lap1;
        if (j == 0) {
            const double a01x = a0x + a1x, a01y = a0y + a1y, a01z = a0z + a1z;
            const double a12x = a1x + a2x, a12y = a1y + a2y, a12z = a1z + a2z;
            const double AT1 = 0.5 * rs / aT1;
            lap1 = (-a01x * ti1x - a01y * ti1y - a01z * ti1z + a12x * TE(j, 0) + a12y * TE(j, 1) + a12z * TE(j, 2)
                    + a02x * ti2x + a02y * ti2y + a02z * ti2z)
                   * AT1;
        } else {
            lap1 = lap2;
        }
        const double a23x = a2x + a3x, a23y = a2y + a3y, a23z = a2z + a3z;
        const double a03x = a0x + a3x, a03y = a0y + a3y, a03z = a0z + a3z;
        lap2 = (-a02x * ti2x - a02y * ti2y - a02z * ti2z + a23x * TE(jp1, 0) + a23y * TE(jp1, 1) + a23z * TE(jp1, 2)
                + a03x * TI(jp2, 0) + a03y * TI(jp2, 1) + a03z * TI(jp2, 2))
               * AT2;
        lapx += (lap1 + dwdz + lap2 + dwdz2) * g.noa[(size_t)(j * 3 + 0) * S + p] * AT0;
        lapy += (lap1 + dwdz + lap2 + dwdz2) * g.noa[(size_t)(j * 3 + 1) * S + p] * AT0;
        lapz += (lap1 + dwdz + lap2 + dwdz2) * g.noa[(size_t)(j * 3 + 2) * S + p] * AT0;
        ti1x = ti2x;
        ti1y = ti2y;
        ti1z = ti2z;
        aT1 = aT2;
    }
#undef AR
#undef TI
#undef TE
    if (LAST) {
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

// Pole point of a stencil kernel: slot s is read from global memory through
// the ring (the pole's five neighbours; slot 6 repeats slot 5 with a zero
// coefficient).
struct PoleCtx {
    int p;
    const int *ring;
    size_t S;
    __device__ __forceinline__ int gi(int s) const { return s == 0 ? p : ring[(size_t)(s - 1) * S + p]; }
};
__device__ __forceinline__ bool pole_ctx(const Geo &g, PoleCtx &c) {
    if (threadIdx.x >= 2) return false;
    c.p = g.P - 2 + threadIdx.x;
    c.ring = g.ring;
    c.S = g.S;
    return true;
}

// ------------------------------------------------ mixed-precision operators
// The hyperdiffusion Laplacian and the divergence-damping operator above are
// linear in their inputs, with coefficients fixed by the geometry and the
// level's rscale. The mixed-precision path applies them as FP32 per-point
// weights (Geo::wlap, wdd, udd), which k_probe_weights derives from the FP64
// operators themselves by feeding them unit inputs.

// Laplacians of NV fields: rs^2 sum_{s>=1} wlap[s] (a_s - a_0). The weights
// of a Laplacian sum to zero, so the differences form keeps FP32 precision
// for fields with a large mean (R*T).
template <int NV, class GetA>
__device__ __forceinline__ void laplacians_w(const Geo &g, int p, float rs2, GetA a, float out[NV]) {
    const size_t S = g.S;
    float w[6];
#pragma unroll
    for (int s = 0; s < 6; s++) w[s] = g.wlap[(size_t)(s + 1) * S + p];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const float a0 = a(v, 0);
        float sum = 0.f;
#pragma unroll
        for (int s = 0; s < 6; s++) sum += w[s] * (a(v, s + 1) - a0);
        out[v] = rs2 * sum;
    }
}

// Divergence-damping pass (DivM_Op): pass 0 (LAST = false) from the momentum
// a and the vertical-velocity differences dw(s) = wl(s) - wt(s); pass 1
// (LAST = true) from pass 0's output, with the radial part removed and
// -Kdhz applied as in the FP64 operator.
template <bool LAST, class GetA, class GetDW>
__device__ __forceinline__ void divm_w(const Geo &g, int p, int lev, GetA a, GetDW dw, float out[3]) {
    const size_t S = g.S;
    const double rsd = rscale_at(g, lev);
    const float rs = (float)rsd, rs2 = (float)(rsd * rsd);
    float av[7][3];
#pragma unroll
    for (int s = 0; s < 7; s++)
#pragma unroll
        for (int k = 0; k < 3; k++) av[s][k] = a(s, k);
    float d[7];
    if (!LAST) {
#pragma unroll
        for (int s = 0; s < 7; s++) d[s] = dw(s);
    }
    const float dz = LAST ? 0.f : (float)(1.0 / (g.Alth[lev] - g.Alth[lev + 1]));
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float sum = 0.f;
#pragma unroll
        for (int s = 0; s < 7; s++)
#pragma unroll
            for (int k = 0; k < 3; k++) sum += g.wdd[(size_t)(c * 21 + s * 3 + k) * S + p] * av[s][k];
        out[c] = rs2 * sum;
        if (!LAST) {
            float su = 0.f;
#pragma unroll
            for (int s = 0; s < 7; s++) su += g.udd[(size_t)(c * 7 + s) * S + p] * d[s];
            out[c] += rs * dz * su;
        }
    }
    if (LAST) {
        const float sdiff = (float)g.Kdhz[lev];
        const float fx = (float)g.fr[p], fy = (float)g.fr[S + p], fz = (float)g.fr[2 * S + p];
        const float lapr = fx * out[0] + fy * out[1] + fz * out[2];
        out[0] = -(out[0] - fx * lapr) * sdiff;
        out[1] = -(out[1] - fy * lapr) * sdiff;
        out[2] = -(out[2] - fz * lapr) * sdiff;
    }
}

// Derives the mixed-precision weights from the FP64 operators: for every
// point, the Laplacian of the 7 unit slot fields and the divergence-damping
// pass 0 of the 21 unit momentum fields and the 7 unit vertical-velocity
// fields, all at rscale = 1 (gu is the grid with DeepModel off) and, for the
// vertical part, divided by level 0's 1/dz. One thread per point.
__global__ __launch_bounds__(256) void k_probe_weights(Geo gu, float *wlap, float *wdd, float *udd) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= gu.P) return;
    const size_t S = gu.S;
    const bool pole = p >= gu.P - 2;
    const int nf = pole ? 5 : gu.nfc[p];
    {
        double out[7];
        laplacians<7>(gu, p, nf, 1.0, [&](int v, int s) { return v == s ? 1.0 : 0.0; }, out);
        for (int s = 0; s < 7; s++) wlap[(size_t)s * S + p] = (float)out[s];
    }
    auto zero = [&](int) { return 0.0; };
    for (int s0 = 0; s0 < 7; s0++) {
        for (int k0 = 0; k0 < 3; k0++) {
            double o3[3];
            divm_nf<false>(
                gu, p, 0, nf, pole, [&](int s, int k) { return (s == s0 && k == k0) ? 1.0 : 0.0; }, zero, zero, o3);
            for (int c = 0; c < 3; c++) wdd[(size_t)(c * 21 + s0 * 3 + k0) * S + p] = (float)o3[c];
        }
        double o3[3];
        divm_nf<false>(
            gu, p, 0, nf, pole, [&](int, int) { return 0.0; }, [&](int s) { return s == s0 ? 1.0 : 0.0; }, zero, o3);
        const double dz0 = 1.0 / (gu.Alth[0] - gu.Alth[1]);
        for (int c = 0; c < 3; c++) udd[(size_t)(c * 7 + s0) * S + p] = (float)(o3[c] / dz0);
    }
}

// ===================================================================== K_A
// Compute_Temperature_H_Pt_Geff (thor_auxiliary.h:55-309, entropy form) and
// the vertical advection + Coriolis part of Compute_Advec_Cori2
// (thor_adv_cor.h:366-536), written to Adv; k_adv_diff1 adds the horizontal
// part. The reference runs both as one serial sweep per column; here one
// thread takes one (column, level) pair, lev in [0, nv], and recomputes the
// values the sweep carries from the level below from their own inputs, so
// every output is the same number the sweep produces.
// Blocks of 256 columns x one level: block i runs on XCD i % 8, each XCD takes
// a contiguous range of column blocks and walks their levels 0..nv
// consecutively, so the neighbouring levels a thread reads are in that XCD's
// L2 from the blocks just before it.
template <bool MX> __global__ __launch_bounds__(256) void k_column_prep(Geo g, Fields f) {
    const int i = blockIdx.x, idx = i >> 3;
    const int pbx = ((g.P + 255) / 256 + 7) / 8;  // column blocks per XCD
    const int pb = (i & 7) * pbx + idx / (g.nv + 1);
    const int lev = idx % (g.nv + 1);  // 0..nv: centre lev (if < nv) and interface lev
    const int p = pb * 256 + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const double Rd = g.Rd, Cp = g.Cp, P_Ref = g.P_Ref, Gravit = g.Gravit, A = g.A;
    const double Cv = Cp - Rd, CvoCp = Cv / Cp;
    const double *Alt = g.Alt, *Alth = g.Alth;
    const double *pk = f.pk, *Rho = f.Rhok, *Wh = f.Whk;
    auto h_at = [&](int l) {
        const double pr = LV(pk, l), r = LV(Rho, l);
        const double temperature = pr / (r * Rd);
        return Cv * temperature + pr / r;
    };
    auto pt_at = [&](int l) {
        const double pr = LV(pk, l), r = LV(Rho, l);
        return (P_Ref / (Rd * r)) * XPOW(pr / P_Ref, CvoCp);
    };
    // interpolation weights of the point xi between xim and xip
    auto wts = [](double xi, double xim, double xip, double &intt, double &intl) {
        intt = (xi - xip) / (xim - xip);
        intl = (xi - xim) / (xip - xim);
    };
    // -(Wh/rho_h) dp/dz at interior interface k (the sweep's gtilht), 0 at k = 0
    auto gtil_tilde = [&](int k) {
        if (k == 0) return 0.0;
        const double dp_dz = (LV(pk, k) - LV(pk, k - 1)) / (Alt[k] - Alt[k - 1]);
        double intt, intl;
        wts(Alth[k], Alt[k - 1], Alt[k], intt, intl);
        const double rhoh = LV(Rho, k - 1) * intt + LV(Rho, k) * intl;
        return -(LV(Wh, k) / rhoh) * dp_dz;
    };

    // ---- centre values
    if (lev < nv) {
        const double pressure = LV(pk, lev), rho = LV(Rho, lev);
        const double temperature = pressure / (rho * Rd);
        LV(f.T, lev) = temperature;
        const double hv = Cv * temperature + pressure / rho;
        const double ptv = (P_Ref / (Rd * rho)) * XPOW(pressure / P_Ref, CvoCp);
        LV(f.h, lev) = hv;
        LV(f.pt, lev) = ptv;
        if constexpr (MX) {  // FP32 copies for the substep flux divergences of the deviation
            f.mhpt[(size_t)lev * S + p] = (float)hv;
            f.mhpt[((size_t)nv + lev) * S + p] = (float)ptv;
        }
        // gtil at this centre: the sweep forms it at interface lev + 1 from the
        // interface values below and above, weighted at (Alt, Alth, Alth+1) of
        // that sweep step
        double intt, intl;
        if (lev < nv - 1) {
            wts(Alt[lev + 1], Alth[lev + 1], Alth[lev + 2], intt, intl);
            LV(f.gtil, lev) = gtil_tilde(lev) * intt + gtil_tilde(lev + 1) * intl;
        } else {
            wts(Alt[nv - 1], Alth[nv - 1], Alth[nv], intt, intl);
            const double gtilht = 0.0;
            LV(f.gtil, nv - 1) = gtil_tilde(nv - 1) * intt + gtilht * intl;
        }
    }
    // ---- interface values
    if (lev == 0) {
        const double rho = LV(Rho, 0), alt = Alt[0];
        const double extr = (-alt - Alt[1]) / (Alt[1] - alt);
        LV(f.hh, 0) = h_at(0);
        double pl;
        if (g.ghv) pl = LV(pk, 1) - rho * Gravit * XPOW(A / (A + Alt[0]), 2) * (-alt - Alt[1]);
        else pl = LV(pk, 1) - rho * Gravit * (-alt - Alt[1]);
        const double rhol = LV(Rho, 1) + (LV(Rho, 1) - rho) * extr;
        const double rhoh = 0.5 * (rho + rhol);
        const double dz = 2.0 * Alt[0];
        const double dpdz = (LV(pk, 0) - pl) / dz;
        LV(f.gtilh, 0) = -(1.0 / rhoh) * dpdz;
        const double pt0 = pt_at(0), pt1 = pt_at(1);
        LV(f.pth, 0) = pt1 + (pt1 - pt0) * (Alth[0] - Alt[1]) / (Alt[1] - Alt[0]);
    } else if (lev == nv) {
        const double alt = Alt[nv - 1], rho = LV(Rho, nv - 1);
        const double extr = (2 * Alth[nv] - alt - Alt[nv - 2]) / (Alt[nv - 1] - Alt[nv - 2]);
        LV(f.hh, nv) = h_at(nv - 1);
        const double rhop = LV(Rho, nv - 2) + (LV(Rho, nv - 1) - LV(Rho, nv - 2)) * extr;
        const double rhoh = 0.5 * (rho + rhop);
        const double ptm = pt_at(nv - 2), ptc = pt_at(nv - 1);
        LV(f.pth, nv) = ptm + (ptc - ptm) * extr;
        const double dz = 2.0 * (Alth[nv] - Alt[nv - 1]);
        double pp;
        if (g.ghv)
            pp = LV(pk, nv - 2)
                 - LV(Rho, nv - 1) * Gravit * XPOW(A / (A + Alt[nv - 1]), 2) * (2 * Alth[nv] - Alt[nv - 1] - Alt[nv - 2]);
        else
            pp = LV(pk, nv - 2) - LV(Rho, nv - 1) * Gravit * (2 * Alth[nv] - Alt[nv - 1] - Alt[nv - 2]);
        const double dpdz = (pp - LV(pk, nv - 1)) / dz;
        LV(f.gtilh, nv) = -(1.0 / rhoh) * dpdz;
    } else {
        const double dz = Alt[lev] - Alt[lev - 1];
        const double dp_dz = (LV(pk, lev) - LV(pk, lev - 1)) / dz;
        double intt, intl;
        wts(Alth[lev], Alt[lev - 1], Alt[lev], intt, intl);
        LV(f.hh, lev) = h_at(lev - 1) * intt + h_at(lev) * intl;
        LV(f.pth, lev) = pt_at(lev - 1) * intt + pt_at(lev) * intl;
        const double rhoh = LV(Rho, lev - 1) * intt + LV(Rho, lev) * intl;
        LV(f.gtilh, lev) = -(1.0 / rhoh) * dp_dz;
    }
    if (lev == nv) return;

    // ---- vertical advection of momentum and Coriolis (Compute_Advec_Cori2)
    const double fr[3] = {g.fr[p], g.fr[S + p], g.fr[2 * S + p]};
    const double Omega = g.Omega;
    auto vel = [&](int l, double v[3]) {
        const double rho = LV(Rho, l), w = LV(f.Wk, l);
#pragma unroll
        for (int c = 0; c < 3; c++) v[c] = VEC(f.Mhk, c, l, p) / rho + (w / rho) * fr[c];
    };
    const double althl = Alth[lev], altht = Alth[lev + 1], alt = Alt[lev];
    double vc[3], dvl[3], dvt[3], intt, intl;
    vel(lev, vc);
    if (lev == 0) {
        dvl[0] = dvl[1] = dvl[2] = 0.0;
    } else {
        double vl[3];
        vel(lev - 1, vl);
        const double whl = LV(Wh, lev);
        wts(althl, Alt[lev - 1], alt, intt, intl);
#pragma unroll
        for (int c = 0; c < 3; c++) dvl[c] = (vl[c] * intt + vc[c] * intl) * whl;
    }
    if (lev == nv - 1) {
        dvt[0] = dvt[1] = dvt[2] = 0.0;
    } else {
        double vt[3];
        vel(lev + 1, vt);
        const double wht = LV(Wh, lev + 1);
        wts(altht, alt, Alt[lev + 1], intt, intl);
#pragma unroll
        for (int c = 0; c < 3; c++) dvt[c] = (vc[c] * intt + vt[c] * intl) * wht;
    }
    double r2p, r2m, r2l;
    if (g.deep) {
        r2p = XPOW(altht + A, 2.0);
        r2m = XPOW(alt + A, 2.0);
        r2l = XPOW(althl + A, 2.0);
    } else {
        r2p = r2m = r2l = 1.0;
    }
    const double dzi = 1.0 / ((althl - altht) * r2m);
    double dav[3];
#pragma unroll
    for (int c = 0; c < 3; c++) dav[c] = (dvl[c] * r2l - dvt[c] * r2p) * dzi;
    const double rho = LV(Rho, lev);
    double Cx, Cy, Cz;
    if (g.deep) {
        Cx = -2.0 * Omega * vc[1] * rho;
        Cy = 2.0 * Omega * vc[0] * rho;
        Cz = 0.0;
    } else {
        Cx = 2 * Omega * (vc[2] * fr[2] * fr[1] - vc[1] * XPOW(fr[2], 2)) * rho;
        Cy = -2 * Omega * (vc[2] * fr[2] * fr[0] - vc[0] * XPOW(fr[2], 2)) * rho;
        Cz = 2 * Omega * (vc[1] * fr[2] * fr[0] - vc[0] * fr[2] * fr[1]) * rho;
    }
    auto *Adv = FS<MX>::Adv(f);
    VEC(Adv, 0, lev, p) = dav[0] + Cx;
    VEC(Adv, 1, lev, p) = dav[1] + Cy;
    VEC(Adv, 2, lev, p) = dav[2] + Cz;
}

// The same outputs with LV consecutive levels per thread: the per-level
// values each output needs from its own and the neighbouring levels
// (temperature, h, pt, the advecting velocity, the interface gtil term) are
// formed once per level into a window in registers instead of once per use,
// with the same expressions, so the outputs equal k_column_prep's bit for bit
// (one power per level instead of three). Block i: XCD i % 8, a contiguous
// range of column blocks per XCD, level groups of one column block
// consecutive.
// blocks per CU of the column preparation (5 spills 18 VGPRs: 78 -> 99 us at g6 L48)
#ifndef THOR_COLPREP_MINB
#define THOR_COLPREP_MINB 1
#endif
template <bool MX, int LV> __global__ __launch_bounds__(256, THOR_COLPREP_MINB) void k_column_prep_lv(Geo g, Fields f) {
    const int nv = g.nv, ngrp = (nv + 1 + LV - 1) / LV;
    const int i = blockIdx.x, idx = i >> 3;
    const int pbx = ((g.P + 255) / 256 + 7) / 8;  // column blocks per XCD
    const int pb = (i & 7) * pbx + idx / ngrp;
    const int l0 = (idx % ngrp) * LV;  // levels l0 .. l0 + LV - 1 (centre if < nv, interface if <= nv)
    const int p = pb * 256 + threadIdx.x;
    if (p >= g.P) return;
    const size_t S = g.S;
    const double Rd = g.Rd, Cp = g.Cp, P_Ref = g.P_Ref, Gravit = g.Gravit, A = g.A;
    const double Cv = Cp - Rd, CvoCp = Cv / Cp;
    const double *Alt = g.Alt, *Alth = g.Alth;
    const double *pk = f.pk, *Rho = f.Rhok, *Wh = f.Whk;
    const double fr[3] = {g.fr[p], g.fr[S + p], g.fr[2 * S + p]};
    auto wts = [](double xi, double xim, double xip, double &intt, double &intl) {
        intt = (xi - xip) / (xim - xip);
        intl = (xi - xim) / (xip - xim);
    };
    // window of levels m = l0 - 1 + j, j = 0 .. LV + 1 (centre values valid for 0 <= m < nv)
    constexpr int NW = LV + 2;
    double wp[NW], wr[NW], wT[NW], wh[NW], wpt[NW], wv[NW][3], wgt[NW], wwh[NW];
#pragma unroll
    for (int j = 0; j < NW; j++) {
        const int m = l0 - 1 + j;
        wwh[j] = (m >= 0 && m <= nv) ? LV(Wh, m) : 0.0;
        if (m >= 0 && m < nv) {
            const double pr = LV(pk, m), r = LV(Rho, m), w = LV(f.Wk, m);
            wp[j] = pr;
            wr[j] = r;
            const double temperature = pr / (r * Rd);
            wT[j] = temperature;
            wh[j] = Cv * temperature + pr / r;
            wpt[j] = (P_Ref / (Rd * r)) * XPOW(pr / P_Ref, CvoCp);
#pragma unroll
            for (int c = 0; c < 3; c++) wv[j][c] = VEC(f.Mhk, c, m, p) / r + (w / r) * fr[c];
        } else {
            wp[j] = wr[j] = wT[j] = wh[j] = wpt[j] = 0.0;
            wv[j][0] = wv[j][1] = wv[j][2] = 0.0;
        }
    }
    // -(Wh/rho_h) dp/dz at interior interface m (the sweep's gtilht), 0 at m = 0
#pragma unroll
    for (int j = 0; j < NW; j++) {
        const int m = l0 - 1 + j;
        if (m >= 1 && m < nv) {
            const double pm1 = j > 0 ? wp[j - 1] : LV(pk, m - 1), rm1 = j > 0 ? wr[j - 1] : LV(Rho, m - 1);
            const double dp_dz = (wp[j] - pm1) / (Alt[m] - Alt[m - 1]);
            double intt, intl;
            wts(Alth[m], Alt[m - 1], Alt[m], intt, intl);
            const double rhoh = rm1 * intt + wr[j] * intl;
            wgt[j] = -(wwh[j] / rhoh) * dp_dz;
        } else {
            wgt[j] = 0.0;
        }
    }
#pragma unroll
    for (int q = 0; q < LV; q++) {
        const int lev = l0 + q, j = q + 1;  // window index of lev
        if (lev > nv) break;
        // ---- centre values
        if (lev < nv) {
            const double pressure = wp[j], rho = wr[j];
            LV(f.T, lev) = wT[j];
            LV(f.h, lev) = wh[j];
            LV(f.pt, lev) = wpt[j];
            if constexpr (MX) {
                f.mhpt[(size_t)lev * S + p] = (float)wh[j];
                f.mhpt[((size_t)nv + lev) * S + p] = (float)wpt[j];
            }
            (void)pressure;
            (void)rho;
            double intt, intl;
            if (lev < nv - 1) {
                wts(Alt[lev + 1], Alth[lev + 1], Alth[lev + 2], intt, intl);
                LV(f.gtil, lev) = wgt[j] * intt + wgt[j + 1] * intl;
            } else {
                wts(Alt[nv - 1], Alth[nv - 1], Alth[nv], intt, intl);
                const double gtilht = 0.0;
                LV(f.gtil, nv - 1) = wgt[j] * intt + gtilht * intl;
            }
        }
        // ---- interface values
        if (lev == 0) {
            const double rho = wr[j], alt = Alt[0];
            const double extr = (-alt - Alt[1]) / (Alt[1] - alt);
            LV(f.hh, 0) = wh[j];
            double pl;
            if (g.ghv) pl = wp[j + 1] - rho * Gravit * XPOW(A / (A + Alt[0]), 2) * (-alt - Alt[1]);
            else pl = wp[j + 1] - rho * Gravit * (-alt - Alt[1]);
            const double rhol = wr[j + 1] + (wr[j + 1] - rho) * extr;
            const double rhoh = 0.5 * (rho + rhol);
            const double dz = 2.0 * Alt[0];
            const double dpdz = (wp[j] - pl) / dz;
            LV(f.gtilh, 0) = -(1.0 / rhoh) * dpdz;
            const double pt0 = wpt[j], pt1 = wpt[j + 1];
            LV(f.pth, 0) = pt1 + (pt1 - pt0) * (Alth[0] - Alt[1]) / (Alt[1] - Alt[0]);
        } else if (lev == nv) {
            // window index j - 1 is nv - 1; nv - 2 is j - 2 when q >= 1, else read directly
            const double alt = Alt[nv - 1], rho = wr[j - 1];
            const double rho2 = q >= 1 ? wr[j - 2] : LV(Rho, nv - 2), p2 = q >= 1 ? wp[j - 2] : LV(pk, nv - 2);
            double ptm;
            if (q >= 1) {
                ptm = wpt[j - 2];
            } else {
                const double pr = LV(pk, nv - 2), r = LV(Rho, nv - 2);
                ptm = (P_Ref / (Rd * r)) * XPOW(pr / P_Ref, CvoCp);
            }
            const double extr = (2 * Alth[nv] - alt - Alt[nv - 2]) / (Alt[nv - 1] - Alt[nv - 2]);
            LV(f.hh, nv) = wh[j - 1];
            const double rhop = rho2 + (wr[j - 1] - rho2) * extr;
            const double rhoh = 0.5 * (rho + rhop);
            const double ptc = wpt[j - 1];
            LV(f.pth, nv) = ptm + (ptc - ptm) * extr;
            const double dz = 2.0 * (Alth[nv] - Alt[nv - 1]);
            double pp;
            if (g.ghv)
                pp = p2 - wr[j - 1] * Gravit * XPOW(A / (A + Alt[nv - 1]), 2) * (2 * Alth[nv] - Alt[nv - 1] - Alt[nv - 2]);
            else
                pp = p2 - wr[j - 1] * Gravit * (2 * Alth[nv] - Alt[nv - 1] - Alt[nv - 2]);
            const double dpdz = (pp - wp[j - 1]) / dz;
            LV(f.gtilh, nv) = -(1.0 / rhoh) * dpdz;
        } else {
            const double dz = Alt[lev] - Alt[lev - 1];
            const double dp_dz = (wp[j] - wp[j - 1]) / dz;
            double intt, intl;
            wts(Alth[lev], Alt[lev - 1], Alt[lev], intt, intl);
            LV(f.hh, lev) = wh[j - 1] * intt + wh[j] * intl;
            LV(f.pth, lev) = wpt[j - 1] * intt + wpt[j] * intl;
            const double rhoh = wr[j - 1] * intt + wr[j] * intl;
            LV(f.gtilh, lev) = -(1.0 / rhoh) * dp_dz;
        }
        if (lev == nv) break;

        // ---- vertical advection of momentum and Coriolis (Compute_Advec_Cori2)
        const double Omega = g.Omega;
        const double althl = Alth[lev], altht = Alth[lev + 1], alt = Alt[lev];
        double dvl[3], dvt[3], intt, intl;
        const double *vc = wv[j];
        if (lev == 0) {
            dvl[0] = dvl[1] = dvl[2] = 0.0;
        } else {
            const double whl = wwh[j];
            wts(althl, Alt[lev - 1], alt, intt, intl);
#pragma unroll
            for (int c = 0; c < 3; c++) dvl[c] = (wv[j - 1][c] * intt + vc[c] * intl) * whl;
        }
        if (lev == nv - 1) {
            dvt[0] = dvt[1] = dvt[2] = 0.0;
        } else {
            const double wht = wwh[j + 1];
            wts(altht, alt, Alt[lev + 1], intt, intl);
#pragma unroll
            for (int c = 0; c < 3; c++) dvt[c] = (vc[c] * intt + wv[j + 1][c] * intl) * wht;
        }
        double r2p, r2m, r2l;
        if (g.deep) {
            r2p = XPOW(altht + A, 2.0);
            r2m = XPOW(alt + A, 2.0);
            r2l = XPOW(althl + A, 2.0);
        } else {
            r2p = r2m = r2l = 1.0;
        }
        const double dzi = 1.0 / ((althl - altht) * r2m);
        double dav[3];
#pragma unroll
        for (int c = 0; c < 3; c++) dav[c] = (dvl[c] * r2l - dvt[c] * r2p) * dzi;
        const double rho = wr[j];
        double Cx, Cy, Cz;
        if (g.deep) {
            Cx = -2.0 * Omega * vc[1] * rho;
            Cy = 2.0 * Omega * vc[0] * rho;
            Cz = 0.0;
        } else {
            Cx = 2 * Omega * (vc[2] * fr[2] * fr[1] - vc[1] * XPOW(fr[2], 2)) * rho;
            Cy = -2 * Omega * (vc[2] * fr[2] * fr[0] - vc[0] * XPOW(fr[2], 2)) * rho;
            Cz = 2 * Omega * (vc[1] * fr[2] * fr[0] - vc[0] * fr[2] * fr[1]) * rho;
        }
        auto *Adv = FS<MX>::Adv(f);
        VEC(Adv, 0, lev, p) = dav[0] + Cx;
        VEC(Adv, 1, lev, p) = dav[1] + Cy;
        VEC(Adv, 2, lev, p) = dav[2] + Cz;
    }
}

// Total-energy form of Compute_Temperature_H_Pt_Geff (thor_auxiliary.h:143-161,
// 212-230, 281-287), run after the column preparation, whose h and hh it
// reads: potential energy ep = g z (g (A / (A + z))^2 z with GravHeightVar)
// and kinetic energy ek = |(Mh, W)|^2 / (2 rho^2) of the stage at the centres,
// at the interfaces the same interpolation and end extrapolations as the
// other interface values; the stage's total energy rho (ep + ek) + Cv p / Rd,
// and the flux variable of the density/pressure update, h + ep + ek at the
// centres and hh + eph + ekh at the interfaces (thor_fastmodes.h:352-372,
// 459-476), in place of pt and pth. Block (x, lev): 256 columns at centre
// lev (if < nv) and interface lev.
__global__ __launch_bounds__(256) void k_energy_prep(Geo g, Fields f) {
    const int p = blockIdx.x * 256 + threadIdx.x, lev = blockIdx.y;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const double Rd = g.Rd, Cp = g.Cp, Gravit = g.Gravit, A = g.A, Cv = Cp - Rd;
    const double *Alt = g.Alt, *Alth = g.Alth;
    auto ep_at = [&](double z) { return g.ghv ? Gravit * MPOW(A / (A + z), 2) * z : Gravit * z; };
    auto ek_at = [&](int l) {
        const double m0 = VEC(f.Mhk, 0, l, p), m1 = VEC(f.Mhk, 1, l, p), m2 = VEC(f.Mhk, 2, l, p);
        const double w = LV(f.Wk, l), r = LV(f.Rhok, l);
        return 0.5 * (m0 * m0 + m1 * m1 + m2 * m2 + w * w) / (r * r);
    };
    if (lev < nv) {
        const double ep = ep_at(Alt[lev]), ek = ek_at(lev);
        LV(f.Etot, lev) = LV(f.Rhok, lev) * (ep + ek) + Cv / Rd * LV(f.pk, lev);
        LV(f.pt, lev) = LV(f.h, lev) + ep + ek;
    }
    double eph, ekh;
    if (lev == 0) {
        eph = ep_at(Alth[0]);
        const double ek0 = ek_at(0), ek1 = ek_at(1);
        ekh = ek1 + (ek1 - ek0) * (Alth[0] - Alt[1]) / (Alt[1] - Alt[0]);
    } else if (lev == nv) {
        eph = ep_at(Alth[nv]);
        const double alt = Alt[nv - 1];
        const double extr = (2 * Alth[nv] - alt - Alt[nv - 2]) / (Alt[nv - 1] - Alt[nv - 2]);
        const double ekm = ek_at(nv - 2), ekc = ek_at(nv - 1);
        ekh = ekm + (ekc - ekm) * extr;
    } else {
        const double xi = Alth[lev], xim = Alt[lev - 1], xip = Alt[lev];
        const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
        eph = ep_at(Alt[lev - 1]) * intt + ep_at(Alt[lev]) * intl;
        ekh = ek_at(lev - 1) * intt + ek_at(lev) * intl;
    }
    LV(f.pth, lev) = LV(f.hh, lev) + eph + ekh;
}

// ===================================================================== K_B
// Horizontal advection (Compute_Advec_Cori1 / _Poles, thor_adv_cor.h:51-362)
// added to the vertical part, and the first Laplacian pass of the 4th-order
// hyperdiffusion (Diffusion_Op stepnum 0, thor_diff.h:54-522) of rho, u, v,
// w = M/rho, W/rho and Rd*T.
template <bool MX, class GetM, class GetV, class GetR, class GetW, class GetT>
__device__ __forceinline__ void adv_diff1_point(const Geo &g, const Fields &f, int p, int lev, int nf, int hydiff,
                                                GetM M, GetV v, GetR rho, GetW W, GetT T) {
    const size_t S = g.S;
    const int nv = g.nv;
    const double rs = rscale_at(g, lev);
    double nf3[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 3; kk++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double sum = g.div[(size_t)(0 * 3 + kk) * S + p] * v(0, c) * M(0, kk);
SLOT_UNROLL(THOR_SLOT_UNROLL)
            for (int s = 1; s < 7; s++) sum = sum + g.div[(size_t)(s * 3 + kk) * S + p] * v(s, c) * M(s, kk);
            nf3[c] += rs * sum;
        }
    }
    auto *Adv = FS<MX>::Adv(f);
#pragma unroll
    for (int c = 0; c < 3; c++) VEC(Adv, c, lev, p) = nf3[c] + VEC(Adv, c, lev, p);
    if (!hydiff) return;
    const double Rd = g.Rd;
    auto a = [&](int var, int s) -> double {
        if (var == 0) return rho(s);
        if (var <= 3) return M(s, var - 1) / rho(s);
        if (var == 4) return W(s) / rho(s);
        if (g.energy) return (g.Cp - Rd) * T(s);  // internal energy (thor_diff.h:196-202)
        return Rd * T(s);
    };
    auto *diff = FS<MX>::diff(f);
    if constexpr (MX) {
        float lap[6];
        laplacians_w<6>(g, p, (float)(rs * rs), [&](int var, int s) { return (float)a(var, s); }, lap);
#pragma unroll
        for (int var = 0; var < 6; var++) VEC(diff, var, lev, p) = lap[var];
    } else {
        double lap[6];
        laplacians<6>(g, p, nf, rs, a, lap);
#pragma unroll
        for (int var = 0; var < 6; var++) VEC(diff, var, lev, p) = lap[var];
    }
}

template <bool MX> __global__ STENCIL_BOUNDS void k_adv_diff1(Geo g, Fields f, int hydiff) {
    int r, lev;
    if (!block_tile(g, r, lev)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    auto vel = [&](int q, int c) {
        const double rho = LVP(f.Rhok, lev, q), w = LVP(f.Wk, lev, q);
        return VEC(f.Mhk, c, lev, q) / rho + (w / rho) * g.fr[(size_t)c * S + q];
    };
    if (r == g.nr) {
        PoleCtx c;
        if (!pole_ctx(g, c)) return;
        adv_diff1_point<MX>(
            g, f, c.p, lev, 5, hydiff, [&](int s, int k) { return VEC(f.Mhk, k, lev, c.gi(s)); },
            [&](int s, int k) { return vel(c.gi(s), k); }, [&](int s) { return LVP(f.Rhok, lev, c.gi(s)); },
            [&](int s) { return LVP(f.Wk, lev, c.gi(s)); }, [&](int s) { return LVP(f.T, lev, c.gi(s)); });
        return;
    }
    __shared__ double sM[3][NHL2], sV[3][NHL2], sR[NHL2], sW[NHL2], sT[NHL2];
    const TileThread T = tile_thread(g, r);
    stage_tile(g, r, T, [&](int pos, int q) {
        if (q < 0) {
            for (int k = 0; k < 3; k++) sM[k][pos] = sV[k][pos] = 0.0;
            sR[pos] = sW[pos] = sT[pos] = 0.0;
            return;
        }
        const double rho = LVP(f.Rhok, lev, q), w = LVP(f.Wk, lev, q);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double m = VEC(f.Mhk, k, lev, q);
            sM[k][pos] = m;
            sV[k][pos] = m / rho + (w / rho) * g.fr[(size_t)k * S + q];
        }
        sR[pos] = rho;
        sW[pos] = w;
        sT[pos] = LVP(f.T, lev, q);
    });
    __syncthreads();
    adv_diff1_point<MX>(
        g, f, T.p, lev, T.nf, hydiff, [&](int s, int k) { return sM[k][T.pos(s)]; },
        [&](int s, int k) { return sV[k][T.pos(s)]; }, [&](int s) { return sR[T.pos(s)]; },
        [&](int s) { return sW[T.pos(s)]; }, [&](int s) { return sT[T.pos(s)]; });
}

// ===================================================================== K_C
// Last hyperdiffusion pass (Diffusion_Op laststep) + Correct_Horizontal
// (thor_diff.h:882-915), and divergence-damping pass 0 on Mhk/Whk
// (DivM_Op laststep = false).
// diff2_tail: the last pass's output at one cell from its six Laplacians
// (the strength and sign, the diffusive sponge, Correct_Horizontal); rr is the
// cell's own density.
template <bool MX>
__device__ __forceinline__ void diff2_tail(const Geo &g, const Fields &f, int p, int lev, bool pole, double sign,
                                           double out[6], double rr) {
    const size_t S = g.S;
    const int nv = g.nv;
    const double sdiff = -g.Kdh4[lev] * sign;
#pragma unroll
    for (int var = 0; var < 6; var++) out[var] = pole ? sdiff * out[var] : out[var] * sdiff;
    if (const auto *dsp = FS<MX>::dsp(f)) {
        // diffusive sponge (thor_diff.h:453-509): (-1)^(order/2 + 1) rho Kdh2
        // times the sponge pass's input at the point, for u, v, w
        const double k2 = g.Kdh2[lev];
        for (int var = 1; var <= 4; var++)
            out[var] += g.ssp * rr * k2 * (double)VEC(dsp, var, lev, p) * g.dsp_scale;
    }
    LV(FS<MX>::diffrh(f), lev) = out[0];
    LV(FS<MX>::diffw(f), lev) = out[4];
    LV(FS<MX>::diffpr(f), lev) = out[5];
    const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
    const double dmhr = fx * out[1] + fy * out[2] + fz * out[3];
    auto *diffmh = FS<MX>::diffmh(f);
    VEC(diffmh, 0, lev, p) = out[1] + -fx * dmhr;
    VEC(diffmh, 1, lev, p) = out[2] + -fy * dmhr;
    VEC(diffmh, 2, lev, p) = out[3] + -fz * dmhr;
}

template <bool MX, class GetD, class GetS, class GetM, class GetWl, class GetWt>
__device__ __forceinline__ void diff2_divm0_point(const Geo &g, const Fields &f, int p, int lev, int nf, bool pole,
                                                  int hydiff, int divdamp, double sign, GetD dif, GetS rho, GetM M,
                                                  GetWl wl, GetWt wt) {
    const size_t S = g.S;
    const int nv = g.nv;
    if (hydiff) {
        const double rs = rscale_at(g, lev);
        double out[6];
        if constexpr (MX) {
            float o[6];
            laplacians_w<6>(g, p, (float)(rs * rs),
                            [&](int var, int s) { return var == 0 ? (float)dif(s, 0) : (float)(dif(s, var) * rho(s)); },
                            o);
#pragma unroll
            for (int var = 0; var < 6; var++) out[var] = o[var];
        } else {
            auto a = [&](int var, int s) -> double { return var == 0 ? dif(s, 0) : dif(s, var) * rho(s); };
            laplacians<6>(g, p, nf, rs, a, out);
        }
        diff2_tail<MX>(g, f, p, lev, pole, sign, out, rho(0));
    }
    if (divdamp) {
        auto *divg = FS<MX>::divg(f);
        if constexpr (MX) {
            float o3[3];
            divm_w<false>(g, p, lev, [&](int s, int k) { return (float)M(s, k); },
                          [&](int s) { return (float)(wl(s) - wt(s)); }, o3);
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(divg, k, lev, p) = o3[k];
        } else {
            double o3[3];
            divm_nf<false>(g, p, lev, nf, pole, M, wl, wt, o3);
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(divg, k, lev, p) = o3[k];
        }
    }
}

template <bool MX>
__global__ STENCIL_BOUNDS void k_diff2_divm0(Geo g, Fields f, int hydiff, int divdamp, double sign) {
    int r, lev;
    if (!block_tile(g, r, lev)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    if (r == g.nr) {
        PoleCtx c;
        if (!pole_ctx(g, c)) return;
        auto *diff = FS<MX>::diff(f);
        diff2_divm0_point<MX>(
            g, f, c.p, lev, 5, true, hydiff, divdamp, sign,
            [&](int s, int var) { return VEC(diff, var, lev, c.gi(s)); },
            [&](int s) { return LVP(f.Rhok, lev, c.gi(s)); },
            [&](int s, int k) { return VEC(f.Mhk, k, lev, c.gi(s)); },
            [&](int s) { return LVP(f.Whk, lev, c.gi(s)); }, [&](int s) { return LVP(f.Whk, lev + 1, c.gi(s)); });
        return;
    }
    using R = typename FS<MX>::R;
    __shared__ R sD[6][NHL2];
    __shared__ double sR[NHL2], sM[3][NHL2], sWl[NHL2], sWt[NHL2];
    const TileThread T = tile_thread(g, r);
    auto *diff = FS<MX>::diff(f);
    stage_tile(g, r, T, [&](int pos, int q) {
        if (q < 0) {
            for (int k = 0; k < 6; k++) sD[k][pos] = 0.0;
            for (int k = 0; k < 3; k++) sM[k][pos] = 0.0;
            sR[pos] = sWl[pos] = sWt[pos] = 0.0;
            return;
        }
        if (hydiff) {
#pragma unroll
            for (int k = 0; k < 6; k++) sD[k][pos] = VEC(diff, k, lev, q);
            sR[pos] = LVP(f.Rhok, lev, q);
        }
        if (divdamp) {
#pragma unroll
            for (int k = 0; k < 3; k++) sM[k][pos] = VEC(f.Mhk, k, lev, q);
            sWl[pos] = LVP(f.Whk, lev, q);
            sWt[pos] = LVP(f.Whk, lev + 1, q);
        }
    });
    __syncthreads();
    diff2_divm0_point<MX>(
        g, f, T.p, lev, T.nf, false, hydiff, divdamp, sign, [&](int s, int var) { return sD[var][T.pos(s)]; },
        [&](int s) { return sR[T.pos(s)]; }, [&](int s, int k) { return sM[k][T.pos(s)]; },
        [&](int s) { return sWl[T.pos(s)]; }, [&](int s) { return sWt[T.pos(s)]; });
}

// Vertical hyperdiffusion (vertical_diff, thor_diff.h:1198-1426, run at
// thor_driver.cu:590-612 when VertHyDiff) and its part of Correct_Horizontal
// (thor_diff.h:882-915), in FP64 arithmetic in both precision forms. Per
// column and variable (0 rho, 1-3 momentum / rho, 4 W / rho, 5 T): a vertical
// Laplacian in flux form with r^2 weights on the column extended by one ghost
// level at each end (a copy of the layer next to it), applied order / 2
// times; before the last application the field is multiplied by sign Kdv6
// (and by rho for variables 1-5), sign = (-1)^(order/2 + 1), and the T result
// is multiplied by Rd. The density result is unused (THOR's diffrv is 0).
// A block is blockDim.x columns x 6 variables; each thread holds its column
// in LDS and updates it in place, carrying the old value below in a register.
// Outputs: the momentum part with its radial component removed goes to
// diffmv (FP64 form) or is added to the FP32 horizontal momentum diffusion
// (mixed form); the w and R T parts are added to diffw and diffpr, which
// THOR only reads as diffw + diffwv and diffpr + diffprv
// (thor_slowmodes.h:348-380, thor_fastmodes.h:504-571).
__host__ __device__ inline int vert_diff_columns(int nv) {
    return (6 * nv + (nv + 2) * 6 * 16) * 8 <= 64 * 1024 ? 16 : 8;
}
__host__ __device__ inline size_t vert_diff_lds(int nv, int cb) { return (size_t)(6 * nv + (nv + 2) * 6 * cb) * 8; }

template <bool MX> __global__ __launch_bounds__(96) void k_vert_diff(Geo g, Fields f, int order) {
    extern __shared__ double vsh[];
    const int nv = g.nv, CB = blockDim.x, c = threadIdx.x, var = threadIdx.y;
    const size_t S = g.S;
    double *fac = vsh;           // [6][nv]: 1/dz, r_up^2, r_low^2, r^2, 1/dz_up, 1/dz_low
    double *col = vsh + 6 * nv;  // [nv + 2][6][CB]: ghost-extended columns
    const double *Alt = g.Alt, *Alth = g.Alth;
    for (int lev = var * CB + c; lev < nv; lev += 6 * CB) {
        double ri = 1.0, rup = 1.0, rlow = 1.0;
        if (g.deep) {
            ri = g.A + Alt[lev];
            rup = g.A + Alth[lev + 1];
            rlow = g.A + Alth[lev];
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
        fac[lev] = 1.0 / (Alth[lev + 1] - Alth[lev]);
        fac[nv + lev] = XPOW(rup, 2);
        fac[2 * nv + lev] = XPOW(rlow, 2);
        fac[3 * nv + lev] = XPOW(ri, 2);
        fac[4 * nv + lev] = dzup;
        fac[5 * nv + lev] = dzlow;
    }
    const int p = blockIdx.x * CB + c;
    const bool on = p < g.P;
    auto A = [&](int i) -> double & { return col[((size_t)i * 6 + var) * CB + c]; };
    if (on) {
        for (int lev = 0; lev < nv; lev++) {
            const double r = LV(f.Rhok, lev);
            double a;
            if (var == 0) a = r;
            else if (var <= 3) a = VEC(f.Mhk, var - 1, lev, p) / r;
            else if (var == 4) a = LV(f.Wk, lev) / r;
            else a = LV(f.T, lev);
            A(lev + 1) = a;
        }
        A(0) = A(1);
        A(nv + 1) = A(nv);
    }
    __syncthreads();
    if (on) {
        const double sign = XPOW(-1.0, order / 2 + 1);
        const int passes = order / 2 - 1;
        for (int l = 0; l <= passes; l++) {
            double am = A(0), a0 = A(1);
            for (int lev = 0; lev < nv; lev++) {
                const double ap = A(lev + 2);
                const double flux =
                    fac[nv + lev] * (ap - a0) * fac[4 * nv + lev] - fac[2 * nv + lev] * (a0 - am) * fac[5 * nv + lev];
                double d;
                if (l < passes) {
                    d = flux * fac[lev] / fac[3 * nv + lev];
                    if (l == passes - 1) d = var == 0 ? sign * g.Kdv6[lev] * d : sign * LV(f.Rhok, lev) * g.Kdv6[lev] * d;
                } else {
                    d = var != 5 ? flux * fac[lev] / fac[3 * nv + lev] : flux * fac[lev] * g.Rd / fac[3 * nv + lev];
                }
                A(lev + 1) = d;
                am = a0;
                a0 = ap;
            }
            A(0) = A(1);
            A(nv + 1) = A(nv);
        }
    }
    __syncthreads();
    if (!on || var == 0) return;
    if (var <= 3) {
        const int k = var - 1;
        const double fk = g.fr[(size_t)k * S + p], fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
        for (int lev = 0; lev < nv; lev++) {
            const double *q = &col[((size_t)(lev + 1) * 6) * CB + c];
            const double dk = q[var * CB];
            const double dmvr = fx * q[1 * CB] + fy * q[2 * CB] + fz * q[3 * CB];
            const double out = dk + -fk * dmvr;
            if constexpr (MX) VEC(f.mdiffmh, k, lev, p) = (float)((double)VEC(f.mdiffmh, k, lev, p) + out);
            else VEC(f.diffmv, k, lev, p) = out;
        }
    } else {
        auto *dst = var == 4 ? FS<MX>::diffw(f) : FS<MX>::diffpr(f);
        for (int lev = 0; lev < nv; lev++) LV(dst, lev) = (double)LV(dst, lev) + A(lev + 1);
    }
}

// Intermediate hyperdiffusion pass of HyDiffOrder >= 6 (Diffusion_Op with
// stepnum > 0 and laststep false, thor_driver.cu:435-470, argument
// thor_diff.h:218-221): the Laplacian of each of the six fields of the
// previous pass, unweighted and unscaled. THOR reads and writes diff_d in
// place in this pass, so a block's halo may already hold this pass's values;
// here each pass reads src and writes dst (the caller alternates two arrays),
// so every value is the Laplacian of the previous pass's output.
// Mixed form: the FP32 output is multiplied by scale (dbar^2), so repeated
// Laplacians (each a factor of about 1/dbar^2) stay far from FP32 underflow
// on large planets; the last pass divides the coefficient by the same factors.
template <bool MX>
__global__ STENCIL_BOUNDS void k_diff_mid(Geo g, const typename FS<MX>::R *src, typename FS<MX>::R *dst,
                                         float scale) {
    int r, lev;
    if (!block_tile(g, r, lev)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const double rs = rscale_at(g, lev);
    auto lap = [&](int p, int nf, auto a) {
        if constexpr (MX) {
            float o[6];
            laplacians_w<6>(g, p, (float)(rs * rs), a, o);
#pragma unroll
            for (int var = 0; var < 6; var++) VEC(dst, var, lev, p) = o[var] * scale;
        } else {
            double o[6];
            laplacians<6>(g, p, nf, rs, a, o);
#pragma unroll
            for (int var = 0; var < 6; var++) VEC(dst, var, lev, p) = o[var];
        }
    };
    if (r == g.nr) {
        PoleCtx c;
        if (!pole_ctx(g, c)) return;
        lap(c.p, 5, [&](int var, int s) { return VEC(src, var, lev, c.gi(s)); });
        return;
    }
    using R = typename FS<MX>::R;
    __shared__ R sD[6][NHL2];
    const TileThread T = tile_thread(g, r);
    stage_tile(g, r, T, [&](int pos, int q) {
#pragma unroll
        for (int k = 0; k < 6; k++) sD[k][pos] = q < 0 ? (R)0 : VEC(src, k, lev, q);
    });
    __syncthreads();
    lap(T.p, T.nf, [&](int var, int s) { return sD[var][T.pos(s)]; });
}

// ===================================================================== K_D
// Divergence-damping pass 1 on the stage values + Compute_Slow_Modes / _Poles
// (thor_slowmodes.h:48-683, entropy form) + UpdateRK (thor_auxiliary.h:311-418)
// for rk > 0, or the zero start of the deviations for rk = 0
// (thor_driver.cu:104-108).
// slow_tail: everything of one cell after the horizontal operator sums
// (DivM, the pressure gradient nflxv and the flux divergences nflxr, nflxp),
// given the cell's own pressure pv0 and momentum M0.
template <bool MX>
__device__ __forceinline__ void slow_tail(const Geo &g, const Fields &f, int p, int lev, bool pole, int rk,
                                          const double DivM[3], const double nflxv_in[3], double nflxr, double nflxp,
                                          double pv0, double M0x, double M0y, double M0z);
template <bool MX, class GetG, class GetP, class GetH, class GetM>
__device__ __forceinline__ void slow_point(const Geo &g, const Fields &f, int p, int lev, int nf, bool pole,
                                           int divdamp, int rk, GetG dg, GetP pv, GetH hv, GetM mv) {
    const size_t S = g.S;
    double DivM[3] = {0.0, 0.0, 0.0};
    if (divdamp) {
        if constexpr (MX) {
            float o3[3];
            divm_w<true>(g, p, lev, dg, [](int) { return 0.f; }, o3);
            for (int k = 0; k < 3; k++) DivM[k] = o3[k];
        } else {
            auto none = [&](int) { return 0.0; };
            divm_nf<true>(g, p, lev, nf, pole, dg, none, none, DivM);
        }
    }
    const double rscale = rscale_at(g, lev);
    double nflxv[3], nflxr = 0.0, nflxp = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double sg = g.grad[(size_t)(0 * 3 + k) * S + p] * pv(0);
        double sr = g.div[(size_t)(0 * 3 + k) * S + p] * mv(0, k);
        double sp = g.div[(size_t)(0 * 3 + k) * S + p] * mv(0, k) * hv(0);
SLOT_UNROLL(THOR_SLOT_UNROLL)
        for (int s = 1; s < 7; s++) {
            const double gd = g.grad[(size_t)(s * 3 + k) * S + p], dd = g.div[(size_t)(s * 3 + k) * S + p];
            sg = sg + gd * pv(s);
            sr = sr + dd * mv(s, k);
            sp = sp + dd * mv(s, k) * hv(s);
        }
        nflxv[k] = rscale * sg;
        nflxr += rscale * sr;
        nflxp += rscale * sp;
    }
    slow_tail<MX>(g, f, p, lev, pole, rk, DivM, nflxv, nflxr, nflxp, pv(0), mv(0, 0), mv(0, 1), mv(0, 2));
}

template <bool MX>
__device__ __forceinline__ void slow_tail(const Geo &g, const Fields &f, int p, int lev, bool pole, int rk,
                                          const double DivM[3], const double nflxv_in[3], double nflxr, double nflxp,
                                          double pv0, double M0x, double M0y, double M0z) {
    const size_t S = g.S;
    const int nv = g.nv;
    const double A = g.A, Gravit = g.Gravit, Rd = g.Rd, Cv = g.Cp - g.Rd;
    const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
    const double *Alt = g.Alt, *Alth = g.Alth;
    const double nflxv[3] = {nflxv_in[0], nflxv_in[1], nflxv_in[2]};
    const auto *Adv = FS<MX>::Adv(f);
    double advx = VEC(Adv, 0, lev, p), advy = VEC(Adv, 1, lev, p), advz = VEC(Adv, 2, lev, p);
    const double rho = LV(f.Rhok, lev);
    double advrl = 0, rhol = 0, altl = 0, pressurel = 0;
    if (lev != 0) {
        const double ax = VEC(Adv, 0, lev - 1, p), ay = VEC(Adv, 1, lev - 1, p), az = VEC(Adv, 2, lev - 1, p);
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
    if (!g.nonhydro) {
        advr = (XPOW(M0x, 2) + XPOW(M0y, 2) + XPOW(M0z, 2)) / (rho * (alt + A));
        if (lev != 0) {
            const double m1x = VEC(f.Mhk, 0, lev - 1, p), m1y = VEC(f.Mhk, 1, lev - 1, p),
                         m1z = VEC(f.Mhk, 2, lev - 1, p);
            advrl = (XPOW(m1x, 2) + XPOW(m1y, 2) + XPOW(m1z, 2)) / (rhol * (altl + A));
        }
    }
    const double dz = altht - althl;
    double r2p, r2m, r2l, dr2dz;
    if (g.deep) {
        r2p = XPOW(altht + A, 2.0);
        r2m = XPOW(alt + A, 2.0);
        r2l = XPOW(althl + A, 2.0);
        dr2dz = (XPOW(altht + A, 3.0) - XPOW(althl + A, 3.0)) / 3;
    } else {
        r2p = r2m = r2l = 1.0;
        dr2dz = r2m * dz;
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
    auto *SlowMh = FS<MX>::SlowMh(f);
    const auto *diffmh = FS<MX>::diffmh(f);
#pragma unroll
    for (int k = 0; k < 3; k++)
        VEC(SlowMh, k, lev, p) = -(nflxv[k] - dpr * fv[k]) - ad[k] + DivM[k] + VEC(diffmh, k, lev, p)
                                 + (f.diffmv ? VEC(f.diffmv, k, lev, p) : 0.0) + VEC(f.dMh, k, lev, p);
    if (lev == 0) {
        LV(f.SlowWh, 0) = 0.0;
        LV(f.SlowWh, nv) = 0.0;
    } else {
        const double xi = althl, xim = altl, xip = alt;
        const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
        double swr;
        if (g.deep) {
            if (g.ghv)
                swr = (advrl + rhol * Gravit * XPOW(A / (A + altl), 2)) * intt
                      + (advr + rho * Gravit * XPOW(A / (A + alt), 2)) * intl;
            else
                swr = (advrl + rhol * Gravit) * intt + (advr + rho * Gravit) * intl;
        } else {
            if (g.ghv)
                swr = (rhol * Gravit * XPOW(A / (A + altl), 2)) * intt + (rho * Gravit * XPOW(A / (A + alt), 2)) * intl;
            else
                swr = (rhol * Gravit) * intt + (rho * Gravit) * intl;
        }
        const auto *diffw = FS<MX>::diffw(f);
        if (g.nonhydro) swr += (-LV(diffw, lev - 1) - 0.0) * intt + (-LV(diffw, lev) - 0.0) * intl;
        const double dzl = alt - altl;
        const double dpdz = (pv0 - pressurel) / dzl;
        swr += dpdz;
        swr -= LV(f.dWh, lev);
        LV(f.SlowWh, lev) = -swr;
    }
    nflxr += dwdz;
    LV(f.SlowRho, lev) = -nflxr + LV(FS<MX>::diffrh(f), lev) + 0.0;
    const double r = 1.0 / rho;
    const double vgp =
        (nflxv[0] - dpr * fx) * M0x * r + (nflxv[1] - dpr * fy) * M0y * r + (nflxv[2] - dpr * fz) * M0z * r;
    double sp = (Rd / Cv) * (-nflxp - LV(f.gtil, lev) - dwhdz + vgp + LV(f.Qheat, lev));
    if (g.energy) sp += Rd / Cv * (LV(FS<MX>::diffpr(f), lev) + 0.0);  // thor_slowmodes.h:375-377
    else sp += LV(FS<MX>::diffpr(f), lev) + 0.0;
    LV(f.Slowp, lev) = sp;

    // Deviations at the start of the stage.
    auto *Mhs = FS<MX>::Mhs(f);
    if (rk == 0) {
#pragma unroll
        for (int k = 0; k < 3; k++) VEC(Mhs, k, lev, p) = 0.0;
        LV(f.Rhos, lev) = 0.0;
        LV(f.ps, lev) = 0.0;
        LV(f.Whs, lev) = 0.0;
        if (lev == nv - 1) LV(f.Whs, nv) = 0.0;
    } else {
        const double Mx = VEC(f.Mh, 0, lev, p) - M0x, My = VEC(f.Mh, 1, lev, p) - M0y, Mz = VEC(f.Mh, 2, lev, p) - M0z;
        const double aux = Mx * fx + My * fy + Mz * fz;
        VEC(Mhs, 0, lev, p) = Mx - aux * fx;
        VEC(Mhs, 1, lev, p) = My - aux * fy;
        VEC(Mhs, 2, lev, p) = Mz - aux * fz;
        LV(f.Rhos, lev) = LV(f.Rho, lev) - rho;
        LV(f.ps, lev) = LV(f.p, lev) - pv0;
        LV(f.Whs, lev) = lev == 0 ? 0.0 : LV(f.Wh, lev) - LV(f.Whk, lev);
        if (lev == nv - 1) LV(f.Whs, nv) = 0.0;
    }
}

template <bool MX> __global__ STENCIL_BOUNDS void k_slow(Geo g, Fields f, int divdamp, int rk) {
    int r, lev;
    if (!block_tile(g, r, lev)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    if (r == g.nr) {
        PoleCtx c;
        if (!pole_ctx(g, c)) return;
        const auto *divg = FS<MX>::divg(f);
        slow_point<MX>(
            g, f, c.p, lev, 5, true, divdamp, rk, [&](int s, int k) { return VEC(divg, k, lev, c.gi(s)); },
            [&](int s) { return LVP(f.pk, lev, c.gi(s)); }, [&](int s) { return LVP(f.h, lev, c.gi(s)); },
            [&](int s, int k) { return VEC(f.Mhk, k, lev, c.gi(s)); });
        return;
    }
    using R = typename FS<MX>::R;
    __shared__ R sG[3][NHL2];
    __shared__ double sP[NHL2], sH[NHL2], sM[3][NHL2];
    const TileThread T = tile_thread(g, r);
    const auto *divg = FS<MX>::divg(f);
    stage_tile(g, r, T, [&](int pos, int q) {
        if (q < 0) {
            for (int k = 0; k < 3; k++) sG[k][pos] = sM[k][pos] = 0.0;
            sP[pos] = sH[pos] = 0.0;
            return;
        }
        if (divdamp) {
#pragma unroll
            for (int k = 0; k < 3; k++) sG[k][pos] = VEC(divg, k, lev, q);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) sM[k][pos] = VEC(f.Mhk, k, lev, q);
        sP[pos] = LVP(f.pk, lev, q);
        sH[pos] = LVP(f.h, lev, q);
    });
    __syncthreads();
    slow_point<MX>(
        g, f, T.p, lev, T.nf, false, divdamp, rk, [&](int s, int k) { return sG[k][T.pos(s)]; },
        [&](int s) { return sP[T.pos(s)]; }, [&](int s) { return sH[T.pos(s)]; },
        [&](int s, int k) { return sM[k][T.pos(s)]; });
}

// ===================================================================== K_E
// Divergence-damping pass 0 on the deviations Mhs / Whs.
template <bool MX> __global__ STENCIL_BOUNDS void k_divm0(Geo g, Fields f) {
    int r, lev;
    if (!block_tile(g, r, lev)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    using R = typename FS<MX>::R;
    const R *Mhs = FS<MX>::Mhs(f);
    R *divg = FS<MX>::divg(f);
    // pass 0 of one point from slot getters, written to divg
    auto run = [&](int p, int nf, bool pole, auto a, auto wl, auto wt) {
        if constexpr (MX) {
            float o3[3];
            divm_w<false>(g, p, lev, a, [&](int s) { return (float)(wl(s) - wt(s)); }, o3);
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(divg, k, lev, p) = o3[k];
        } else {
            double o3[3];
            divm_nf<false>(g, p, lev, nf, pole, a, wl, wt, o3);
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(divg, k, lev, p) = o3[k];
        }
    };
    if (r == g.nr) {
        PoleCtx c;
        if (!pole_ctx(g, c)) return;
        run(c.p, 5, true, [&](int s, int k) { return VEC(Mhs, k, lev, c.gi(s)); },
            [&](int s) { return LVP(f.Whs, lev, c.gi(s)); }, [&](int s) { return LVP(f.Whs, lev + 1, c.gi(s)); });
        return;
    }
    __shared__ R sM[3][NHL2];
    __shared__ double sWl[NHL2], sWt[NHL2];
    const TileThread T = tile_thread(g, r);
    stage_tile(g, r, T, [&](int pos, int q) {
        if (q < 0) {
            for (int k = 0; k < 3; k++) sM[k][pos] = 0.0;
            sWl[pos] = sWt[pos] = 0.0;
            return;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) sM[k][pos] = VEC(Mhs, k, lev, q);
        sWl[pos] = LVP(f.Whs, lev, q);
        sWt[pos] = LVP(f.Whs, lev + 1, q);
    });
    __syncthreads();
    run(T.p, T.nf, false, [&](int s, int k) { return sM[k][T.pos(s)]; }, [&](int s) { return sWl[T.pos(s)]; },
        [&](int s) { return sWt[T.pos(s)]; });
}

// ===================================================================== K_F
// Divergence-damping pass 1 on the deviations + Momentum_Eq / _Poles
// (thor_fastmodes.h:54-253). momentum_tail: the update of one cell from its
// pressure-gradient sums nf3 and DivM.
template <bool MX>
__device__ __forceinline__ void momentum_tail(const Geo &g, const Fields &f, int p, int lev, const double nf3[3],
                                              const double DivM[3], double dt);
template <bool MX, class GetG, class GetS>
__device__ __forceinline__ void momentum_point(const Geo &g, const Fields &f, int p, int lev, int nf, bool pole,
                                               int divdamp, double dt, GetG dg, GetS pv) {
    const size_t S = g.S;
    double DivM[3] = {0.0, 0.0, 0.0};
    if (divdamp) {
        if constexpr (MX) {
            float o3[3];
            divm_w<true>(g, p, lev, dg, [](int) { return 0.f; }, o3);
            for (int k = 0; k < 3; k++) DivM[k] = o3[k];
        } else {
            auto none = [&](int) { return 0.0; };
            divm_nf<true>(g, p, lev, nf, pole, dg, none, none, DivM);
        }
    }
    const double rs = rscale_at(g, lev);
    double nf3[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double s0 = g.grad[(size_t)(0 * 3 + k) * S + p] * pv(0);
SLOT_UNROLL(THOR_SLOT_UNROLL)
        for (int s = 1; s < 7; s++) s0 = s0 + g.grad[(size_t)(s * 3 + k) * S + p] * pv(s);
        nf3[k] = rs * s0;
    }
    momentum_tail<MX>(g, f, p, lev, nf3, DivM, dt);
}

template <bool MX>
__device__ __forceinline__ void momentum_tail(const Geo &g, const Fields &f, int p, int lev, const double nf3[3],
                                              const double DivM[3], double dt) {
    const size_t S = g.S;
    const int nv = g.nv;
    const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
    const auto *SlowMh = FS<MX>::SlowMh(f);
    double Mx = (-nf3[0] + VEC(SlowMh, 0, lev, p) + DivM[0]) * dt;
    double My = (-nf3[1] + VEC(SlowMh, 1, lev, p) + DivM[1]) * dt;
    double Mz = (-nf3[2] + VEC(SlowMh, 2, lev, p) + DivM[2]) * dt;
    const double vr = Mx * fx + My * fy + Mz * fz;
    Mx += -vr * fx;
    My += -vr * fy;
    Mz += -vr * fz;
    auto *Mhs = FS<MX>::Mhs(f);
    VEC(Mhs, 0, lev, p) += Mx;
    VEC(Mhs, 1, lev, p) += My;
    VEC(Mhs, 2, lev, p) += Mz;
}

template <bool MX> __global__ STENCIL_BOUNDS void k_momentum(Geo g, Fields f, int divdamp, double dt) {
    int r, lev;
    if (!block_tile(g, r, lev)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    if (r == g.nr) {
        PoleCtx c;
        if (!pole_ctx(g, c)) return;
        const auto *divg = FS<MX>::divg(f);
        momentum_point<MX>(
            g, f, c.p, lev, 5, true, divdamp, dt, [&](int s, int k) { return VEC(divg, k, lev, c.gi(s)); },
            [&](int s) { return LVP(f.ps, lev, c.gi(s)); });
        return;
    }
    using R = typename FS<MX>::R;
    __shared__ R sG[3][NHL2];
    __shared__ double sP[NHL2];
    const TileThread T = tile_thread(g, r);
    const auto *divg = FS<MX>::divg(f);
    stage_tile(g, r, T, [&](int pos, int q) {
        if (q < 0) {
            for (int k = 0; k < 3; k++) sG[k][pos] = 0.0;
            sP[pos] = 0.0;
            return;
        }
        if (divdamp) {
#pragma unroll
            for (int k = 0; k < 3; k++) sG[k][pos] = VEC(divg, k, lev, q);
        }
        sP[pos] = LVP(f.ps, lev, q);
    });
    __syncthreads();
    momentum_point<MX>(
        g, f, T.p, lev, T.nf, false, divdamp, dt, [&](int s, int k) { return sG[k][T.pos(s)]; },
        [&](int s) { return sP[T.pos(s)]; });
}

// ===================================================================== K_G1
// Per column: Prepare_Implicit_Vertical (thor_vertical_int.h:566-771) for the
// source terms Sp, Sd and Vertical_Eq (thor_vertical_int.h:52-563): the
// tridiagonal solve for Whs at interfaces 1..nv-1 with the modify-in-place
// Thomas algorithm; cc/dd live in LDS.
// Columns per Thomas block: 4 LDS planes of (nv + 1) x CB doubles must fit 64 KiB.
__host__ __device__ constexpr int kThomasCB(int nv) { return 4 * (nv + 1) * 32 * 8 <= 65536 ? 32 : 16; }

// Prepare_Implicit_Vertical + _Poles (thor_vertical_int.h:566-771): source
// terms Sp = -(Rd/Cv) div(Mhs h) + Slowpressure, Sd = -div(Mhs) + SlowRho.
// ... and the horizontal flux divergences of this substep's
// Density_Pressure_Eqs (thor_fastmodes.h:255-847), which see the same Mhs:
// hflx_r = rscale sum div Mhs (the Sd sum, accumulated with the opposite sign)
// and hflx_pt = rscale sum div (Mhs + Mhk) pt.
// wv(s, k) is the divergence coefficient of slot s, component k; U is the
// unroll count of the slot loop (7 when wv reads a register array).
// SPLIT (mixed path): v1 is the momentum deviation alone and the stage part
// dptk = rs sum_k div_k(Mhk_k pt) of the pressure flux divergence, formed once
// per stage, is added at the end.
template <int U, bool SPLIT = false, class GetW, class GetM, class GetH, class GetV1, class GetPt>
__device__ __forceinline__ void prepare_point_w(const Geo &g, const Fields &f, int p, int lev, GetW wv, GetM mv,
                                                GetH hv, GetV1 v1, GetPt ptv, double slowp, double slowrho,
                                                double rs, double dptk = 0.0) {
    const size_t S = g.S;
    const double RoC = g.Rd / (g.Cp - g.Rd);
    double nflxr = 0.0, nflxp = 0.0, dr = 0.0, dpt = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double d0 = wv(0, k);
        double sr = d0 * mv(0, k);
        double sp = d0 * mv(0, k) * hv(0);
        double st = d0 * v1(0, k) * ptv(0);
#pragma unroll U
        for (int s = 1; s < 7; s++) {
            const double dd = wv(s, k);
            sr = sr + dd * mv(s, k);
            sp = sp + dd * mv(s, k) * hv(s);
            st = st + dd * v1(s, k) * ptv(s);
        }
        nflxr -= rs * sr;
        nflxp -= (RoC)*rs * sp;
        dr += rs * sr;
        dpt += rs * st;
    }
    LV(f.Sp, lev) = nflxp + slowp;
    LV(f.Sd, lev) = nflxr + slowrho;
    LV(f.hflx_r, lev) = dr;
    LV(f.hflx_pt, lev) = SPLIT ? dpt + dptk : dpt;
}

template <class GetM, class GetH, class GetV1, class GetPt>
__device__ __forceinline__ void prepare_point(const Geo &g, const Fields &f, int p, int lev, GetM mv, GetH hv,
                                              GetV1 v1, GetPt ptv) {
    const size_t S = g.S;
    prepare_point_w<THOR_SLOT_UNROLL>(
        g, f, p, lev, [&](int s, int k) { return g.div[(size_t)(s * 3 + k) * S + p]; }, mv, hv, v1, ptv,
        LV(f.Slowp, lev), LV(f.SlowRho, lev), rscale_at(g, lev));
}

template <bool MX> __global__ STENCIL_BOUNDS void k_prepare(Geo g, Fields f) {
    int r, lev;
    if (!block_tile(g, r, lev)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    if (r == g.nr) {
        PoleCtx c;
        if (!pole_ctx(g, c)) return;
        const auto *Mhs = FS<MX>::Mhs(f);
        prepare_point(
            g, f, c.p, lev, [&](int s, int k) { return (double)VEC(Mhs, k, lev, c.gi(s)); },
            [&](int s) { return LVP(f.h, lev, c.gi(s)); },
            [&](int s, int k) { return (double)VEC(Mhs, k, lev, c.gi(s)) + VEC(f.Mhk, k, lev, c.gi(s)); },
            [&](int s) { return LVP(f.pt, lev, c.gi(s)); });
        return;
    }
    using R = typename FS<MX>::R;
    __shared__ R sM[3][NHL2];
    __shared__ double sH[NHL2], sV1[3][NHL2], sPt[NHL2];
    const TileThread T = tile_thread(g, r);
    const auto *Mhs = FS<MX>::Mhs(f);
    stage_tile(g, r, T, [&](int pos, int q) {
        if (q < 0) {
            for (int k = 0; k < 3; k++) sM[k][pos] = sV1[k][pos] = 0.0;
            sH[pos] = sPt[pos] = 0.0;
            return;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const R m = VEC(Mhs, k, lev, q);
            sM[k][pos] = m;
            sV1[k][pos] = (double)m + VEC(f.Mhk, k, lev, q);
        }
        sH[pos] = LVP(f.h, lev, q);
        sPt[pos] = LVP(f.pt, lev, q);
    });
    __syncthreads();
    prepare_point(g, f, T.p, lev, [&](int s, int k) { return (double)sM[k][T.pos(s)]; },
                  [&](int s) { return sH[T.pos(s)]; }, [&](int s, int k) { return sV1[k][T.pos(s)]; },
                  [&](int s) { return sPt[T.pos(s)]; });
}

// ============================================== mixed precision, level chunks
// The substep stencils of the mixed-precision path (momentum, prepare and the
// divergence-damping pass 0) run one block per (rhombus tile, chunk of lc
// consecutive levels). Their per-point operator weights do not depend on the
// level, so each thread loads them once into registers and reuses them at
// every level of its chunk. The tile of level lev + 1 and the thread's own
// inputs for lev + 1 are loaded while level lev is computed and then written
// to the second of two LDS buffers, so each level costs one barrier and the
// loads of the next level overlap the arithmetic of this one. Block i runs on
// XCD i % 8 and each XCD takes a contiguous range of tiles. The pole block
// walks the same levels with direct gathers.
__device__ __forceinline__ bool block_tile_lc(const Geo &g, int lc, int &r, int &l0, int &l1) {
    const int i = blockIdx.x, xcd = i & 7, idx = i >> 3;
    const int nch = (g.nv + lc - 1) / lc;
    r = xcd * g.tiles_per_xcd + idx / nch;
    l0 = (idx % nch) * lc;
    l1 = min(g.nv, l0 + lc);
    return r <= g.nr;
}

// Buffer loads: one 128-bit resource descriptor per array (scalar registers),
// a 32-bit per-lane byte offset (the point) and a wave-uniform byte offset
// (level and component). A level-chunked kernel then keeps one 32-bit offset
// per staged point instead of a 64-bit address per array. A load outside
// [0, bytes) returns 0 instead of faulting.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0,
                                             (int)(bytes < 0x7fffffffu ? bytes : 0x7fffffffu), 0x00020000);
}
__device__ __forceinline__ float ldb_f(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ double ldb_d(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void stb_f(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void stb_d(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, double v) {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)voff, (int)soff, 0);
}

// Pole block of a level-chunked kernel: thread t takes pole P - 2 + (t & 1) at
// level l0 + t / 2 (two poles x lc levels <= 256 threads).
__device__ __forceinline__ bool pole_ctx_lc(const Geo &g, int l0, int l1, PoleCtx &c, int &lev) {
    const int t = threadIdx.x;
    lev = l0 + (t >> 1);
    if (lev >= l1) return false;
    c.p = g.P - 2 + (t & 1);
    c.ring = g.ring;
    c.S = g.S;
    return true;
}

// Thread of a level-chunked tile block: its point p, the LDS positions of its
// 7 stencil slots and its own and halo staging positions, and 32-bit byte
// offsets of the point and of its halo point (threads 0..65). A thread with
// no halo point, or whose halo point is a missing pentagon corner, gets an
// offset past the end of every buffer, so its halo loads return 0 without a
// branch; a missing corner is never read (a pentagon's slot 5 reads slot 6).
struct LcThread {
    int p, ir, hp;  // hp = -1: no halo point to stage
    int pos[7];
    unsigned o4, o8, h4, h8;
};
__device__ __forceinline__ LcThread lc_thread(const Geo &g, int r) {
    const TileThread T = tile_thread(g, r);
    LcThread L;
    L.p = T.p;
    L.ir = T.ir;
#pragma unroll
    for (int s = 0; s < 7; s++) L.pos[s] = T.pos(s);
    const int t = threadIdx.x;
    const int q = t < NHALO ? g.halo[r * NHALO + t] : -1;
    L.hp = t < NHALO ? halo_pos(t) : -1;
    L.o4 = 4u * (unsigned)T.p;
    L.o8 = 8u * (unsigned)T.p;
    L.h4 = q >= 0 ? 4u * (unsigned)q : 0x80000000u;
    L.h8 = q >= 0 ? 8u * (unsigned)q : 0x80000000u;
    return L;
}

// Per-level constants of a chunk (Geo::lvc rows for levels l0..l1-1) staged in
// LDS before the level loop: read as LDS values inside the loop, they do not
// join the in-order vector-memory counter that the pipelined loads rely on.
constexpr int kLcMax = 64;  // largest level chunk
__device__ __forceinline__ void stage_lvc(const Geo &g, int l0, int l1, float (*sL)[kLcMax]) {
    const int t = threadIdx.x, n = l1 - l0;
    if (t < 4 * n) sL[t / n][t % n] = g.lvc[(t / n) * g.nv + l0 + t % n];
}


// Momentum_Eq (thor_fastmodes.h:54-253) after the divergence-damping pass 1,
// in FP32: gradient of the FP64 pressure deviation with FP32 weights, DivM
// from the probed FP32 weights, FP32 momentum deviation.
__global__ __launch_bounds__(256, 3) void k_momentum_lc(Geo g, Fields f, int divdamp, double dt, int lc) {
#pragma clang fp contract(fast)
    int r, l0, l1;
    if (!block_tile_lc(g, lc, r, l0, l1)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    float *Mhs = f.mMhs;
    const float *divg = f.mdivg, *SlowMh = f.mSlowMh;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (pole_ctx_lc(g, l0, l1, c, lev))
            momentum_point<true>(
                g, f, c.p, lev, 5, true, divdamp, dt, [&](int s, int k) { return VEC(divg, k, lev, c.gi(s)); },
                [&](int s) { return LVP(f.ps, lev, c.gi(s)); });
        return;
    }
    __shared__ float sG[2][3][NHL2], sP[2][NHL2], sL[4][kLcMax];
    stage_lvc(g, l0, l1, sL);
    const LcThread T = lc_thread(g, r);
    const int p = T.p;
    float wg[21], wd[63];
#pragma unroll
    for (int i = 0; i < 21; i++) wg[i] = (float)g.grad[(size_t)i * S + p];
#pragma unroll
    for (int i = 0; i < 63; i++) wd[i] = divdamp ? g.wdd[(size_t)i * S + p] : 0.f;
    const float fx = (float)g.fr[p], fy = (float)g.fr[S + p], fz = (float)g.fr[2 * S + p];
    const float dtf = (float)dt;
    const size_t pl4 = S * 4, pl8 = S * 8;  // bytes of one FP32 / FP64 level plane
    const auto rG = rsrc(divg, 3 * nv * pl4), rP = rsrc(f.ps, nv * pl8), rS = rsrc(SlowMh, 3 * nv * pl4),
               rM = rsrc(Mhs, 3 * nv * pl4);
    struct In {
        float sm[3], mh[3], og[3], op, hg[3], hpv;
    };
    auto load = [&](In &v, int lev) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v.sm[k] = ldb_f(rS, T.o4, (unsigned)((k * nv + lev) * pl4));
            v.mh[k] = ldb_f(rM, T.o4, (unsigned)((k * nv + lev) * pl4));
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v.og[k] = ldb_f(rG, T.o4, (unsigned)((k * nv + lev) * pl4));
            v.hg[k] = ldb_f(rG, T.h4, (unsigned)((k * nv + lev) * pl4));
        }
        v.op = (float)ldb_d(rP, T.o8, (unsigned)(lev * pl8));
        v.hpv = (float)ldb_d(rP, T.h8, (unsigned)(lev * pl8));
    };
    auto put = [&](const In &v, int b) {
#pragma unroll
        for (int k = 0; k < 3; k++) sG[b][k][T.ir] = v.og[k];
        sP[b][T.ir] = v.op;
        if (T.hp >= 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) sG[b][k][T.hp] = v.hg[k];
            sP[b][T.hp] = v.hpv;
        }
    };
    float sm[3], mh[3];
    auto body = [&](int lev, In &R, In &L) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            sm[k] = L.sm[k];
            mh[k] = L.mh[k];
        }
        load(L, min(lev + 2, l1 - 1));
        const int b = (lev - l0) & 1;
        const float rs = sL[0][lev - l0], rs2 = sL[1][lev - l0];
        float DivM[3] = {0.f, 0.f, 0.f};
        if (divdamp) {
            float a[7][3];
#pragma unroll
            for (int s = 0; s < 7; s++)
#pragma unroll
                for (int k = 0; k < 3; k++) a[s][k] = sG[b][k][T.pos[s]];
            float o[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float sum = wd[c * 21] * a[0][0];
#pragma unroll
                for (int i = 1; i < 21; i++) sum += wd[c * 21 + i] * a[i / 3][i % 3];
                o[c] = rs2 * sum;
            }
            const float sdiff = sL[3][lev - l0];
            const float lapr = fx * o[0] + fy * o[1] + fz * o[2];
            DivM[0] = -(o[0] - fx * lapr) * sdiff;
            DivM[1] = -(o[1] - fy * lapr) * sdiff;
            DivM[2] = -(o[2] - fz * lapr) * sdiff;
        }
        float pv[7], M[3];
#pragma unroll
        for (int s = 0; s < 7; s++) pv[s] = sP[b][T.pos[s]];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float s0 = wg[k] * pv[0];
#pragma unroll
            for (int s = 1; s < 7; s++) s0 += wg[s * 3 + k] * pv[s];
            M[k] = (-rs * s0 + sm[k] + DivM[k]) * dtf;
        }
        const float vr = M[0] * fx + M[1] * fy + M[2] * fz;
        stb_f(rM, T.o4, (unsigned)((0 * nv + lev) * pl4), mh[0] + (M[0] - vr * fx));
        stb_f(rM, T.o4, (unsigned)((1 * nv + lev) * pl4), mh[1] + (M[1] - vr * fy));
        stb_f(rM, T.o4, (unsigned)((2 * nv + lev) * pl4), mh[2] + (M[2] - vr * fz));
        if (lev + 1 < l1) put(R, b ^ 1);
        __syncthreads();
    };
    In A, B;
    load(A, l0);
    load(B, min(l0 + 1, l1 - 1));
    put(A, 0);
    __syncthreads();
    for (int lev = l0; lev < l1; lev += 2) {
        body(lev, B, A);
        if (lev + 1 < l1) body(lev + 1, A, B);
    }
}

// Prepare_Implicit_Vertical (thor_vertical_int.h:566-771) and the horizontal
// flux divergences of Density_Pressure_Eqs in FP64 arithmetic, as in
// prepare_point (the momentum deviation is read from its FP32 storage).
// SPLIT: the stage part of the pressure flux divergence comes from
// f.hflx_ptk (formed by k_slow_lc), so the stage momentum Mhk is not staged.
// blocks per CU of the split mixed prepare (4 spills: 33.7 -> 38.8 us at g6 L48)
#ifndef THOR_PREP_LC_MINB
#define THOR_PREP_LC_MINB 3
#endif
template <bool SPLIT>
__global__ __launch_bounds__(256, SPLIT ? THOR_PREP_LC_MINB : 2) void k_prepare_lc(Geo g, Fields f, int lc) {
#pragma clang fp contract(fast)
    int r, l0, l1;
    if (!block_tile_lc(g, lc, r, l0, l1)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const float *Mhs = f.mMhs;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (pole_ctx_lc(g, l0, l1, c, lev))
            prepare_point(
                g, f, c.p, lev, [&](int s, int k) { return (double)VEC(Mhs, k, lev, c.gi(s)); },
                [&](int s) { return LVP(f.h, lev, c.gi(s)); },
                [&](int s, int k) { return (double)VEC(Mhs, k, lev, c.gi(s)) + VEC(f.Mhk, k, lev, c.gi(s)); },
                [&](int s) { return LVP(f.pt, lev, c.gi(s)); });
        return;
    }
    // staged point: (Mhs, 0) as one 16-byte LDS word and (h, pt) as another
    __shared__ float4 sM4[2][NHL2];
    using HT = std::conditional_t<SPLIT, float2, double2>;
    __shared__ HT sHT[2][NHL2];
    __shared__ double sV1[2][SPLIT ? 1 : 3][NHL2], sRs[kLcMax];
    if ((int)threadIdx.x < l1 - l0) sRs[threadIdx.x] = rscale_at(g, l0 + threadIdx.x);
    const LcThread T = lc_thread(g, r);
    const int p = T.p;
    double wv[21];
#pragma unroll
    for (int i = 0; i < 21; i++) wv[i] = g.div[(size_t)i * S + p];
    const size_t pl4 = S * 4, pl8 = S * 8;  // bytes of one FP32 / FP64 level plane
    const auto rM = rsrc(Mhs, 3 * nv * pl4), rMk = rsrc(f.Mhk, 3 * nv * pl8), rH = rsrc(f.h, nv * pl8),
               rPt = rsrc(f.pt, nv * pl8), rSp = rsrc(f.Slowp, nv * pl8), rSr = rsrc(f.SlowRho, nv * pl8),
               rPk = rsrc(f.hflx_ptk, nv * pl8), rHP = rsrc(f.mhpt, 2 * nv * pl4);
    struct In {
        double slowp, slowrho;
        float om[3], hm[3];
        double omk[SPLIT ? 1 : 3], hmk[SPLIT ? 1 : 3];
        std::conditional_t<SPLIT, float, double> oh, hh, opt, hpt;
    };
    auto load = [&](In &v, int lev) {
        v.slowp = ldb_d(rSp, T.o8, (unsigned)(lev * pl8));
        v.slowrho = ldb_d(rSr, T.o8, (unsigned)(lev * pl8));
        if constexpr (SPLIT) v.omk[0] = ldb_d(rPk, T.o8, (unsigned)(lev * pl8));  // the stage part of hflx_pt
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v.om[k] = ldb_f(rM, T.o4, (unsigned)((k * nv + lev) * pl4));
            v.hm[k] = ldb_f(rM, T.h4, (unsigned)((k * nv + lev) * pl4));
            if constexpr (!SPLIT) {
                v.omk[k] = ldb_d(rMk, T.o8, (unsigned)((k * nv + lev) * pl8));
                v.hmk[k] = ldb_d(rMk, T.h8, (unsigned)((k * nv + lev) * pl8));
            }
        }
        if constexpr (SPLIT) {  // h and pt of the stage in FP32 (they multiply the FP32 deviation)
            v.oh = ldb_f(rHP, T.o4, (unsigned)(lev * pl4));
            v.opt = ldb_f(rHP, T.o4, (unsigned)((nv + lev) * pl4));
            v.hh = ldb_f(rHP, T.h4, (unsigned)(lev * pl4));
            v.hpt = ldb_f(rHP, T.h4, (unsigned)((nv + lev) * pl4));
        } else {
            v.oh = ldb_d(rH, T.o8, (unsigned)(lev * pl8));
            v.opt = ldb_d(rPt, T.o8, (unsigned)(lev * pl8));
            v.hh = ldb_d(rH, T.h8, (unsigned)(lev * pl8));
            v.hpt = ldb_d(rPt, T.h8, (unsigned)(lev * pl8));
        }
    };
    auto put = [&](const In &v, int b) {
        sM4[b][T.ir] = make_float4(v.om[0], v.om[1], v.om[2], 0.f);
        sHT[b][T.ir] = HT{v.oh, v.opt};
        if constexpr (!SPLIT) {
#pragma unroll
            for (int k = 0; k < 3; k++) sV1[b][k][T.ir] = (double)v.om[k] + v.omk[k];
        }
        if (T.hp >= 0) {
            sM4[b][T.hp] = make_float4(v.hm[0], v.hm[1], v.hm[2], 0.f);
            sHT[b][T.hp] = HT{v.hh, v.hpt};
            if constexpr (!SPLIT) {
#pragma unroll
                for (int k = 0; k < 3; k++) sV1[b][k][T.hp] = (double)v.hm[k] + v.hmk[k];
            }
        }
    };
    auto body = [&](int lev, In &R, In &L) {
        const double slowp = L.slowp, slowrho = L.slowrho, dptk = SPLIT ? L.omk[0] : 0.0;
        load(L, min(lev + 2, l1 - 1));
        const int b = (lev - l0) & 1;
        // prepare_point_w's sums, slot by slot (each component's sum keeps its
        // slot order): one 16-byte LDS read of (Mhs, 0) and one of (h, pt) per slot
        double sr[3], sp[3], st[3];
#pragma unroll
        for (int s = 0; s < 7; s++) {
            const float4 mq = sM4[b][T.pos[s]];
            const HT hq = sHT[b][T.pos[s]];
            const double m[3] = {mq.x, mq.y, mq.z};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double dd = wv[s * 3 + k];
                const double v1 = SPLIT ? m[k] : sV1[b][SPLIT ? 0 : k][T.pos[s]];
                if (s == 0) {
                    sr[k] = dd * m[k];
                    sp[k] = dd * m[k] * (double)hq.x;
                    st[k] = dd * v1 * (double)hq.y;
                } else {
                    sr[k] = sr[k] + dd * m[k];
                    sp[k] = sp[k] + dd * m[k] * (double)hq.x;
                    st[k] = st[k] + dd * v1 * (double)hq.y;
                }
            }
        }
        {
            const double rs = sRs[lev - l0], RoC = g.Rd / (g.Cp - g.Rd);
            double nflxr = 0.0, nflxp = 0.0, dr = 0.0, dpt = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                nflxr -= rs * sr[k];
                nflxp -= (RoC)*rs * sp[k];
                dr += rs * sr[k];
                dpt += rs * st[k];
            }
            LV(f.Sp, lev) = nflxp + slowp;
            LV(f.hflx_r, lev) = dr;  // Sd = -dr + SlowRho is formed by the vertical solve
            LV(f.hflx_pt, lev) = SPLIT ? dpt + dptk : dpt;
        }
        if (lev + 1 < l1) put(R, b ^ 1);
        __syncthreads();
    };
    In A, B;
    load(A, l0);
    load(B, min(l0 + 1, l1 - 1));
    put(A, 0);
    __syncthreads();
    for (int lev = l0; lev < l1; lev += 2) {
        body(lev, B, A);
        if (lev + 1 < l1) body(lev + 1, A, B);
    }
}

// Prepare_Implicit_Vertical (thor_vertical_int.h:566-771) and the horizontal
// flux divergences of Density_Pressure_Eqs of the FP64 form over level
// chunks: prepare_point's sums with FP64 inputs, in its operation order and
// without contraction (bit for bit k_prepare and k_prepare_ml). The divergence
// weights stay in registers for the whole chunk and the next level's tile is
// loaded while the current one is summed, as in k_prepare_lc.
__global__ __launch_bounds__(256, 2) void k_prepare_lc64(Geo g, Fields f, int lc) {
    int r, l0, l1;
    if (!block_tile_lc(g, lc, r, l0, l1)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const double *Mhs = f.Mhs;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (pole_ctx_lc(g, l0, l1, c, lev))
            prepare_point(
                g, f, c.p, lev, [&](int s, int k) { return VEC(Mhs, k, lev, c.gi(s)); },
                [&](int s) { return LVP(f.h, lev, c.gi(s)); },
                [&](int s, int k) { return VEC(Mhs, k, lev, c.gi(s)) + VEC(f.Mhk, k, lev, c.gi(s)); },
                [&](int s) { return LVP(f.pt, lev, c.gi(s)); });
        return;
    }
    __shared__ double sM[2][3][NHL2], sV1[2][3][NHL2], sRs[kLcMax];
    __shared__ double2 sHT[2][NHL2];
    if ((int)threadIdx.x < l1 - l0) sRs[threadIdx.x] = rscale_at(g, l0 + threadIdx.x);
    const LcThread T = lc_thread(g, r);
    const int p = T.p;
    double wv[21];
#pragma unroll
    for (int i = 0; i < 21; i++) wv[i] = g.div[(size_t)i * S + p];
    const size_t pl8 = S * 8;
    const auto rM = rsrc(Mhs, 3 * nv * pl8), rMk = rsrc(f.Mhk, 3 * nv * pl8), rH = rsrc(f.h, nv * pl8),
               rPt = rsrc(f.pt, nv * pl8), rSp = rsrc(f.Slowp, nv * pl8), rSr = rsrc(f.SlowRho, nv * pl8);
    struct In {
        double slowp, slowrho, om[3], hm[3], omk[3], hmk[3], oh, hh, opt, hpt;
    };
    auto load = [&](In &v, int lev) {
        v.slowp = ldb_d(rSp, T.o8, (unsigned)(lev * pl8));
        v.slowrho = ldb_d(rSr, T.o8, (unsigned)(lev * pl8));
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v.om[k] = ldb_d(rM, T.o8, (unsigned)((k * nv + lev) * pl8));
            v.hm[k] = ldb_d(rM, T.h8, (unsigned)((k * nv + lev) * pl8));
            v.omk[k] = ldb_d(rMk, T.o8, (unsigned)((k * nv + lev) * pl8));
            v.hmk[k] = ldb_d(rMk, T.h8, (unsigned)((k * nv + lev) * pl8));
        }
        v.oh = ldb_d(rH, T.o8, (unsigned)(lev * pl8));
        v.opt = ldb_d(rPt, T.o8, (unsigned)(lev * pl8));
        v.hh = ldb_d(rH, T.h8, (unsigned)(lev * pl8));
        v.hpt = ldb_d(rPt, T.h8, (unsigned)(lev * pl8));
    };
    auto put = [&](const In &v, int b) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            sM[b][k][T.ir] = v.om[k];
            sV1[b][k][T.ir] = v.om[k] + v.omk[k];
        }
        sHT[b][T.ir] = make_double2(v.oh, v.opt);
        if (T.hp >= 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                sM[b][k][T.hp] = v.hm[k];
                sV1[b][k][T.hp] = v.hm[k] + v.hmk[k];
            }
            sHT[b][T.hp] = make_double2(v.hh, v.hpt);
        }
    };
    auto body = [&](int lev, In &R, In &L) {
        const double slowp = L.slowp, slowrho = L.slowrho;
        load(L, min(lev + 2, l1 - 1));
        const int b = (lev - l0) & 1;
        double sr[3], sp[3], st[3];
#pragma unroll
        for (int s = 0; s < 7; s++) {
            const double2 hq = sHT[b][T.pos[s]];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double dd = wv[s * 3 + k], m = sM[b][k][T.pos[s]], v1 = sV1[b][k][T.pos[s]];
                if (s == 0) {
                    sr[k] = dd * m;
                    sp[k] = dd * m * hq.x;
                    st[k] = dd * v1 * hq.y;
                } else {
                    sr[k] = sr[k] + dd * m;
                    sp[k] = sp[k] + dd * m * hq.x;
                    st[k] = st[k] + dd * v1 * hq.y;
                }
            }
        }
        {
            const double rs = sRs[lev - l0], RoC = g.Rd / (g.Cp - g.Rd);
            double nflxr = 0.0, nflxp = 0.0, dr = 0.0, dpt = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                nflxr -= rs * sr[k];
                nflxp -= (RoC)*rs * sp[k];
                dr += rs * sr[k];
                dpt += rs * st[k];
            }
            LV(f.Sp, lev) = nflxp + slowp;
            LV(f.Sd, lev) = nflxr + slowrho;
            LV(f.hflx_r, lev) = dr;
            LV(f.hflx_pt, lev) = dpt;
        }
        if (lev + 1 < l1) put(R, b ^ 1);
        __syncthreads();
    };
    In A, B;
    load(A, l0);
    load(B, min(l0 + 1, l1 - 1));
    put(A, 0);
    __syncthreads();
    for (int lev = l0; lev < l1; lev += 2) {
        body(lev, B, A);
        if (lev + 1 < l1) body(lev + 1, A, B);
    }
}

// Last Diffusion_Op pass with Correct_Horizontal (thor_diff.h:524-915) of
// the mixed-precision path over level chunks: the FP32 Laplacian of the first
// pass's output (variables 1..5 weighted by the stage density, as
// diff2_divm0_point forms them), scaled by -Kdh4 sign, with the radial part
// of the momentum components removed.
__global__ __launch_bounds__(256, 3) void k_diff2_lc(Geo g, Fields f, double sign, int lc) {
#pragma clang fp contract(fast)
    int r, l0, l1;
    if (!block_tile_lc(g, lc, r, l0, l1)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const float *diff = f.mdiff;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (pole_ctx_lc(g, l0, l1, c, lev))
            diff2_divm0_point<true>(
                g, f, c.p, lev, 5, true, 1, 0, sign, [&](int s, int var) { return VEC(diff, var, lev, c.gi(s)); },
                [&](int s) { return LVP(f.Rhok, lev, c.gi(s)); }, [&](int, int) { return 0.0; },
                [&](int) { return 0.0; }, [&](int) { return 0.0; });
        return;
    }
    __shared__ float sA[2][6][NHL2], sL[4][kLcMax];
    __shared__ double sK[kLcMax];
    stage_lvc(g, l0, l1, sL);
    if ((int)threadIdx.x < l1 - l0) sK[threadIdx.x] = -g.Kdh4[l0 + threadIdx.x] * sign;
    const LcThread T = lc_thread(g, r);
    const int p = T.p;
    float wl[6];
#pragma unroll
    for (int s = 0; s < 6; s++) wl[s] = g.wlap[(size_t)(s + 1) * S + p];
    const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
    const size_t pl4 = S * 4, pl8 = S * 8;
    const auto rDf = rsrc(diff, 6 * nv * pl4), rR = rsrc(f.Rhok, nv * pl8), rRh = rsrc(f.mdiffrh, nv * pl4),
               rWd = rsrc(f.mdiffw, nv * pl4), rPd = rsrc(f.mdiffpr, nv * pl4), rMd = rsrc(f.mdiffmh, 3 * nv * pl4);
    struct In {
        float od[6], hd[6];
        double orho, hrho;
    };
    auto load = [&](In &v, int lev) {
#pragma unroll
        for (int k = 0; k < 6; k++) {
            v.od[k] = ldb_f(rDf, T.o4, (unsigned)((k * nv + lev) * pl4));
            v.hd[k] = ldb_f(rDf, T.h4, (unsigned)((k * nv + lev) * pl4));
        }
        v.orho = ldb_d(rR, T.o8, (unsigned)(lev * pl8));
        v.hrho = ldb_d(rR, T.h8, (unsigned)(lev * pl8));
    };
    auto put = [&](const In &v, int b) {
        sA[b][0][T.ir] = v.od[0];
#pragma unroll
        for (int k = 1; k < 6; k++) sA[b][k][T.ir] = (float)(v.od[k] * v.orho);
        if (T.hp >= 0) {
            sA[b][0][T.hp] = v.hd[0];
#pragma unroll
            for (int k = 1; k < 6; k++) sA[b][k][T.hp] = (float)(v.hd[k] * v.hrho);
        }
    };
    auto body = [&](int lev, In &R, In &L) {
        load(L, min(lev + 2, l1 - 1));
        const int b = (lev - l0) & 1, i = lev - l0;
        const float rs2 = sL[1][i];
        const double sdiff = sK[i];
        double out[6];
#pragma unroll
        for (int var = 0; var < 6; var++) {
            const float a0 = sA[b][var][T.pos[0]];
            float sum = wl[0] * (sA[b][var][T.pos[1]] - a0);
#pragma unroll
            for (int s = 2; s < 7; s++) sum += wl[s - 1] * (sA[b][var][T.pos[s]] - a0);
            out[var] = (double)(rs2 * sum) * sdiff;
        }
        if (f.mdsp) {  // diffusive sponge, as diff2_divm0_point
            const double rr = LV(f.Rhok, lev), k2 = g.Kdh2[lev];
            for (int var = 1; var <= 4; var++)
                out[var] += g.ssp * rr * k2 * (double)VEC(f.mdsp, var, lev, p) * g.dsp_scale;
        }
        const unsigned l4 = (unsigned)(lev * pl4);
        stb_f(rRh, T.o4, l4, (float)out[0]);
        stb_f(rWd, T.o4, l4, (float)out[4]);
        stb_f(rPd, T.o4, l4, (float)out[5]);
        const double dmhr = fx * out[1] + fy * out[2] + fz * out[3];
        stb_f(rMd, T.o4, (unsigned)((0 * nv + lev) * pl4), (float)(out[1] + -fx * dmhr));
        stb_f(rMd, T.o4, (unsigned)((1 * nv + lev) * pl4), (float)(out[2] + -fy * dmhr));
        stb_f(rMd, T.o4, (unsigned)((2 * nv + lev) * pl4), (float)(out[3] + -fz * dmhr));
        if (lev + 1 < l1) put(R, b ^ 1);
        __syncthreads();
    };
    In A, B;
    load(A, l0);
    load(B, min(l0 + 1, l1 - 1));
    put(A, 0);
    __syncthreads();
    for (int lev = l0; lev < l1; lev += 2) {
        body(lev, B, A);
        if (lev + 1 < l1) body(lev + 1, A, B);
    }
}

// Divergence-damping pass 0 (DivM_Op) on the deviations Mhs / Whs in FP32,
// as divm_w<false>. Whs lives on interfaces: the top value of level lev is
// the bottom value of lev + 1, so each level loads one new interface.
// STAGE: pass 0 of the slow part, on the stage values Mhk / Whk (FP64 storage,
// rounded to FP32 as diff2_divm0_point does).
// blocks per CU of the mixed divm0 (4 spills: 24 -> 36 us at g6 L48)
#ifndef THOR_DIVM0_LC_MINB
#define THOR_DIVM0_LC_MINB 3
#endif
template <bool STAGE>
__global__ __launch_bounds__(256, THOR_DIVM0_LC_MINB) void k_divm0_lc(Geo g, Fields f, int lc) {
#pragma clang fp contract(fast)
    int r, l0, l1;
    if (!block_tile_lc(g, lc, r, l0, l1)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const float *Mhs = f.mMhs;
    float *divg = f.mdivg;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (pole_ctx_lc(g, l0, l1, c, lev)) {
            float o3[3];
            const double *Wv = STAGE ? f.Whk : f.Whs;
            divm_w<false>(
                g, c.p, lev,
                [&](int s, int k) {
                    if constexpr (STAGE) return (float)VEC(f.Mhk, k, lev, c.gi(s));
                    else return VEC(Mhs, k, lev, c.gi(s));
                },
                [&](int s) { return (float)(LVP(Wv, lev, c.gi(s)) - LVP(Wv, lev + 1, c.gi(s))); }, o3);
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(divg, k, lev, c.p) = o3[k];
        }
        return;
    }
    // staged point: (M_x, M_y, M_z, Wl - Wt) as one 16-byte LDS word per slot
    __shared__ float4 sMD[2][NHL2];
    __shared__ float sL[4][kLcMax];
    stage_lvc(g, l0, l1, sL);
    const LcThread T = lc_thread(g, r);
    const int p = T.p;
    float wd[63], wu[21];
#pragma unroll
    for (int i = 0; i < 63; i++) wd[i] = g.wdd[(size_t)i * S + p];
#pragma unroll
    for (int i = 0; i < 21; i++) wu[i] = g.udd[(size_t)i * S + p];
    const size_t pl4 = S * 4, pl8 = S * 8;  // bytes of one FP32 / FP64 level plane
    const auto rM = STAGE ? rsrc(f.Mhk, 3 * nv * pl8) : rsrc(Mhs, 3 * nv * pl4),
               rW = rsrc(STAGE ? f.Whk : f.Whs, (nv + 1) * pl8), rD = rsrc(divg, 3 * nv * pl4);
    // loaded values of one level: Mhs and the top interface's Whs
    struct In {
        float om[3], hm[3];
        double ow, hw;
    };
    auto load = [&](In &v, int lev) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if constexpr (STAGE) {
                v.om[k] = (float)ldb_d(rM, T.o8, (unsigned)((k * nv + lev) * pl8));
                v.hm[k] = (float)ldb_d(rM, T.h8, (unsigned)((k * nv + lev) * pl8));
            } else {
                v.om[k] = ldb_f(rM, T.o4, (unsigned)((k * nv + lev) * pl4));
                v.hm[k] = ldb_f(rM, T.h4, (unsigned)((k * nv + lev) * pl4));
            }
        }
        v.ow = ldb_d(rW, T.o8, (unsigned)((lev + 1) * pl8));
        v.hw = ldb_d(rW, T.h8, (unsigned)((lev + 1) * pl8));
    };
    // Whs at the bottom interface of the next level to be put
    double owl = ldb_d(rW, T.o8, (unsigned)(l0 * pl8)), hwl = ldb_d(rW, T.h8, (unsigned)(l0 * pl8));
    auto put = [&](const In &v, int b) {
        sMD[b][T.ir] = make_float4(v.om[0], v.om[1], v.om[2], (float)(owl - v.ow));
        owl = v.ow;
        if (T.hp >= 0) sMD[b][T.hp] = make_float4(v.hm[0], v.hm[1], v.hm[2], (float)(hwl - v.hw));
        hwl = v.hw;
    };
    auto body = [&](int lev, In &R, In &L) {
        load(L, min(lev + 2, l1 - 1));
        const int b = (lev - l0) & 1;
        const float rs = sL[0][lev - l0], rs2 = sL[1][lev - l0], dz = sL[2][lev - l0];
        float a[7][3], d[7];
#pragma unroll
        for (int s = 0; s < 7; s++) {
            const float4 q = sMD[b][T.pos[s]];
            a[s][0] = q.x;
            a[s][1] = q.y;
            a[s][2] = q.z;
            d[s] = q.w;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float sum = wd[c * 21] * a[0][0];
#pragma unroll
            for (int i = 1; i < 21; i++) sum += wd[c * 21 + i] * a[i / 3][i % 3];
            float su = wu[c * 7] * d[0];
#pragma unroll
            for (int s = 1; s < 7; s++) su += wu[c * 7 + s] * d[s];
            stb_f(rD, T.o4, (unsigned)((c * nv + lev) * pl4), rs2 * sum + rs * dz * su);
        }
        if (lev + 1 < l1) put(R, b ^ 1);
        __syncthreads();
    };
    In A, B;
    load(A, l0);
    load(B, min(l0 + 1, l1 - 1));
    put(A, 0);
    __syncthreads();
    for (int lev = l0; lev < l1; lev += 2) {
        body(lev, B, A);
        if (lev + 1 < l1) body(lev + 1, A, B);
    }
}

// Divergence-damping pass 1 of the slow part (the DivM term of
// Compute_Slow_Modes, thor_slowmodes.h:82-105) in FP32 over level chunks, with
// the operation order of divm_w<true>, written to dm [3][nv] for k_slow_lc.
// Tiles only: the pole block of k_slow_lc forms its own DivM.
__global__ __launch_bounds__(256, 3) void k_divm1_lc(Geo g, Fields f, float *dm, int lc) {
    int r, l0, l1;
    if (!block_tile_lc(g, lc, r, l0, l1) || r == g.nr) return;
    const size_t S = g.S;
    const int nv = g.nv;
    __shared__ float sG[2][3][NHL2], sL[4][kLcMax];
    stage_lvc(g, l0, l1, sL);
    const LcThread T = lc_thread(g, r);
    const int p = T.p;
    float wd[63];
#pragma unroll
    for (int i = 0; i < 63; i++) wd[i] = g.wdd[(size_t)i * S + p];
    const float fx = (float)g.fr[p], fy = (float)g.fr[S + p], fz = (float)g.fr[2 * S + p];
    const size_t pl4 = S * 4;
    const auto rG = rsrc(f.mdivg, 3 * nv * pl4), rD = rsrc(dm, 3 * nv * pl4);
    struct In {
        float og[3], hg[3];
    };
    auto load = [&](In &v, int lev) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v.og[k] = ldb_f(rG, T.o4, (unsigned)((k * nv + lev) * pl4));
            v.hg[k] = ldb_f(rG, T.h4, (unsigned)((k * nv + lev) * pl4));
        }
    };
    auto put = [&](const In &v, int b) {
#pragma unroll
        for (int k = 0; k < 3; k++) sG[b][k][T.ir] = v.og[k];
        if (T.hp >= 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) sG[b][k][T.hp] = v.hg[k];
        }
    };
    auto body = [&](int lev, In &R, In &L) {
        load(L, min(lev + 2, l1 - 1));
        const int b = (lev - l0) & 1;
        const float rs2 = sL[1][lev - l0], sdiff = sL[3][lev - l0];
        float av[7][3];
#pragma unroll
        for (int s = 0; s < 7; s++)
#pragma unroll
            for (int k = 0; k < 3; k++) av[s][k] = sG[b][k][T.pos[s]];
        float o[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float sum = 0.f;
#pragma unroll
            for (int s = 0; s < 7; s++)
#pragma unroll
                for (int k = 0; k < 3; k++) sum += wd[c * 21 + s * 3 + k] * av[s][k];
            o[c] = rs2 * sum;
        }
        const float lapr = fx * o[0] + fy * o[1] + fz * o[2];
        stb_f(rD, T.o4, (unsigned)((0 * nv + lev) * pl4), -(o[0] - fx * lapr) * sdiff);
        stb_f(rD, T.o4, (unsigned)((1 * nv + lev) * pl4), -(o[1] - fy * lapr) * sdiff);
        stb_f(rD, T.o4, (unsigned)((2 * nv + lev) * pl4), -(o[2] - fz * lapr) * sdiff);
        if (lev + 1 < l1) put(R, b ^ 1);
        __syncthreads();
    };
    In A, B;
    load(A, l0);
    load(B, min(l0 + 1, l1 - 1));
    put(A, 0);
    __syncthreads();
    for (int lev = l0; lev < l1; lev += 2) {
        body(lev, B, A);
        if (lev + 1 < l1) body(lev + 1, A, B);
    }
}

// Compute_Slow_Modes (thor_slowmodes.h:51-313) of the mixed-precision path
// over level chunks, with slow_point's arithmetic and operation order, so its
// results equal k_slow<true>'s bit for bit. The FP64 gradient weights of the
// point stay in registers for the chunk and also serve the divergence: THOR's
// div and grad operators share their neighbour weights (grid.cu:2656-2831,
// 3015-3160 differ only in the centre weight; the host checks this before
// choosing this kernel). DivM comes from k_divm1_lc. Each level carries its
// lower neighbour's values (advection, density, pressure, momentum, diffw) to
// the next, and stages the stage momentum, pressure and h of the tile.
__global__ __launch_bounds__(256, 1) void k_slow_lc(Geo g, Fields f, const float *dm, int divdamp, int rk, int lc) {
    constexpr bool MX = true;
    int r, l0, l1;
    if (!block_tile_lc(g, lc, r, l0, l1)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (pole_ctx_lc(g, l0, l1, c, lev)) {
            const float *divg = f.mdivg;
            slow_point<true>(
                g, f, c.p, lev, 5, true, divdamp, rk, [&](int s, int k) { return VEC(divg, k, lev, c.gi(s)); },
                [&](int s) { return LVP(f.pk, lev, c.gi(s)); }, [&](int s) { return LVP(f.h, lev, c.gi(s)); },
                [&](int s, int k) { return VEC(f.Mhk, k, lev, c.gi(s)); });
        }
        return;
    }
    __shared__ double sM[2][3][NHL2], sP[2][NHL2], sH[2][NHL2], sT[2][NHL2];
    __shared__ double sW[24][kTile];  // the thread's 21 gradient weights and 3 divergence centre weights
    // Per-level constants of the chunk, formed once with slow_point's
    // arithmetic: Alt[l0 - 1 .. l1 - 1], Alth[l0 .. l1], (A / (A + Alt))^2 at
    // the same levels, and per level r2p, r2m, r2l, rscale and dr2dz.
    __shared__ double sAlt[kLcMax + 1], sAlth[kLcMax + 1], sG2[kLcMax + 1], sR[5][kLcMax];
    const double A = g.A, Gravit = g.Gravit, Rd = g.Rd, Cv = g.Cp - g.Rd;
    {
        const int t = threadIdx.x, n = l1 - l0;
        if (t <= n) {
            const double alt = l0 - 1 + t >= 0 ? g.Alt[l0 - 1 + t] : 0.0;
            sAlt[t] = alt;
            sAlth[t] = g.Alth[l0 + t];
            sG2[t] = g.ghv ? XPOW(A / (A + alt), 2) : 0.0;
        }
        if (t < n) {
            const int lev = l0 + t;
            const double altht = g.Alth[lev + 1], althl = g.Alth[lev], alt = g.Alt[lev];
            if (g.deep) {
                sR[0][t] = XPOW(altht + A, 2.0);
                sR[1][t] = XPOW(alt + A, 2.0);
                sR[2][t] = XPOW(althl + A, 2.0);
                sR[3][t] = A / (alt + A);
                sR[4][t] = (XPOW(altht + A, 3.0) - XPOW(althl + A, 3.0)) / 3;
            } else {
                sR[0][t] = sR[1][t] = sR[2][t] = 1.0;
                sR[3][t] = 1.0;
                sR[4][t] = 1.0 * (altht - althl);
            }
        }
    }
    const LcThread T = lc_thread(g, r);
    const int p = T.p;
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < 21; i++) sW[i][t] = g.grad[(size_t)i * S + p];
#pragma unroll
        for (int k = 0; k < 3; k++) sW[21 + k][t] = g.div[(size_t)k * S + p];
    }
    auto wg = [&](int i) { return sW[i][threadIdx.x]; };
    auto wd0 = [&](int k) { return sW[21 + k][threadIdx.x]; };
    const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
    const size_t pl4 = S * 4, pl8 = S * 8;
    const auto rMk = rsrc(f.Mhk, 3 * nv * pl8), rPk = rsrc(f.pk, nv * pl8), rH = rsrc(f.h, nv * pl8),
               rPt = rsrc(f.pt, nv * pl8),
               rAdv = rsrc(f.mAdv, 3 * nv * pl4), rRk = rsrc(f.Rhok, nv * pl8),
               rWk = rsrc(f.Whk, (nv + 1) * pl8), rHh = rsrc(f.hh, (nv + 1) * pl8),
               rDmh = rsrc(f.mdiffmh, 3 * nv * pl4), rdM = rsrc(f.dMh, 3 * nv * pl8),
               rDw = rsrc(f.mdiffw, nv * pl4), rdW = rsrc(f.dWh, (nv + 1) * pl8),
               rDr = rsrc(f.mdiffrh, nv * pl4), rDp = rsrc(f.mdiffpr, nv * pl4), rGt = rsrc(f.gtil, nv * pl8),
               rQ = rsrc(f.Qheat, nv * pl8), rDM = rsrc(dm, 3 * nv * pl4), rM = rsrc(f.Mh, 3 * nv * pl8),
               rR = rsrc(f.Rho, nv * pl8), rPr = rsrc(f.p, nv * pl8), rW = rsrc(f.Wh, (nv + 1) * pl8);
    struct Stg {  // staged tile values: own point and halo point, two levels ahead
        double om[3], op, oh, ot, hm[3], hp, hh, ht;
    };
    struct Pt {  // the point's own values, one level ahead
        float adv[3], dmh[3], dM[3], dw, drh, dpr;
        double rhok, whk1, hh1, dmv[3], dwh, gtil, q;
        double mh[3], rho, pr, wh;  // state, for the deviations of rk > 0
    };
    auto load_stg = [&](Stg &v, int lev) {
        const unsigned l8 = (unsigned)(lev * pl8);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v.om[k] = ldb_d(rMk, T.o8, (unsigned)((k * nv + lev) * pl8));
            v.hm[k] = ldb_d(rMk, T.h8, (unsigned)((k * nv + lev) * pl8));
        }
        v.op = ldb_d(rPk, T.o8, l8);
        v.hp = ldb_d(rPk, T.h8, l8);
        v.oh = ldb_d(rH, T.o8, l8);
        v.hh = ldb_d(rH, T.h8, l8);
        v.ot = ldb_d(rPt, T.o8, l8);
        v.ht = ldb_d(rPt, T.h8, l8);
    };
    auto load_pt = [&](Pt &v, int lev) {
        const unsigned l8 = (unsigned)(lev * pl8), l4 = (unsigned)(lev * pl4), u8 = (unsigned)((lev + 1) * pl8);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v.adv[k] = ldb_f(rAdv, T.o4, (unsigned)((k * nv + lev) * pl4));
            v.dmh[k] = ldb_f(rDmh, T.o4, (unsigned)((k * nv + lev) * pl4));
            v.dM[k] = divdamp ? ldb_f(rDM, T.o4, (unsigned)((k * nv + lev) * pl4)) : 0.f;
            v.dmv[k] = ldb_d(rdM, T.o8, (unsigned)((k * nv + lev) * pl8));
        }
        v.dw = ldb_f(rDw, T.o4, l4);
        v.drh = ldb_f(rDr, T.o4, l4);
        v.dpr = ldb_f(rDp, T.o4, l4);
        v.rhok = ldb_d(rRk, T.o8, l8);
        v.whk1 = ldb_d(rWk, T.o8, u8);
        v.hh1 = ldb_d(rHh, T.o8, u8);
        v.dwh = ldb_d(rdW, T.o8, l8);
        v.gtil = ldb_d(rGt, T.o8, l8);
        v.q = ldb_d(rQ, T.o8, l8);
        if (rk) {
#pragma unroll
            for (int k = 0; k < 3; k++) v.mh[k] = ldb_d(rM, T.o8, (unsigned)((k * nv + lev) * pl8));
            v.rho = ldb_d(rR, T.o8, l8);
            v.pr = ldb_d(rPr, T.o8, l8);
            v.wh = ldb_d(rW, T.o8, l8);
        }
    };
    auto put = [&](const Stg &v, int b) {
#pragma unroll
        for (int k = 0; k < 3; k++) sM[b][k][T.ir] = v.om[k];
        sP[b][T.ir] = v.op;
        sH[b][T.ir] = v.oh;
        sT[b][T.ir] = v.ot;
        if (T.hp >= 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) sM[b][k][T.hp] = v.hm[k];
            sP[b][T.hp] = v.hp;
            sH[b][T.hp] = v.hh;
            sT[b][T.hp] = v.ht;
        }
    };
    // values of level lev - 1 carried to level lev
    float cadv[3] = {0.f, 0.f, 0.f}, cdw = 0.f;
    double crho = 0.0, cpk = 0.0, cm[3] = {0.0, 0.0, 0.0}, cwh, chh;
    if (l0 > 0) {
        const int lm = l0 - 1;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            cadv[k] = ldb_f(rAdv, T.o4, (unsigned)((k * nv + lm) * pl4));
            cm[k] = ldb_d(rMk, T.o8, (unsigned)((k * nv + lm) * pl8));
        }
        cdw = ldb_f(rDw, T.o4, (unsigned)(lm * pl4));
        crho = ldb_d(rRk, T.o8, (unsigned)(lm * pl8));
        cpk = ldb_d(rPk, T.o8, (unsigned)(lm * pl8));
    }
    cwh = ldb_d(rWk, T.o8, (unsigned)(l0 * pl8));  // Whk and hh at the chunk's lowest interface
    chh = ldb_d(rHh, T.o8, (unsigned)(l0 * pl8));
    auto *SlowMh = f.mSlowMh;
    float *Mhs = f.mMhs;
    const auto rSm = rsrc(SlowMh, 3 * nv * pl4);
    auto body = [&](int lev, Stg &R, Stg &L, const Pt &v, Pt &nxt) {
        load_pt(nxt, min(lev + 1, l1 - 1));
        load_stg(L, min(lev + 2, l1 - 1));
        const int b = (lev - l0) & 1, i = lev - l0;
        double advx = v.adv[0], advy = v.adv[1], advz = v.adv[2];
        const double rho = v.rhok;
        double advrl = 0, rhol = 0, altl = 0, pressurel = 0;
        if (lev != 0) {
            const double ax = cadv[0], ay = cadv[1], az = cadv[2];
            advrl = ax * fx + ay * fy + az * fz;
            rhol = crho;
            altl = sAlt[i];
            pressurel = cpk;
        }
        double advr = advx * fx + advy * fy + advz * fz;
        advx += -advr * fx;
        advy += -advr * fy;
        advz += -advr * fz;
        const double altht = sAlth[i + 1], althl = sAlth[i], alt = sAlt[i + 1];
        const double M0x = sM[b][0][T.pos[0]], M0y = sM[b][1][T.pos[0]], M0z = sM[b][2][T.pos[0]];
        if (!g.nonhydro) {
            advr = (XPOW(M0x, 2) + XPOW(M0y, 2) + XPOW(M0z, 2)) / (rho * (alt + A));
            if (lev != 0) advrl = (XPOW(cm[0], 2) + XPOW(cm[1], 2) + XPOW(cm[2], 2)) / (rhol * (altl + A));
        }
        const double dz = altht - althl;
        const double r2p = sR[0][i], r2m = sR[1][i], r2l = sR[2][i], rscale = sR[3][i], dr2dz = sR[4][i];
        double pv[7], hv[7], tv[7];
#pragma unroll
        for (int s = 0; s < 7; s++) {
            pv[s] = sP[b][T.pos[s]];
            hv[s] = sH[b][T.pos[s]];
            tv[s] = sT[b][T.pos[s]];
        }
        // nflxv, nflxr, nflxp as slow_point; dptk: the stage part of the
        // pressure flux divergence of the substeps (prepare_point_w's st with
        // the stage momentum), formed here once per stage
        double nflxv[3], nflxr = 0.0, nflxp = 0.0, dptk = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double m0 = sM[b][k][T.pos[0]];
            double sg = wg(k) * pv[0];
            double sr = wd0(k) * m0;
            double sp = wd0(k) * m0 * hv[0];
            double st = wd0(k) * m0 * tv[0];
#pragma unroll
            for (int s = 1; s < 7; s++) {
                const double w = wg(s * 3 + k), m = sM[b][k][T.pos[s]];
                sg = sg + w * pv[s];
                sr = sr + w * m;
                sp = sp + w * m * hv[s];
                st = st + w * m * tv[s];
            }
            nflxv[k] = rscale * sg;
            nflxr += rscale * sr;
            nflxp += rscale * sp;
            dptk += rscale * st;
        }
        LV(f.hflx_ptk, lev) = dptk;
        double whl, wht, hhl, hht;
        if (lev == 0) {
            whl = 0.0;
            wht = v.whk1;
            hhl = 0.0;
            hht = v.hh1;
        } else {
            whl = cwh;
            wht = v.whk1;
            hhl = chh;
            hht = v.hh1;
        }
        const double dwdz = (wht * r2p - whl * r2l) / (dr2dz);
        const double dwhdz = (wht * r2p * hht - whl * r2l * hhl) / (dz * r2m);
        const double dpr = nflxv[0] * fx + nflxv[1] * fy + nflxv[2] * fz;
        const double fv[3] = {fx, fy, fz}, ad[3] = {advx, advy, advz};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double DivM = v.dM[k];
            const float o = -(nflxv[k] - dpr * fv[k]) - ad[k] + DivM + v.dmh[k] + 0.0 + v.dmv[k];
            stb_f(rSm, T.o4, (unsigned)((k * nv + lev) * pl4), o);
        }
        if (lev == 0) {
            LV(f.SlowWh, 0) = 0.0;
            LV(f.SlowWh, nv) = 0.0;
        } else {
            const double xi = althl, xim = altl, xip = alt;
            const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            double swr;
            if (g.deep) {
                if (g.ghv)
                    swr = (advrl + rhol * Gravit * sG2[i]) * intt + (advr + rho * Gravit * sG2[i + 1]) * intl;
                else
                    swr = (advrl + rhol * Gravit) * intt + (advr + rho * Gravit) * intl;
            } else {
                if (g.ghv)
                    swr = (rhol * Gravit * sG2[i]) * intt + (rho * Gravit * sG2[i + 1]) * intl;
                else
                    swr = (rhol * Gravit) * intt + (rho * Gravit) * intl;
            }
            if (g.nonhydro) swr += (-cdw - 0.0) * intt + (-v.dw - 0.0) * intl;
            const double dzl = alt - altl;
            const double dpdz = (pv[0] - pressurel) / dzl;
            swr += dpdz;
            swr -= v.dwh;
            LV(f.SlowWh, lev) = -swr;
        }
        nflxr += dwdz;
        LV(f.SlowRho, lev) = -nflxr + v.drh + 0.0;
        const double rr = 1.0 / rho;
        const double vgp = (nflxv[0] - dpr * fx) * M0x * rr + (nflxv[1] - dpr * fy) * M0y * rr
                           + (nflxv[2] - dpr * fz) * M0z * rr;
        double sp = (Rd / Cv) * (-nflxp - v.gtil - dwhdz + vgp + v.q);
        sp += v.dpr + 0.0;
        LV(f.Slowp, lev) = sp;
        if (rk == 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(Mhs, k, lev, p) = 0.0;
            LV(f.Rhos, lev) = 0.0;
            LV(f.ps, lev) = 0.0;
            LV(f.Whs, lev) = 0.0;
            if (lev == nv - 1) LV(f.Whs, nv) = 0.0;
        } else {
            const double Mx = v.mh[0] - M0x, My = v.mh[1] - M0y, Mz = v.mh[2] - M0z;
            const double aux = Mx * fx + My * fy + Mz * fz;
            VEC(Mhs, 0, lev, p) = Mx - aux * fx;
            VEC(Mhs, 1, lev, p) = My - aux * fy;
            VEC(Mhs, 2, lev, p) = Mz - aux * fz;
            LV(f.Rhos, lev) = v.rho - rho;
            LV(f.ps, lev) = v.pr - pv[0];
            LV(f.Whs, lev) = lev == 0 ? 0.0 : v.wh - cwh;
            if (lev == nv - 1) LV(f.Whs, nv) = 0.0;
        }
        // carry this level's values to the next
#pragma unroll
        for (int k = 0; k < 3; k++) cadv[k] = v.adv[k];
        cm[0] = M0x;
        cm[1] = M0y;
        cm[2] = M0z;
        cdw = v.dw;
        crho = rho;
        cpk = pv[0];
        cwh = v.whk1;
        chh = v.hh1;
        if (lev + 1 < l1) put(R, b ^ 1);
        __syncthreads();
    };
    Stg S0, S1;
    Pt P0, P1;
    load_pt(P0, l0);
    load_stg(S0, l0);
    load_stg(S1, min(l0 + 1, l1 - 1));
    put(S0, 0);
    __syncthreads();
    for (int lev = l0; lev < l1; lev += 2) {
        body(lev, S1, S0, P0, P1);
        if (lev + 1 < l1) body(lev + 1, S0, S1, P1, P0);
    }
}

// k_slow_lc for both precision forms, the mixed one with two blocks per CU:
// the per-level constants come from a table in HBM (k_slow_levels, the same
// expressions), the staged tile values are loaded one level ahead instead of
// two, and the values used late in a level are read at its start. The
// arithmetic is slow_point's, so the results equal k_slow<MX>'s bit for bit.
// dm: DivM (divergence-damping pass 1) of the tile points, from k_divm1_lc
// (mixed) or k_divm1_ml (FP64).
template <bool MX>
__global__ __launch_bounds__(256, MX ? 2 : 1) void k_slow_lc2(Geo g, Fields f, const typename FS<MX>::R *dm,
                                                              int divdamp, int rk, int lc) {
    using R = typename FS<MX>::R;
    int r, l0, l1;
    if (!block_tile_lc(g, lc, r, l0, l1)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (pole_ctx_lc(g, l0, l1, c, lev)) {
            const auto *divg = FS<MX>::divg(f);
            slow_point<MX>(
                g, f, c.p, lev, 5, true, divdamp, rk, [&](int s, int k) { return VEC(divg, k, lev, c.gi(s)); },
                [&](int s) { return LVP(f.pk, lev, c.gi(s)); }, [&](int s) { return LVP(f.h, lev, c.gi(s)); },
                [&](int s, int k) { return VEC(f.Mhk, k, lev, c.gi(s)); });
        }
        return;
    }
    __shared__ double sM[2][3][NHL2], sP[2][NHL2], sH[2][NHL2], sT[2][NHL2];
    __shared__ double sW[24][kTile];  // the thread's 21 gradient weights and 3 divergence centre weights
    const double A = g.A, Gravit = g.Gravit, Rd = g.Rd, Cv = g.Cp - g.Rd;
    const double *slv = g.slv;  // per-level constants (k_slow_levels)
    const int nvl = nv;
    const LcThread T = lc_thread(g, r);
    const int p = T.p;
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < 21; i++) sW[i][t] = g.grad[(size_t)i * S + p];
#pragma unroll
        for (int k = 0; k < 3; k++) sW[21 + k][t] = g.div[(size_t)k * S + p];
    }
    auto wg = [&](int i) { return sW[i][threadIdx.x]; };
    auto wd0 = [&](int k) { return sW[21 + k][threadIdx.x]; };
    const double fx = g.fr[p], fy = g.fr[S + p], fz = g.fr[2 * S + p];
    const size_t pl4 = S * 4, pl8 = S * 8, plR = S * sizeof(R);
    // a value of an array of the form's precision at the thread's point
    auto ldR = [&](__amdgpu_buffer_rsrc_t rs, unsigned off) -> R {
        if constexpr (MX) return ldb_f(rs, T.o4, off);
        else return ldb_d(rs, T.o8, off);
    };
    const auto rMk = rsrc(f.Mhk, 3 * nv * pl8), rPk = rsrc(f.pk, nv * pl8), rH = rsrc(f.h, nv * pl8),
               rPt = rsrc(f.pt, nv * pl8),
               rAdv = rsrc(FS<MX>::Adv(f), 3 * nv * plR), rRk = rsrc(f.Rhok, nv * pl8),
               rWk = rsrc(f.Whk, (nv + 1) * pl8), rHh = rsrc(f.hh, (nv + 1) * pl8),
               rDmh = rsrc(FS<MX>::diffmh(f), 3 * nv * plR), rdM = rsrc(f.dMh, 3 * nv * pl8),
               rDw = rsrc(FS<MX>::diffw(f), nv * plR), rdW = rsrc(f.dWh, (nv + 1) * pl8),
               rDr = rsrc(FS<MX>::diffrh(f), nv * plR), rDp = rsrc(FS<MX>::diffpr(f), nv * plR),
               rGt = rsrc(f.gtil, nv * pl8), rQ = rsrc(f.Qheat, nv * pl8), rDM = rsrc(dm, 3 * nv * plR),
               rM = rsrc(f.Mh, 3 * nv * pl8),
               rR = rsrc(f.Rho, nv * pl8), rPr = rsrc(f.p, nv * pl8), rW = rsrc(f.Wh, (nv + 1) * pl8);
    struct Stg {  // staged tile values: own point and halo point, two levels ahead
        double om[3], op, oh, ot, hm[3], hp, hh, ht;
    };
    struct Pt {  // the point's own values, one level ahead
        R adv[3], dmh[3], dM[3], dw;
        double rhok, whk1, hh1, dmv[3];
    };
    auto load_stg = [&](Stg &v, int lev) {
        const unsigned l8 = (unsigned)(lev * pl8);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v.om[k] = ldb_d(rMk, T.o8, (unsigned)((k * nv + lev) * pl8));
            v.hm[k] = ldb_d(rMk, T.h8, (unsigned)((k * nv + lev) * pl8));
        }
        v.op = ldb_d(rPk, T.o8, l8);
        v.hp = ldb_d(rPk, T.h8, l8);
        v.oh = ldb_d(rH, T.o8, l8);
        v.hh = ldb_d(rH, T.h8, l8);
        if constexpr (MX) {  // pt: the stage part of the substeps' pressure flux divergence
            v.ot = ldb_d(rPt, T.o8, l8);
            v.ht = ldb_d(rPt, T.h8, l8);
        } else {
            v.ot = v.ht = 0.0;
        }
    };
    auto load_pt = [&](Pt &v, int lev) {
        const unsigned l8 = (unsigned)(lev * pl8), u8 = (unsigned)((lev + 1) * pl8);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v.adv[k] = ldR(rAdv, (unsigned)((k * nv + lev) * plR));
            v.dmh[k] = ldR(rDmh, (unsigned)((k * nv + lev) * plR));
            v.dM[k] = divdamp ? ldR(rDM, (unsigned)((k * nv + lev) * plR)) : (R)0;
            v.dmv[k] = ldb_d(rdM, T.o8, (unsigned)((k * nv + lev) * pl8));
        }
        v.dw = ldR(rDw, (unsigned)(lev * plR));
        v.rhok = ldb_d(rRk, T.o8, l8);
        v.whk1 = ldb_d(rWk, T.o8, u8);
        v.hh1 = ldb_d(rHh, T.o8, u8);
    };
    auto put = [&](const Stg &v, int b) {
#pragma unroll
        for (int k = 0; k < 3; k++) sM[b][k][T.ir] = v.om[k];
        sP[b][T.ir] = v.op;
        sH[b][T.ir] = v.oh;
        sT[b][T.ir] = v.ot;
        if (T.hp >= 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) sM[b][k][T.hp] = v.hm[k];
            sP[b][T.hp] = v.hp;
            sH[b][T.hp] = v.hh;
            sT[b][T.hp] = v.ht;
        }
    };
    // values of level lev - 1 carried to level lev
    R cadv[3] = {0, 0, 0}, cdw = 0;
    double crho = 0.0, cpk = 0.0, cm[3] = {0.0, 0.0, 0.0}, cwh, chh;
    if (l0 > 0) {
        const int lm = l0 - 1;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            cadv[k] = ldR(rAdv, (unsigned)((k * nv + lm) * plR));
            cm[k] = ldb_d(rMk, T.o8, (unsigned)((k * nv + lm) * pl8));
        }
        cdw = ldR(rDw, (unsigned)(lm * plR));
        crho = ldb_d(rRk, T.o8, (unsigned)(lm * pl8));
        cpk = ldb_d(rPk, T.o8, (unsigned)(lm * pl8));
    }
    cwh = ldb_d(rWk, T.o8, (unsigned)(l0 * pl8));  // Whk and hh at the chunk's lowest interface
    chh = ldb_d(rHh, T.o8, (unsigned)(l0 * pl8));
    auto *SlowMh = FS<MX>::SlowMh(f);
    auto *Mhs = FS<MX>::Mhs(f);
    const auto rSm = rsrc(SlowMh, 3 * nv * plR);
    auto body = [&](int lev, Stg &Rs, const Pt &v, Pt &nxt) {
        load_pt(nxt, min(lev + 1, l1 - 1));
        load_stg(Rs, min(lev + 1, l1 - 1));
        const int b = (lev - l0) & 1;
        // values used late in the level: read at its start
        const R v_drh = ldR(rDr, (unsigned)(lev * plR)), v_dpr = ldR(rDp, (unsigned)(lev * plR));
        const double v_dwh = ldb_d(rdW, T.o8, (unsigned)(lev * pl8)), v_gtil = ldb_d(rGt, T.o8, (unsigned)(lev * pl8)),
                     v_q = ldb_d(rQ, T.o8, (unsigned)(lev * pl8));
        double advx = v.adv[0], advy = v.adv[1], advz = v.adv[2];
        const double rho = v.rhok;
        double advrl = 0, rhol = 0, altl = 0, pressurel = 0;
        if (lev != 0) {
            const double ax = cadv[0], ay = cadv[1], az = cadv[2];
            advrl = ax * fx + ay * fy + az * fz;
            rhol = crho;
            altl = g.Alt[lev - 1];
            pressurel = cpk;
        }
        double advr = advx * fx + advy * fy + advz * fz;
        advx += -advr * fx;
        advy += -advr * fy;
        advz += -advr * fz;
        const double altht = g.Alth[lev + 1], althl = g.Alth[lev], alt = g.Alt[lev];
        const double M0x = sM[b][0][T.pos[0]], M0y = sM[b][1][T.pos[0]], M0z = sM[b][2][T.pos[0]];
        if (!g.nonhydro) {
            advr = (XPOW(M0x, 2) + XPOW(M0y, 2) + XPOW(M0z, 2)) / (rho * (alt + A));
            if (lev != 0) advrl = (XPOW(cm[0], 2) + XPOW(cm[1], 2) + XPOW(cm[2], 2)) / (rhol * (altl + A));
        }
        const double dz = altht - althl;
        const double r2p = slv[lev], r2m = slv[nvl + lev], r2l = slv[2 * nvl + lev], rscale = slv[3 * nvl + lev],
                     dr2dz = slv[4 * nvl + lev];
        // nflxv, nflxr, nflxp as slow_point; dptk: the stage part of the
        // pressure flux divergence of the substeps (prepare_point_w's st with
        // the stage momentum), formed here once per stage. The stencil slots
        // are walked once for all three components (each sum still adds its
        // slots in order 0..6).
        const double p0 = sP[b][T.pos[0]];
        double sg[3], sr[3], sp3[3], st3[3];
        {
            const double h0 = sH[b][T.pos[0]], t0 = sT[b][T.pos[0]];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double m0 = sM[b][k][T.pos[0]];
                sg[k] = wg(k) * p0;
                sr[k] = wd0(k) * m0;
                sp3[k] = wd0(k) * m0 * h0;
                st3[k] = wd0(k) * m0 * t0;
            }
        }
#pragma unroll
        for (int s = 1; s < 7; s++) {
            const double ps_ = sP[b][T.pos[s]], hs_ = sH[b][T.pos[s]], ts_ = sT[b][T.pos[s]];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double w = wg(s * 3 + k), m = sM[b][k][T.pos[s]];
                sg[k] = sg[k] + w * ps_;
                sr[k] = sr[k] + w * m;
                sp3[k] = sp3[k] + w * m * hs_;
                st3[k] = st3[k] + w * m * ts_;
            }
        }
        double nflxv[3], nflxr = 0.0, nflxp = 0.0, dptk = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            nflxv[k] = rscale * sg[k];
            nflxr += rscale * sr[k];
            nflxp += rscale * sp3[k];
            dptk += rscale * st3[k];
        }
        if constexpr (MX) LV(f.hflx_ptk, lev) = dptk;
        double whl, wht, hhl, hht;
        if (lev == 0) {
            whl = 0.0;
            wht = v.whk1;
            hhl = 0.0;
            hht = v.hh1;
        } else {
            whl = cwh;
            wht = v.whk1;
            hhl = chh;
            hht = v.hh1;
        }
        const double dwdz = (wht * r2p - whl * r2l) / (dr2dz);
        const double dwhdz = (wht * r2p * hht - whl * r2l * hhl) / (dz * r2m);
        const double dpr = nflxv[0] * fx + nflxv[1] * fy + nflxv[2] * fz;
        const double fv[3] = {fx, fy, fz}, ad[3] = {advx, advy, advz};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double DivM = v.dM[k];
            // the vertical momentum hyperdiffusion has its own array in the FP64
            // form only (the mixed form adds it to diffmh)
            double dmv = 0.0;
            if constexpr (!MX) dmv = f.diffmv ? VEC(f.diffmv, k, lev, p) : 0.0;
            const double o = -(nflxv[k] - dpr * fv[k]) - ad[k] + DivM + v.dmh[k] + dmv + v.dmv[k];
            if constexpr (MX) stb_f(rSm, T.o4, (unsigned)((k * nv + lev) * plR), (float)o);
            else VEC(SlowMh, k, lev, p) = o;
        }
        if (lev == 0) {
            LV(f.SlowWh, 0) = 0.0;
            LV(f.SlowWh, nv) = 0.0;
        } else {
            const double xi = althl, xim = altl, xip = alt;
            const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            double swr;
            if (g.deep) {
                if (g.ghv)
                    swr = (advrl + rhol * Gravit * slv[5 * nvl + lev - 1]) * intt
                          + (advr + rho * Gravit * slv[5 * nvl + lev]) * intl;
                else
                    swr = (advrl + rhol * Gravit) * intt + (advr + rho * Gravit) * intl;
            } else {
                if (g.ghv)
                    swr = (rhol * Gravit * slv[5 * nvl + lev - 1]) * intt + (rho * Gravit * slv[5 * nvl + lev]) * intl;
                else
                    swr = (rhol * Gravit) * intt + (rho * Gravit) * intl;
            }
            if (g.nonhydro) swr += (-cdw - 0.0) * intt + (-v.dw - 0.0) * intl;
            const double dzl = alt - altl;
            const double dpdz = (p0 - pressurel) / dzl;
            swr += dpdz;
            swr -= v_dwh;
            LV(f.SlowWh, lev) = -swr;
        }
        nflxr += dwdz;
        LV(f.SlowRho, lev) = -nflxr + v_drh + 0.0;
        const double rr = 1.0 / rho;
        const double vgp = (nflxv[0] - dpr * fx) * M0x * rr + (nflxv[1] - dpr * fy) * M0y * rr
                           + (nflxv[2] - dpr * fz) * M0z * rr;
        double sp = (Rd / Cv) * (-nflxp - v_gtil - dwhdz + vgp + v_q);
        if (!MX && g.energy) sp += Rd / Cv * (v_dpr + 0.0);  // thor_slowmodes.h:375-377
        else sp += v_dpr + 0.0;
        LV(f.Slowp, lev) = sp;
        if (rk == 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(Mhs, k, lev, p) = 0.0;
            LV(f.Rhos, lev) = 0.0;
            LV(f.ps, lev) = 0.0;
            LV(f.Whs, lev) = 0.0;
            if (lev == nv - 1) LV(f.Whs, nv) = 0.0;
        } else {
            const unsigned l8 = (unsigned)(lev * pl8);
            const double Mx = ldb_d(rM, T.o8, (unsigned)((0 * nv + lev) * pl8)) - M0x,
                         My = ldb_d(rM, T.o8, (unsigned)((1 * nv + lev) * pl8)) - M0y,
                         Mz = ldb_d(rM, T.o8, (unsigned)((2 * nv + lev) * pl8)) - M0z;
            const double aux = Mx * fx + My * fy + Mz * fz;
            VEC(Mhs, 0, lev, p) = Mx - aux * fx;
            VEC(Mhs, 1, lev, p) = My - aux * fy;
            VEC(Mhs, 2, lev, p) = Mz - aux * fz;
            LV(f.Rhos, lev) = ldb_d(rR, T.o8, l8) - rho;
            LV(f.ps, lev) = ldb_d(rPr, T.o8, l8) - p0;
            LV(f.Whs, lev) = lev == 0 ? 0.0 : ldb_d(rW, T.o8, l8) - cwh;
            if (lev == nv - 1) LV(f.Whs, nv) = 0.0;
        }
        // carry this level's values to the next
#pragma unroll
        for (int k = 0; k < 3; k++) cadv[k] = v.adv[k];
        cm[0] = M0x;
        cm[1] = M0y;
        cm[2] = M0z;
        cdw = v.dw;
        crho = rho;
        cpk = p0;
        cwh = v.whk1;
        chh = v.hh1;
        if (lev + 1 < l1) put(Rs, b ^ 1);
        __syncthreads();
    };
    Stg S0;
    Pt P0, P1;
    load_pt(P0, l0);
    load_stg(S0, l0);
    put(S0, 0);
    __syncthreads();
    for (int lev = l0; lev < l1; lev += 2) {
        body(lev, S0, P0, P1);
        if (lev + 1 < l1) body(lev + 1, S0, P1, P0);
    }
}

// Per-level constants of k_slow_lc2, with k_slow_lc's expressions: [6][nv]
// r2p, r2m, r2l, rscale, dr2dz and (ghv) (A / (A + Alt))^2.
__global__ void k_slow_levels(Geo g, double *out) {
    const int lev = blockIdx.x * blockDim.x + threadIdx.x, nv = g.nv;
    if (lev >= nv) return;
    const double A = g.A;
    const double altht = g.Alth[lev + 1], althl = g.Alth[lev], alt = g.Alt[lev];
    if (g.deep) {
        out[lev] = XPOW(altht + A, 2.0);
        out[nv + lev] = XPOW(alt + A, 2.0);
        out[2 * nv + lev] = XPOW(althl + A, 2.0);
        out[3 * nv + lev] = A / (alt + A);
        out[4 * nv + lev] = (XPOW(altht + A, 3.0) - XPOW(althl + A, 3.0)) / 3;
    } else {
        out[lev] = out[nv + lev] = out[2 * nv + lev] = 1.0;
        out[3 * nv + lev] = 1.0;
        out[4 * nv + lev] = 1.0 * (altht - althl);
    }
    out[5 * nv + lev] = g.ghv ? XPOW(A / (A + alt), 2) : 0.0;
}

// Compute_Advec_Cori1 (thor_adv_cor.h:51-235) and the first Diffusion_Op pass
// (thor_diff.h:54-522) of the mixed-precision path, in FP32 over level
// chunks. Each staged point carries its momentum, the advecting velocity
// v = M / rho + (W / rho) r and the diffused variables rho, M / rho, W / rho
// and Rd T (the reference's variables 0..5), formed once per point from the
// FP64 stage values and rounded to FP32, so no quotient is recomputed per
// stencil slot. The horizontal advection is added to Adv (which holds the
// vertical part from k_column_prep).
__global__ __launch_bounds__(256, 3) void k_adv_diff1_lc(Geo g, Fields f, int hydiff, int lc) {
#pragma clang fp contract(fast)
    int r, l0, l1;
    if (!block_tile_lc(g, lc, r, l0, l1)) return;
    const size_t S = g.S;
    const int nv = g.nv;
    float *Adv = f.mAdv, *diff = f.mdiff;
    if (r == g.nr) {
        PoleCtx c;
        int lev;
        if (pole_ctx_lc(g, l0, l1, c, lev)) {
            auto vel = [&](int q, int k) {
                const double rho = LVP(f.Rhok, lev, q), w = LVP(f.Wk, lev, q);
                return VEC(f.Mhk, k, lev, q) / rho + (w / rho) * g.fr[(size_t)k * S + q];
            };
            adv_diff1_point<true>(
                g, f, c.p, lev, 5, hydiff, [&](int s, int k) { return VEC(f.Mhk, k, lev, c.gi(s)); },
                [&](int s, int k) { return vel(c.gi(s), k); }, [&](int s) { return LVP(f.Rhok, lev, c.gi(s)); },
                [&](int s) { return LVP(f.Wk, lev, c.gi(s)); }, [&](int s) { return LVP(f.T, lev, c.gi(s)); });
        }
        return;
    }
    __shared__ float sM[2][3][NHL2], sV[2][3][NHL2], sA[2][6][NHL2], sL[4][kLcMax];
    stage_lvc(g, l0, l1, sL);
    const LcThread T = lc_thread(g, r);
    const int p = T.p;
    float wd[21], wl[6];
#pragma unroll
    for (int i = 0; i < 21; i++) wd[i] = (float)g.div[(size_t)i * S + p];
#pragma unroll
    for (int s = 0; s < 6; s++) wl[s] = g.wlap[(size_t)(s + 1) * S + p];
    // unit radial vector of the point and of its halo point
    const int qh = T.hp >= 0 ? g.halo[r * NHALO + threadIdx.x] : -1;
    float ofr[3], hfr[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        ofr[k] = (float)g.fr[(size_t)k * S + p];
        hfr[k] = qh >= 0 ? (float)g.fr[(size_t)k * S + qh] : 0.f;
    }
    const double Rd = g.Rd;
    const size_t pl4 = S * 4, pl8 = S * 8;  // bytes of one FP32 / FP64 level plane
    const auto rAd = rsrc(Adv, 3 * nv * pl4), rDf = rsrc(diff, 6 * nv * pl4), rM = rsrc(f.Mhk, 3 * nv * pl8),
               rR = rsrc(f.Rhok, nv * pl8), rW = rsrc(f.Wk, nv * pl8), rT = rsrc(f.T, nv * pl8);
    struct In {
        float adv[3];
        double om[3], orho, ow, ot, hm[3], hrho, hw, ht;
    };
    auto load = [&](In &v, int lev) {
#pragma unroll
        for (int k = 0; k < 3; k++) v.adv[k] = ldb_f(rAd, T.o4, (unsigned)((k * nv + lev) * pl4));
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v.om[k] = ldb_d(rM, T.o8, (unsigned)((k * nv + lev) * pl8));
            v.hm[k] = ldb_d(rM, T.h8, (unsigned)((k * nv + lev) * pl8));
        }
        v.orho = ldb_d(rR, T.o8, (unsigned)(lev * pl8));
        v.hrho = ldb_d(rR, T.h8, (unsigned)(lev * pl8));
        v.ow = ldb_d(rW, T.o8, (unsigned)(lev * pl8));
        v.hw = ldb_d(rW, T.h8, (unsigned)(lev * pl8));
        v.ot = ldb_d(rT, T.o8, (unsigned)(lev * pl8));
        v.ht = ldb_d(rT, T.h8, (unsigned)(lev * pl8));
    };
    // one staged point: M, v, and the six diffused variables
    auto put1 = [&](int b, int pos, const double m[3], double rho, double w, double t, const float fr[3]) {
        const double wr = w / rho;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double u = m[k] / rho;
            sM[b][k][pos] = (float)m[k];
            sV[b][k][pos] = (float)(u + wr * fr[k]);
            sA[b][1 + k][pos] = (float)u;
        }
        sA[b][0][pos] = (float)rho;
        sA[b][4][pos] = (float)wr;
        sA[b][5][pos] = (float)(Rd * t);
    };
    auto put = [&](const In &v, int b) {
        put1(b, T.ir, v.om, v.orho, v.ow, v.ot, ofr);
        if (T.hp >= 0) {
            if (qh >= 0) {
                put1(b, T.hp, v.hm, v.hrho, v.hw, v.ht, hfr);
            } else {
#pragma unroll
                for (int k = 0; k < 3; k++) sM[b][k][T.hp] = sV[b][k][T.hp] = 0.f;
#pragma unroll
                for (int k = 0; k < 6; k++) sA[b][k][T.hp] = 0.f;
            }
        }
    };
    float adv[3];
    auto body = [&](int lev, In &R, In &L) {
#pragma unroll
        for (int k = 0; k < 3; k++) adv[k] = L.adv[k];
        load(L, min(lev + 2, l1 - 1));
        const int b = (lev - l0) & 1;
        const float rs = sL[0][lev - l0], rs2 = sL[1][lev - l0];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float sum = 0.f;
#pragma unroll
            for (int kk = 0; kk < 3; kk++) {
                float sk = wd[kk] * sV[b][c][T.pos[0]] * sM[b][kk][T.pos[0]];
#pragma unroll
                for (int s = 1; s < 7; s++) sk += wd[s * 3 + kk] * sV[b][c][T.pos[s]] * sM[b][kk][T.pos[s]];
                sum += rs * sk;
            }
            stb_f(rAd, T.o4, (unsigned)((c * nv + lev) * pl4), sum + adv[c]);
        }
        if (hydiff) {
#pragma unroll
            for (int var = 0; var < 6; var++) {
                const float a0 = sA[b][var][T.pos[0]];
                float sum = wl[0] * (sA[b][var][T.pos[1]] - a0);
#pragma unroll
                for (int s = 2; s < 7; s++) sum += wl[s - 1] * (sA[b][var][T.pos[s]] - a0);
                stb_f(rDf, T.o4, (unsigned)((var * nv + lev) * pl4), rs2 * sum);
            }
        }
        if (lev + 1 < l1) put(R, b ^ 1);
        __syncthreads();
    };
    In A, B;
    load(A, l0);
    load(B, min(l0 + 1, l1 - 1));
    put(A, 0);
    __syncthreads();
    for (int lev = l0; lev < l1; lev += 2) {
        body(lev, B, A);
        if (lev + 1 < l1) body(lev + 1, A, B);
    }
}

// Vertical_Eq (thor_vertical_int.h:52-563): per column, the tridiagonal solve
// for Whs at interfaces 1..nv-1 with the modify-in-place Thomas algorithm;
// cc/dd live in LDS.
// The tridiagonal matrix depends only on stage-start quantities (hh, gtilh)
// and the substep length, so its Thomas factorisation - the modified upper
// diagonal cc' and the pivots t = 1 / (bb - cc'_{k-1} aa) of the reference's
// modify-in-place algorithm - is computed once per RK stage and reused by every
// acoustic substep; the values are those the reference recomputes each substep.
struct Plane {
    double *b;
    size_t S;
    int p;
    __device__ __forceinline__ double &operator[](int lev) const { return b[(size_t)lev * S + p]; }
};

// Thomas kernels: a block of 256 threads serves CB columns. First every
// thread computes per-level terms (coefficients or right-hand sides) for a
// share of the levels of its column into LDS, all loads in flight at once;
// then one lane per column runs the serial recurrence out of LDS. The
// arithmetic of each term and of the recurrence is the reference's.
// LDS: k_thomas_factor 4 planes of (nv + 1) x kThomasCB(nv) doubles;
// k_thomas_dp one plane of (nv + 1) x CB doubles.
template <int CB, bool MX>
__global__ __launch_bounds__(256) void k_thomas_factor(Geo g, Fields f, double deltat) {
    extern __shared__ double sm[];
    const int nv = g.nv, t = threadIdx.x, c = t % CB, row = t / CB;
    constexpr int NROW = 256 / CB;
    const int p = blockIdx.x * CB + c;
    const size_t S = g.S;
    double *sA = sm, *sB = sm + (nv + 1) * CB, *sC = sm + 2 * (nv + 1) * CB;
    const double A = g.A, Gravit = g.Gravit, Rd = g.Rd, Cp = g.Cp;
    const double *Alt = g.Alt, *Alth = g.Alth;
    const bool ghv = g.ghv, nonh = g.nonhydro;
    const double CRdd_c = (Cp - Rd) / (Rd * deltat * deltat);
    auto gcor = [&](double altx) {
        return ghv ? Gravit * XPOW(A / (A + altx), 2) * (Cp - Rd) / Rd : Gravit * (Cp - Rd) / Rd;
    };
    if (p < g.P) {
        for (int lev = 1 + row; lev < nv; lev += NROW) {
            const double altht = Alth[lev + 1], alth = Alth[lev], althl = Alth[lev - 1];
            const double alt = Alt[lev], altl = Alt[lev - 1];
            double hp = LV(f.hh, lev + 1), h = LV(f.hh, lev), hm = LV(f.hh, lev - 1);
            const double gp = LV(f.gtilh, lev + 1), gg = LV(f.gtilh, lev), gm = LV(f.gtilh, lev - 1);
            const double GCoRl = gcor(altl), GCoRu = gcor(alt);
            const double dzp = 1.0 / (altht - alth), dzh = 1.0 / (alt - altl), dzm = 1.0 / (alth - althl);
            const double dzph = dzp * dzh, dzmh = dzm * dzh;
            const double CRdd = (CRdd_c * (alt - alth) + CRdd_c * (alth - altl)) / (alt - altl);
            const double GCoR = (GCoRl * (alt - alth) + GCoRu * (alth - altl)) / (alt - altl);
            const double intl = (alt - alth) * dzp * dzh;
            const double inttm = -(altl - alth) * dzm * dzh;
            double aa, bb, cc;
            if (g.deep) {
                hp = hp * XPOW(altht + A, 2);
                h = h * XPOW(alth + A, 2);
                hm = hm * XPOW(althl + A, 2);
                const double or2 = 1 / XPOW(alth + A, 2);
                const double tor3 = 2 / XPOW(alth + A, 3);
                cc = -dzph * or2 * hp - inttm * (gp + GCoR - tor3 * hp);
                if (nonh) bb = CRdd + (dzph + dzmh) * or2 * h + (intl - inttm) * (gg + GCoR - tor3 * h);
                else bb = (dzph + dzmh) * or2 * h + (intl - inttm) * (gg + GCoR - tor3 * h);
                aa = -dzmh * or2 * hm + intl * (gm + GCoR - tor3 * hm);
            } else {
                cc = -dzph * hp - inttm * (gp + GCoR);
                if (nonh) bb = CRdd + (dzph + dzmh) * h + (intl - inttm) * (gg + GCoR);
                else bb = (dzph + dzmh) * h + (intl - inttm) * (gg + GCoR);
                aa = -dzmh * hm + intl * (gm + GCoR);
            }
            sA[lev * CB + c] = aa;
            sB[lev * CB + c] = bb;
            sC[lev * CB + c] = cc;
        }
    }
    __syncthreads();
    if (row != 0 || p >= g.P) return;
    const Plane AA{f.tAA, S, p}, CC{f.tCC, S, p}, TT{f.tTT, S, p};
    double ccp = 0.0;
    for (int lev = 1; lev < nv; lev++) {
        const double aa = sA[lev * CB + c], bb = sB[lev * CB + c], cc = sC[lev * CB + c];
        if (lev == 1) {
            ccp = cc / bb;
            TT[1] = bb;
        } else {
            const double tt = 1.0 / (bb - ccp * aa);
            ccp = cc * tt;
            TT[lev] = tt;
            AA[lev] = aa;
        }
        CC[lev] = ccp;
    }
}

// Mixed path, segment form of k_thomas_factor: row r of the block owns the
// contiguous interfaces [lo, hi) of its column (as in thomas_sweeps_rows),
// forms their aa, bb, cc in registers with k_thomas_factor's arithmetic, and
// the recurrence of the modified upper diagonal,
//   c'_1 = cc_1 / bb_1,  c'_l = cc_l / (bb_l - aa_l c'_{l-1}),
// a Moebius map of c'_{l-1} with matrix [[0, cc_l], [-aa_l, bb_l]], is
// composed over the segment (normalised after each product), the rows' maps
// are chained in order through LDS (sX: 4 x NROW x CB doubles), and each row
// then runs the recurrence over its segment from its true start value. The
// pivots and c' agree with the one-lane recurrence to FP64 rounding.
template <int CB, int SEG>
__global__ __launch_bounds__(256) void k_thomas_factor_seg(Geo g, Fields f, double deltat) {
    constexpr int NROW = 256 / CB;
    __shared__ double sX[4][NROW][CB];
    const int nv = g.nv, t = threadIdx.x, c = t % CB, row = t / CB;
    const int p = blockIdx.x * CB + c;
    const bool valid = p < g.P;
    const size_t S = g.S;
    const int seg = (nv - 1 + NROW - 1) / NROW;
    const int lo = 1 + row * seg, hi = min(nv, lo + seg);
    const double A = g.A, Gravit = g.Gravit, Rd = g.Rd, Cp = g.Cp;
    const double *Alt = g.Alt, *Alth = g.Alth;
    const bool ghv = g.ghv, nonh = g.nonhydro;
    const double CRdd_c = (Cp - Rd) / (Rd * deltat * deltat);
    auto gcor = [&](double altx) {
        return ghv ? Gravit * XPOW(A / (A + altx), 2) * (Cp - Rd) / Rd : Gravit * (Cp - Rd) / Rd;
    };
    double ra[SEG], rb[SEG], rc[SEG];
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        const int lev = lo + i;
        ra[i] = 0.0;
        rb[i] = 1.0;
        rc[i] = 0.0;
        if (valid && lev < hi) {
            const double altht = Alth[lev + 1], alth = Alth[lev], althl = Alth[lev - 1];
            const double alt = Alt[lev], altl = Alt[lev - 1];
            double hp = LV(f.hh, lev + 1), h = LV(f.hh, lev), hm = LV(f.hh, lev - 1);
            const double gp = LV(f.gtilh, lev + 1), gg = LV(f.gtilh, lev), gm = LV(f.gtilh, lev - 1);
            const double GCoRl = gcor(altl), GCoRu = gcor(alt);
            const double dzp = 1.0 / (altht - alth), dzh = 1.0 / (alt - altl), dzm = 1.0 / (alth - althl);
            const double dzph = dzp * dzh, dzmh = dzm * dzh;
            const double CRdd = (CRdd_c * (alt - alth) + CRdd_c * (alth - altl)) / (alt - altl);
            const double GCoR = (GCoRl * (alt - alth) + GCoRu * (alth - altl)) / (alt - altl);
            const double intl = (alt - alth) * dzp * dzh;
            const double inttm = -(altl - alth) * dzm * dzh;
            double aa, bb, cc;
            if (g.deep) {
                hp = hp * XPOW(altht + A, 2);
                h = h * XPOW(alth + A, 2);
                hm = hm * XPOW(althl + A, 2);
                const double or2 = 1 / XPOW(alth + A, 2);
                const double tor3 = 2 / XPOW(alth + A, 3);
                cc = -dzph * or2 * hp - inttm * (gp + GCoR - tor3 * hp);
                if (nonh) bb = CRdd + (dzph + dzmh) * or2 * h + (intl - inttm) * (gg + GCoR - tor3 * h);
                else bb = (dzph + dzmh) * or2 * h + (intl - inttm) * (gg + GCoR - tor3 * h);
                aa = -dzmh * or2 * hm + intl * (gm + GCoR - tor3 * hm);
            } else {
                cc = -dzph * hp - inttm * (gp + GCoR);
                if (nonh) bb = CRdd + (dzph + dzmh) * h + (intl - inttm) * (gg + GCoR);
                else bb = (dzph + dzmh) * h + (intl - inttm) * (gg + GCoR);
                aa = -dzmh * hm + intl * (gm + GCoR);
            }
            ra[i] = aa;
            rb[i] = bb;
            rc[i] = cc;
        }
    }
    // the segment's map x_{hi-1} = (m00 x + m01) / (m10 x + m11) of x_{lo-1}
    double m00 = 1.0, m01 = 0.0, m10 = 0.0, m11 = 1.0;
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        if (lo + i < hi) {
            const double a = lo + i == 1 ? 0.0 : -ra[i];  // c'_0 does not enter the first interface
            // [[0, cc], [a, bb]] x M
            const double n00 = rc[i] * m10, n01 = rc[i] * m11;
            const double n10 = a * m00 + rb[i] * m10, n11 = a * m01 + rb[i] * m11;
            const double sc = 1.0 / fmax(fmax(fabs(n00), fabs(n01)), fmax(fabs(n10), fabs(n11)));
            m00 = n00 * sc;
            m01 = n01 * sc;
            m10 = n10 * sc;
            m11 = n11 * sc;
        }
    }
    sX[0][row][c] = m00;
    sX[1][row][c] = m01;
    sX[2][row][c] = m10;
    sX[3][row][c] = m11;
    __syncthreads();
    double x = 0.0;  // c' at interface lo - 1
    for (int r = 0; r < row; r++) x = (sX[0][r][c] * x + sX[1][r][c]) / (sX[2][r][c] * x + sX[3][r][c]);
    if (!valid) return;
    const Plane AA{f.tAA, S, p}, CC{f.tCC, S, p}, TT{f.tTT, S, p};
#pragma unroll
    for (int i = 0; i < SEG; i++) {
        const int lev = lo + i;
        if (lev < hi) {
            if (lev == 1) {
                x = rc[i] / rb[i];
                TT[1] = rb[i];
            } else {
                const double tt = 1.0 / (rb[i] - x * ra[i]);
                x = rc[i] * tt;
                TT[lev] = tt;
                AA[lev] = ra[i];
            }
            CC[lev] = x;
        }
    }
}

// ===================================================================== K_G2
// Density_Pressure_Eqs / _Poles (thor_fastmodes.h:255-847, entropy form),
// column part: the horizontal flux divergences nflxr = rscale sum div M' and
// nflxpt = rscale sum div (Mk + M') pt come from k_prepare (same sums, same
// order), the vertical fluxes from the freshly solved Whs of the column
// (whs(k) = Whs at interface k).
// rhos, pss, slowrho: Rhos, ps and SlowRho of the cell, read by the caller
// (the segment kernel keeps them from its right-hand-side phase).
template <bool MX, class GetW>
__device__ __forceinline__ void dp_column_in(const Geo &g, const Fields &f, int p, int lev, bool pole, double dt,
                                             double nflxr, double nflxpt, GetW whs, double rhos, double pss,
                                             double slowrho) {
    const size_t S = g.S;
    const double A = g.A, P_Ref = g.P_Ref, Rd = g.Rd, Cp = g.Cp, Cv = Cp - Rd;
    const double altht = g.Alth[lev + 1], althl = g.Alth[lev], dz = altht - althl;
    double r2p, r2m, r2l, dr2dz;
    if (g.deep) {
        const double alt = g.Alt[lev];
        r2p = XPOW(altht + A, 2.0);
        r2m = XPOW(alt + A, 2.0);
        r2l = XPOW(althl + A, 2.0);
        dr2dz = (XPOW(altht + A, 3.0) - XPOW(althl + A, 3.0)) / 3.0;
    } else {
        r2p = r2m = r2l = 1.0;
        dr2dz = r2m * dz;
    }
    double whl, wht, whl2, wht2, phl, pht;
    if (lev == 0) {
        whl = 0.0;
        wht = whs(lev + 1);
        whl2 = 0.0;
        wht2 = whs(lev + 1) + LV(f.Whk, lev + 1);
        phl = 0.0;
        pht = LV(f.pth, lev + 1);
    } else {
        whl = whs(lev);
        wht = whs(lev + 1);
        whl2 = whs(lev) + LV(f.Whk, lev);
        wht2 = whs(lev + 1) + LV(f.Whk, lev + 1);
        phl = LV(f.pth, lev);
        pht = LV(f.pth, lev + 1);
    }
    const double dwdz = pole ? (wht * r2p - whl * r2l) / (dz * r2m) : (wht * r2p - whl * r2l) / (dr2dz);
    const double dwptdz = (wht2 * pht * r2p - whl2 * phl * r2l) / (dz * r2m);
    double aux = -(nflxpt + dwptdz) * dt;
    const double pkv = LV(f.pk, lev);
    const double r = LV(f.Rhok, lev) + rhos;
    nflxr += dwdz;
    const double rhosn = rhos + (slowrho - nflxr) * dt;
    LV(f.Rhos, lev) = rhosn;
    double psn;
    if constexpr (!MX) {
        if (g.energy) {
            // total energy (thor_fastmodes.h:501-533): advance it, then the
            // pressure from it less the potential and kinetic energy of the
            // new density, momentum (Mhs + Mhk at the point) and vertical
            // momentum (the new Whs + Whk interpolated to the centre)
            const double *Alt = g.Alt, *Alth = g.Alth, Gravit = g.Gravit;
            const int nv = g.nv;
            aux += LV(f.Etot, lev) + (LV(f.diffpr, lev) + 0.0) * dt + LV(f.Qheat, lev) * dt;
            LV(f.Etot, lev) = aux;
            const double w =
                (whl2 * (Alth[lev + 1] - Alt[lev]) + wht2 * (Alt[lev] - Alth[lev])) / (Alth[lev + 1] - Alth[lev]);
            const double v0 = VEC(f.Mhs, 0, lev, p) + VEC(f.Mhk, 0, lev, p),
                         v1 = VEC(f.Mhs, 1, lev, p) + VEC(f.Mhk, 1, lev, p),
                         v2 = VEC(f.Mhs, 2, lev, p) + VEC(f.Mhk, 2, lev, p);
            const double rn = rhosn + LV(f.Rhok, lev);
            double epk = g.ghv ? rn * Gravit * XPOW(A / (A + Alt[lev]), 2) * Alt[lev] : rn * Gravit * Alt[lev];
            epk = epk + 0.5 * (XPOW(v0, 2) + XPOW(v1, 2) + XPOW(v2, 2) + XPOW(w, 2)) / rn;
            psn = Rd / Cv * (aux - epk) - pkv;
            LV(f.ps, lev) = psn;
            if (isnan(psn)) atomicOr(f.flags, FLAG_NAN_P);
            return;
        }
    }
    if constexpr (MX) {
        // The same update with one power. With pf = pss + pkv, q = (pf / P_Ref)^(Cv/Cp)
        // and x = -Rd (nflxpt + dwptdz) dt / (P_Ref q): Rd aux / P_Ref = q (1 + x),
        // so pnew = pf (1 + x)^(Cp/Cv) and pnew - pkv = pss + pf ((1 + x)^(Cp/Cv) - 1),
        // formed without the cancellation of pnew - pkv. For |x| < 2^-7 the
        // binomial series (terms to x^9) gives (1 + x)^gamma - 1.
        const double pf = pss + pkv, gam = Cp / Cv;
        const double q = pm::pow(pf / P_Ref, Cv / Cp);
        const double x = aux * Rd / (P_Ref * q);
        double g1;
        if (fabs(x) < 0.0078125) {
            double c[10];
            c[1] = gam;
#pragma unroll
            for (int k = 2; k <= 9; k++) c[k] = c[k - 1] * (gam - (k - 1)) / k;
            double h = c[9];
#pragma unroll
            for (int k = 8; k >= 1; k--) h = h * x + c[k];
            g1 = h * x;
        } else {
            g1 = pm::pow(1.0 + x, gam) - 1.0;
        }
        psn = pss + pf * g1 + (LV(FS<MX>::diffpr(f), lev) + 0.0) * dt + Rd / Cv * LV(f.Qheat, lev) * dt;
    } else {
        const double ptc = (P_Ref / (Rd * r)) * XPOW((pss + pkv) / P_Ref, Cv / Cp);
        aux += ptc * r;
        const double pnew = P_Ref * XPOW(Rd * aux / P_Ref, Cp / Cv);
        psn = pnew - pkv + (LV(FS<MX>::diffpr(f), lev) + 0.0) * dt + Rd / Cv * LV(f.Qheat, lev) * dt;
    }
    LV(f.ps, lev) = psn;
    if (isnan(psn)) atomicOr(f.flags, FLAG_NAN_P);
}

template <bool MX, class GetW>
__device__ __forceinline__ void dp_column(const Geo &g, const Fields &f, int p, int lev, bool pole, double dt,
                                          double nflxr, double nflxpt, GetW whs) {
    const size_t S = g.S;
    dp_column_in<MX>(g, f, p, lev, pole, dt, nflxr, nflxpt, whs, LV(f.Rhos, lev), LV(f.ps, lev),
                     LV(f.SlowRho, lev));
}

// Vertical_Eq (thor_vertical_int.h:52-563) with the stage's factorisation
// (right-hand side C0 from the substep's sources, forward elimination into
// Whs as the reference's dd, back substitution Whs_k = -cc'_k Whs_{k+1} + dd_k,
// Whs = 0 at the bottom and top interfaces), followed by the column part of
// the substep's Density_Pressure_Eqs on the new Whs, which the block keeps in
// LDS: all threads take a share of the levels again after the solve.
// DP = false: the vertical solve alone (Whs written, density/pressure left to
// k_dp), for a physics hook that runs between the two as in the reference.
// Mixed path: both sweeps of the vertical solve on all rows of the block, each
// row on a contiguous segment of at most kSeg interfaces of its column. The
// forward elimination x_l = (-C_l - x_{l-1} AA_l) TT_l and the back
// substitution w_l = -CC_l w_{l+1} + x_l are first-order linear recurrences:
// each row composes its segment into one affine map, the rows' maps are
// chained in order through LDS (sX: 2 x NROW x CB doubles), and each row then
// runs its segment from its true starting value with the recurrence's own
// arithmetic. The loads of the factorisation are spread over all rows and the
// serial chain is a quarter of the one-lane sweep's. All threads call it.
constexpr int kSeg = 12;
template <int CB, int KS = kSeg>
__device__ __forceinline__ void thomas_sweeps_rows(double *sD, double *sX, const Plane &AA, const Plane &CC,
                                                   const Plane &TT, int nv, int c, int row, bool valid) {
    constexpr int NROW = 256 / CB;
    const int seg = (nv - 1 + NROW - 1) / NROW;
    const int lo = 1 + row * seg, hi = min(nv, lo + seg);  // this row's interfaces [lo, hi)
    double *sa = sX, *sb = sX + NROW * CB;
    // forward elimination: the segment's map x_{hi-1} = a x_{lo-1} + b; the
    // back substitution's coefficients are loaded here too, so both sweeps
    // wait for memory once
    double a = 1.0, b = 0.0;
    double ccl[KS];
    {
        double al[KS], tl[KS], cl[KS];
#pragma unroll
        for (int i = 0; i < KS; i++) {
            const int l = lo + i;
            const bool in = valid && l < hi;
            al[i] = in ? AA[l] : 0.0;
            tl[i] = in ? TT[l] : 1.0;
            ccl[i] = in ? CC[l] : 0.0;
            cl[i] = in ? sD[l * CB + c] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < KS; i++) {
            const int l = lo + i;
            if (l < hi) {
                if (l == 1) {
                    a = 0.0;
                    b = (-cl[i]) / tl[i];
                } else {
                    const double m = -al[i] * tl[i];
                    a = m * a;
                    b = m * b + (-cl[i]) * tl[i];
                }
            }
        }
    }
    sa[row * CB + c] = a;
    sb[row * CB + c] = b;
    __syncthreads();
    double x = 0.0;  // x at interface lo - 1
    for (int r = 0; r < row; r++) x = sa[r * CB + c] * x + sb[r * CB + c];
    {
        double al[KS], tl[KS];
#pragma unroll
        for (int i = 0; i < KS; i++) {
            const int l = lo + i;
            const bool in = valid && l < hi;
            al[i] = in ? AA[l] : 0.0;
            tl[i] = in ? TT[l] : 1.0;
        }
#pragma unroll
        for (int i = 0; i < KS; i++) {
            const int l = lo + i;
            if (l < hi) {
                x = l == 1 ? (-sD[CB + c]) / tl[i] : (-sD[l * CB + c] - x * al[i]) * tl[i];
                sD[l * CB + c] = x;
            }
        }
    }
    __syncthreads();
    // back substitution: the segment's map w_lo = e w_hi + f
    double e = 1.0, fo = 0.0;
    {
#pragma unroll
        for (int i = KS - 1; i >= 0; i--) {
            const int l = lo + i;
            if (l < hi) {
                e = -ccl[i] * e;
                fo = -ccl[i] * fo + sD[l * CB + c];
            }
        }
        sa[row * CB + c] = e;
        sb[row * CB + c] = fo;
        __syncthreads();
        double w = 0.0;  // w at interface hi (0 at the top, interface nv)
        for (int r = NROW - 1; r > row; r--) w = sa[r * CB + c] * w + sb[r * CB + c];
#pragma unroll
        for (int i = KS - 1; i >= 0; i--) {
            const int l = lo + i;
            if (l < hi) {
                w = (-ccl[i] * w + sD[l * CB + c]);
                sD[l * CB + c] = w;
            }
        }
    }
}

// Iterations of the serial sweeps unrolled together: their factorisation loads
// do not depend on the recurrence and are issued ahead of it.
#ifndef THOR_SWEEP_UNROLL
#define THOR_SWEEP_UNROLL 8
#endif
template <int CB, bool MX, bool DP = true, int UR = 1>
__global__ __launch_bounds__(256) void k_thomas_dp(Geo g, Fields f, double deltat) {
    extern __shared__ double sm[];
    const int nv = g.nv, t = threadIdx.x, c = t % CB, row = t / CB;
    constexpr int NROW = 256 / CB;
    const int p = blockIdx.x * CB + c;
    const size_t S = g.S;
    double *sD = sm;
    const double A = g.A, Gravit = g.Gravit, Rd = g.Rd, Cp = g.Cp;
    const double *Alt = g.Alt, *Alth = g.Alth;
    const bool ghv = g.ghv, nonh = g.nonhydro;
    const double CRdd_c = (Cp - Rd) / (Rd * deltat * deltat);
    const double dt2 = XPOW(deltat, 2.0);
    if (p < g.P) {
#pragma unroll UR
        for (int lev = 1 + row; lev < nv; lev += NROW) {
            const double alth = Alth[lev], alt = Alt[lev], altl = Alt[lev - 1];
            // mixed form: Sd = -div(Mhs) + SlowRho from the flux divergence the
            // density update reads anyway (the same sum, opposite sign: equal bit
            // for bit), so the prepare kernel writes no Sd plane
            auto sd_at = [&](int l) {
                if constexpr (MX) return -LV(f.hflx_r, l) + LV(f.SlowRho, l);
                else return LV(f.Sd, l);
            };
            const double Sp = LV(f.Sp, lev), Sd = sd_at(lev), pr = LV(f.ps, lev), r = LV(f.Rhos, lev);
            const double Spl = LV(f.Sp, lev - 1), Sdl = sd_at(lev - 1), pl = LV(f.ps, lev - 1),
                         rl = LV(f.Rhos, lev - 1);
            const double dzh = 1.0 / (alt - altl);
            const double CRdd = (CRdd_c * (alt - alth) + CRdd_c * (alth - altl)) / (alt - altl);
            const double dSpdz = (Sp - Spl) * dzh, dPdz = (pr - pl) * dzh;
            const double intt = (alth - alt) / (altl - alt), intl = (alth - altl) / (alt - altl);
            const double rhohs = rl * intt + r * intl;
            const double Sdh = Sdl * intt + Sd * intl;
            const double srh = LV(f.SlowWh, lev);
            double C0;
            if (!nonh) {
                if (ghv)
                    C0 = (dt2 * dSpdz + dt2 * Gravit * XPOW(A / (A + alt), 2) * Sdh
                          + Gravit * XPOW(A / (A + alt), 2) * deltat * rhohs + deltat * dPdz - deltat * srh)
                         * (CRdd);
                else
                    C0 = (dt2 * dSpdz + dt2 * Gravit * Sdh + Gravit * deltat * rhohs + deltat * dPdz
                          - deltat * srh)
                         * (CRdd);
            } else {
                const double whs = LV(f.Whs, lev);
                if (ghv)
                    C0 = (-whs + dt2 * dSpdz + dt2 * Gravit * XPOW(A / (A + alt), 2) * Sdh
                          + Gravit * XPOW(A / (A + alt), 2) * deltat * rhohs + deltat * dPdz - deltat * srh)
                         * (CRdd);
                else
                    C0 = (-whs + dt2 * dSpdz + dt2 * Gravit * Sdh + Gravit * deltat * rhohs + deltat * dPdz
                          - deltat * srh)
                         * (CRdd);
            }
            sD[lev * CB + c] = C0;
        }
    }
    __syncthreads();
    // One lane per column: forward sweep and back substitution out of LDS,
    // with the stage's factorisation read straight from HBM (its loads do not
    // depend on the recurrence, so they are issued ahead of it). Staging the
    // factorisation in LDS instead measured no faster at g6 L48.
    const Plane AA{f.tAA, S, p}, CC{f.tCC, S, p}, TT{f.tTT, S, p};
    bool rows_done = false;
    if constexpr (MX && NROW * kSeg >= 8) {
        if (f.thomas_rows && nv > 1 && nv - 1 <= NROW * kSeg) {
            thomas_sweeps_rows<CB>(sD, sm + (size_t)(nv + 1) * CB, AA, CC, TT, nv, c, row, p < g.P);
            if (row == 0) {
                sD[c] = 0.0;
                sD[nv * CB + c] = 0.0;
            }
            rows_done = true;
        }
    }
    if (!rows_done && row == 0 && p < g.P && nv > 1) {
        double ddp = (-sD[CB + c]) / TT[1];
        sD[CB + c] = ddp;
#pragma unroll THOR_SWEEP_UNROLL
        for (int lev = 2; lev < nv; lev++) {
            ddp = (-sD[lev * CB + c] - ddp * AA[lev]) * TT[lev];
            sD[lev * CB + c] = ddp;
        }
    }
    if (!rows_done && row == 0 && p < g.P) {
        double wup = 0.0;
#pragma unroll THOR_SWEEP_UNROLL
        for (int lev = nv - 1; lev > 0; lev--) {
            const double w = (-CC[lev] * wup + sD[lev * CB + c]);
            sD[lev * CB + c] = w;
            wup = w;
        }
        sD[c] = 0.0;
        sD[nv * CB + c] = 0.0;
    }
    __syncthreads();
    if (p >= g.P) return;
    {
        const Plane W{f.Whs, S, p};
        for (int lev = row; lev <= nv; lev += NROW) W[lev] = sD[lev * CB + c];
    }
    if (!DP) return;
    const bool pole = p >= g.P - 2;
#pragma unroll UR
    for (int lev = row; lev < nv; lev += NROW)
        dp_column<MX>(g, f, p, lev, pole, deltat, LV(f.hflx_r, lev), LV(f.hflx_pt, lev),
                      [&](int k) { return sD[k * CB + c]; });
}

// Segment form of k_thomas_dp: each row of the block owns a
// contiguous segment of interfaces [lo, hi) of its column in all three phases
// (the segments of thomas_sweeps_rows) and the cells lo - 1 .. hi - 1 under and
// above them. The right-hand side of interface l reads the cells l - 1 and l,
// all in the row's segment, and the cells' Rhos, ps, SlowRho and hflx_r stay
// in registers for the density/pressure update, which the row then applies to
// the cells lo - 1 .. hi - 2 (and nv - 1 on the top row): no cell value is
// read twice from HBM. C0, the sweeps and the update are k_thomas_dp's
// arithmetic: the mixed form (MX) runs both sweeps on all rows
// (thomas_sweeps_rows) and equals k_thomas_dp<CB, true> with the same rows;
// the FP64 form runs them on one lane per column in the reference's order and
// equals k_thomas_dp<.., false> bit for bit.
// SEG: interfaces per row at most; the host launches it with
// (nv - 1 + NROW - 1) / NROW <= SEG. DP = false: the vertical solve alone
// (Whs written; k_dp applies the update after a physics hook). Up to three
// interfaces per row at 16 columns: five waves per SIMD (96 VGPRs, no spills;
// 79 against 83 us at g6 L48 with four, and six spill); longer rows spill at
// five and keep the compiler's choice.
template <int CB, int SEG, bool DP = true, bool MX = true>
__global__ __launch_bounds__(256, (MX && CB == 16 && SEG <= 3) ? 5 : 1) void k_thomas_dp_seg(Geo g, Fields f,
                                                                                             double deltat) {
    constexpr int NROW = 256 / CB;
    extern __shared__ double sm[];
    double *sD = sm;
    const int nv = g.nv, t = threadIdx.x, c = t % CB, row = t / CB;
    const int p = blockIdx.x * CB + c;
    const bool valid = p < g.P;
    const size_t S = g.S;
    const int seg = (nv - 1 + NROW - 1) / NROW;
    const int lo = 1 + row * seg, hi = min(nv, lo + seg);
    const int ncell = hi >= lo ? hi - lo + 1 : 0;  // cells lo - 1 .. hi - 1
    const double A = g.A, Gravit = g.Gravit, Rd = g.Rd, Cp = g.Cp;
    const double *Alt = g.Alt, *Alth = g.Alth;
    const bool ghv = g.ghv, nonh = g.nonhydro;
    const double CRdd_c = (Cp - Rd) / (Rd * deltat * deltat);
    const double dt2 = XPOW(deltat, 2.0);
    double cR[SEG + 1], cP[SEG + 1], cS[SEG + 1], cF[SEG + 1], cSp[SEG + 1], cD[SEG + 1];
#pragma unroll
    for (int i = 0; i <= SEG; i++) {
        const int lev = lo - 1 + i;
        const bool in = valid && i < ncell;
        cR[i] = in ? LV(f.Rhos, lev) : 0.0;
        cP[i] = in ? LV(f.ps, lev) : 0.0;
        cS[i] = in ? LV(f.SlowRho, lev) : 0.0;
        cF[i] = in ? LV(f.hflx_r, lev) : 0.0;
        cSp[i] = in ? LV(f.Sp, lev) : 0.0;
        if constexpr (!MX) cD[i] = in ? LV(f.Sd, lev) : 0.0;  // the FP64 prepare writes Sd
    }
#pragma unroll
    for (int i = 1; i <= SEG; i++) {
        const int lev = lo - 1 + i;  // interface
        if (valid && i < ncell) {
            const double alth = Alth[lev], alt = Alt[lev], altl = Alt[lev - 1];
            const double Sp = cSp[i], Sd = MX ? -cF[i] + cS[i] : cD[i], pr = cP[i], r = cR[i];
            const double Spl = cSp[i - 1], Sdl = MX ? -cF[i - 1] + cS[i - 1] : cD[i - 1], pl = cP[i - 1],
                         rl = cR[i - 1];
            const double dzh = 1.0 / (alt - altl);
            const double CRdd = (CRdd_c * (alt - alth) + CRdd_c * (alth - altl)) / (alt - altl);
            const double dSpdz = (Sp - Spl) * dzh, dPdz = (pr - pl) * dzh;
            const double intt = (alth - alt) / (altl - alt), intl = (alth - altl) / (alt - altl);
            const double rhohs = rl * intt + r * intl;
            const double Sdh = Sdl * intt + Sd * intl;
            const double srh = LV(f.SlowWh, lev);
            double C0;
            if (!nonh) {
                if (ghv)
                    C0 = (dt2 * dSpdz + dt2 * Gravit * XPOW(A / (A + alt), 2) * Sdh
                          + Gravit * XPOW(A / (A + alt), 2) * deltat * rhohs + deltat * dPdz - deltat * srh)
                         * (CRdd);
                else
                    C0 = (dt2 * dSpdz + dt2 * Gravit * Sdh + Gravit * deltat * rhohs + deltat * dPdz
                          - deltat * srh)
                         * (CRdd);
            } else {
                const double whs = LV(f.Whs, lev);
                if (ghv)
                    C0 = (-whs + dt2 * dSpdz + dt2 * Gravit * XPOW(A / (A + alt), 2) * Sdh
                          + Gravit * XPOW(A / (A + alt), 2) * deltat * rhohs + deltat * dPdz - deltat * srh)
                         * (CRdd);
                else
                    C0 = (-whs + dt2 * dSpdz + dt2 * Gravit * Sdh + Gravit * deltat * rhohs + deltat * dPdz
                          - deltat * srh)
                         * (CRdd);
            }
            sD[lev * CB + c] = C0;
        }
    }
    __syncthreads();
    const Plane AA{f.tAA, S, p}, CC{f.tCC, S, p}, TT{f.tTT, S, p};
    if constexpr (MX) {
        thomas_sweeps_rows<CB, SEG>(sD, sm + (size_t)(nv + 1) * CB, AA, CC, TT, nv, c, row, valid);
    } else if (row == 0 && valid) {
        // one lane per column: forward elimination and back substitution in
        // the reference's order (the factorisation's loads ahead of the chain)
        double ddp = (-sD[CB + c]) / TT[1];
        sD[CB + c] = ddp;
#pragma unroll THOR_SWEEP_UNROLL
        for (int lev = 2; lev < nv; lev++) {
            ddp = (-sD[lev * CB + c] - ddp * AA[lev]) * TT[lev];
            sD[lev * CB + c] = ddp;
        }
        double wup = 0.0;
#pragma unroll THOR_SWEEP_UNROLL
        for (int lev = nv - 1; lev > 0; lev--) {
            const double w = (-CC[lev] * wup + sD[lev * CB + c]);
            sD[lev * CB + c] = w;
            wup = w;
        }
    }
    if (row == 0) {
        sD[c] = 0.0;
        sD[nv * CB + c] = 0.0;
    }
    __syncthreads();
    if (!valid) return;
    {
        const Plane W{f.Whs, S, p};
        for (int lev = row; lev <= nv; lev += NROW) W[lev] = sD[lev * CB + c];
    }
    if (!DP) return;
    const bool pole = p >= g.P - 2;
    const int nlast = hi == nv ? ncell : ncell - 1;  // the top row also updates cell nv - 1
#pragma unroll
    for (int i = 0; i < SEG + 1; i++) {
        const int lev = lo - 1 + i;
        if (i < nlast)
            dp_column_in<MX>(g, f, p, lev, pole, deltat, cF[i], LV(f.hflx_pt, lev),
                               [&](int k) { return sD[k * CB + c]; }, cR[i], cP[i], cS[i]);
    }
}

// Column part of Density_Pressure_Eqs after a separate vertical solve
// (k_thomas_dp<.., false>): one thread per (point, level), Whs from HBM.
template <bool MX> __global__ __launch_bounds__(256) void k_dp(Geo g, Fields f, double deltat) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x, lev = blockIdx.y;
    if (p >= g.P) return;
    const size_t S = g.S;
    dp_column<MX>(g, f, p, lev, p >= g.P - 2, deltat, LV(f.hflx_r, lev), LV(f.hflx_pt, lev),
                  [&](int k) { return LVP(f.Whs, k, p); });
}

// Physics-module arrays advanced by the RK scheme (UpdateRK / UpdateRK2,
// thor_auxiliary.h:376-387, 482-494), THOR layout [(p * nv + lev) * dims + i].
struct RKArrays {
    int n;
    double *a[10], *ak[10], *ai[10];
    long count[10];  // P * nv * dims
};
// mode 0 (start of RK stages 2, 3): a = ai - ak; mode 1 (end of a stage): ak += a.
__global__ __launch_bounds__(256) void k_rk_arrays(RKArrays r, int mode) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (int j = 0; j < r.n; j++) {
        if (i >= r.count[j]) continue;
        if (mode == 0) r.a[j][i] = r.ai[j][i] - r.ak[j][i];
        else r.ak[j][i] += r.a[j][i];
    }
}

// ===================================================================== K_H
// UpdateRK2 (thor_auxiliary.h:420-529): y(k) += deviation. Whk is written to
// a separate buffer so each level can form W at the layer centre from both of
// its interfaces without a race.
template <int LC, bool MX>
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
    const auto *Mhs = FS<MX>::Mhs(f);
    for (int lev = l0; lev < l1; lev++) {
        const double Mx = VEC(Mhs, 0, lev, p) + VEC(f.Mhk, 0, lev, p);
        const double My = VEC(Mhs, 1, lev, p) + VEC(f.Mhk, 1, lev, p);
        const double Mz = VEC(Mhs, 2, lev, p) + VEC(f.Mhk, 2, lev, p);
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
// Newtonian-relaxation parameters in the form the kernel uses (logarithms of
// the pressures and time scales).
struct NRPar {
    double T_mean, dT, lon_sub, lnp_top, lnp_bot, lnp_con, lntau_top, lntau_bot;
};

// Equilibrium temperature and relaxation time of the hot-Jupiter Newtonian
// forcing at pressure pre (see thor_newtonian_params).
__device__ __forceinline__ void nr_profile(const NRPar &q, double lon, double lat, double pre, double &Teq,
                                           double &tau) {
    const double lnp = MLOG(pre);
    double w = (q.lnp_bot - lnp) / (q.lnp_bot - q.lnp_con);
    w = fmin(1.0, fmax(0.0, w));
    Teq = q.T_mean + q.dT * w * MCOS(lon - q.lon_sub) * MCOS(lat);
    double fr = (lnp - q.lnp_top) / (q.lnp_bot - q.lnp_top);
    fr = fmin(1.0, fmax(0.0, fr));
    tau = MEXP(q.lntau_top + fr * (q.lntau_bot - q.lntau_top));
}

// ProfX with a temperature-relaxation forcing: Recompute_W
// (profx_auxiliary.h:110-130), T = p / (rho Rd), the forcing, p = rho Rd T
// (profx_auxiliary.h:95-108). F = THOR_FORCING_HELD_SUAREZ: held_suarez
// (profx_held_suarez.h:51-117); F = THOR_FORCING_NEWTONIAN: nr_profile.
// Block of 64 columns x 4 level rows; the surface pressure of each column is
// read before any level of it is overwritten (one barrier), then every
// (column, level) is independent.
// Boundary-layer module, RayleighHS type (boundary_layer.cu:441-490): an
// implicit Rayleigh drag on the momentum of the layers near the surface,
// kv = surf_drag max(0, (sigma - bl_sigma) / (1 - bl_sigma)) with sigma = p / ps,
// ps extrapolated from the two lowest layers as the reference does (with the
// gravity factor under GravHeightVar). Runs in ProfX after the forcing and
// before the new pressure (phy_modules_phy_loop, profx_driver.cu:403-416), so
// p is the pressure ProfX started from.
struct BLPar {
    int on;
    double surf_drag, bl_sigma;
};
__device__ __forceinline__ double bl_surface_pressure(const Geo &g, const Fields &f, int p) {
    const size_t S = g.S;
    const double *Alt = g.Alt, Gravit = g.Gravit, A = g.A;
    const double psm1 = g.ghv ? LV(f.p, 1) - LV(f.Rho, 0) * Gravit * XPOW(A / (A + Alt[0]), 2) * (-Alt[0] - Alt[1])
                              : LV(f.p, 1) - LV(f.Rho, 0) * Gravit * (-Alt[0] - Alt[1]);
    return 0.5 * (LV(f.p, 0) + psm1);
}
__device__ __forceinline__ void bl_drag(const Geo &g, const Fields &f, const BLPar &bl, int p, int lev, double pre,
                                        double ps, double time_step) {
    const size_t S = g.S;
    const int nv = g.nv;
    const double sigma = (pre / ps);
    const double kv_hs = bl.surf_drag * fmax(0.0, (sigma - bl.bl_sigma) / (1.0 - bl.bl_sigma));
#pragma unroll
    for (int k = 0; k < 3; k++) VEC(f.Mh, k, lev, p) = VEC(f.Mh, k, lev, p) / (1.0 + kv_hs * time_step);
}
// The module alone (forcings that leave Mh alone run after it: the same numbers).
__global__ __launch_bounds__(256) void k_bl_rayleigh(Geo g, Fields f, BLPar bl, double time_step) {
    constexpr int CB = 64, NROW = 4;
    const int c = threadIdx.x % CB, row = threadIdx.x / CB;
    const int p = blockIdx.x * CB + c;
    if (p >= g.P) return;
    const size_t S = g.S;
    const double ps = bl_surface_pressure(g, f, p);
    for (int lev = row; lev < g.nv; lev += NROW) bl_drag(g, f, bl, p, lev, LV(f.p, lev), ps, time_step);
}

template <int F>
__global__ __launch_bounds__(256) void k_profx_relax(Geo g, Fields f, double time_step, int zero_q, NRPar nr,
                                                     BLPar bl) {
    constexpr int CB = 64, NROW = 4;
    const int c = threadIdx.x % CB, row = threadIdx.x / CB;
    const int p = blockIdx.x * CB + c;
    const bool valid = p < g.P;
    const size_t S = g.S;
    const int nv = g.nv;
    const double *Alt = g.Alt, *Alth = g.Alth;
    const double Rd = g.Rd, Cp = g.Cp, Gravit = g.Gravit;
    const double kappa = Rd / Cp;
    const double sigmab = 0.7, ka = (1.0 / 40.0) * (1.0 / 86400.0), kf = 1.0 / 86400.0,
                 ks = (1.0 / 4.0) * (1.0 / 86400.0), dTy = 60.0, dthetaz = 10.0, p0 = 100000.0;
    double ps = 0.0, lat = 0.0, lon = 0.0, ps_bl = 0.0;
    if (valid) {
        lon = g.lonlat[p];
        lat = g.lonlat[S + p];
        const double psm1 = LV(f.p, 1) - LV(f.Rho, 0) * Gravit * (-Alt[0] - Alt[1]);
        ps = 0.5 * (LV(f.p, 0) + psm1);
        if (bl.on) ps_bl = bl_surface_pressure(g, f, p);
    }
    __syncthreads();
    if (!valid) return;
    bool nan_t = false;
    for (int lev = row; lev < nv; lev += NROW) {
        {
            const double whl = LV(f.Wh, lev), wht = LV(f.Wh, lev + 1);
            const double xi = Alt[lev], xim = Alth[lev], xip = Alth[lev + 1];
            const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            if (!f.skip_w) LV(f.W, lev) = whl * intt + wht * intl;
        }
        const double pre = LV(f.p, lev), rho = LV(f.Rho, lev);
        double T = pre / (Rd * rho);
        if constexpr (F == THOR_FORCING_HELD_SUAREZ) {
            const double sigma = (pre / ps), sigma0 = (pre / p0);
            const double Teq = fmax(200.0, (315.0 - dTy * MPOW(MSIN(lat), 2.0) - dthetaz * MLOG(sigma0) * MPOW(MCOS(lat), 2.0))
                                               * MPOW(sigma0, kappa));
            const double kt = ka + (ks - ka) * fmax(0.0, (sigma - sigmab) / (1.0 - sigmab)) * MPOW(MCOS(lat), 4.0);
            const double kv = kf * fmax(0.0, (sigma - sigmab) / (1.0 - sigmab));
#pragma unroll
            for (int k = 0; k < 3; k++) VEC(f.Mh, k, lev, p) = VEC(f.Mh, k, lev, p) / (1.0 + kv * time_step);
            T -= kt * time_step * (T - Teq);
        } else {
            double Teq, tau;
            nr_profile(nr, lon, lat, pre, Teq, tau);
            T = (T + time_step * Teq / tau) / (1.0 + time_step / tau);
        }
        if (bl.on) bl_drag(g, f, bl, p, lev, pre, ps_bl, time_step);  // the module after the forcing
        nan_t |= isnan(T);
        LV(f.p, lev) = T * Rd * rho;
        if (zero_q) LV(f.Qheat, lev) = 0.0;
    }
    if (nan_t) atomicOr(f.flags, FLAG_NAN_T);
}

// ------------------------------------------------- double-gray radiation
struct RTPar {
    double kappa_sw, kappa_lw, n_sw, n_lw, f_lw, diff_ang, Tint, albedo, incflx, ps0, lon_sub;
};

// source_func_lin (profx_RT.h:133-160): layer emission with a source linear
// in optical depth between the layer's bottom, centre and top values.
__device__ __forceinline__ double rt_source(double bb, double bl, double bt, double tau, double diff_ang) {
    double e1 = 0.0, e2 = 0.0;
    if (tau >= 1e-10) {
        e1 = bb - bl + (bl - (diff_ang / (tau / 2.0)) * (bb - bl)) * (1.0 - MEXP(-(tau / 2.0) / diff_ang));
        e2 = bl - bt + (bt - (diff_ang / (tau / 2.0)) * (bl - bt)) * (1.0 - MEXP(-(tau / 2.0) / diff_ang));
    } else {
        // optically thin layer: the reference's 5-term series; (-1)^(i+2) and
        // (i+2)! are exact integers
        double fac = 1.0;
        for (int i = 0; i < 5; i++) {
            fac *= (double)(i + 2);
            const double sgn = (i % 2 == 0) ? 1.0 : -1.0;
            const double xp = MPOW((tau / 2.0) / diff_ang, i + 1.0);
            e1 = e1 + sgn * ((bl + (i + 1) * bl) / fac) * xp;
            e2 = e2 + sgn * ((bt + (i + 1) * bb) / fac) * xp;
        }
    }
    return e1 + e2 * MEXP(-(tau / 2.0) / diff_ang);
}

// rtm_dual_band (profx_RT.h:296-587) without surface, for one column:
// interface pressure and temperature, optical depths (computetau), shortwave
// down with an extra layer above the top (radcsw; no upward shortwave),
// longwave down and up (radclw), heating = shortwave then longwave flux
// divergence [W/m3], clipped so p stays >= 0, added to Qheat; and the
// column's insolation, ASR and OLR. One thread per column. Shortwave and
// longwave downward fluxes go to two interface planes (fdn_sw, fdn_lw); the
// upward longwave flux is carried level to level.
// Double-gray ProfX, CB columns per 256-thread block. Phase 1 (all threads,
// a share of the levels each): W and T, and per layer the shortwave and
// longwave transmissions and the longwave source terms of both directions
// into LDS; phase 2 (one lane per column): the three serial recurrences of
// the reference (shortwave down, longwave down, longwave up) over those
// terms; phase 3 (all threads): the layer heating, the pressure clip and
// p = rho Rd T. Each term and each recurrence step is the reference's
// arithmetic, so the results are those of the one-thread-per-column form.
// LDS: 4 planes of (nv + 1) x CB doubles (A: SW transmission -> SW flux,
// B: LW transmission, C: LW down source -> LW down flux, D: LW up source ->
// LW up flux). list != nullptr: just the *count columns list[0 .. *count)
// (the surrogate's fallback), block b taking entries b CB .. b CB + CB - 1.
// The optical-depth exponents n - 1 of the RT are 0 or 1 in the usual set-ups
// (n_sw = 1, n_lw = 2): those powers are 1 and x exactly, as pow returns them,
// without the call.
__device__ __forceinline__ double rt_pow(double x, double y) {
    if (y == 0.0) return 1.0;
    if (y == 1.0) return x;
    return MPOW(x, y);
}

template <int CB>
__global__ __launch_bounds__(256) void k_profx_rt(Geo g, Fields f, RTPar q, double time_step, double *insol,
                                                  double *asr, double *olr, const int *list, const int *count) {
    extern __shared__ double rts[];
    constexpr int NROW = 256 / CB;
    const int c = threadIdx.x % CB, row = threadIdx.x / CB;
    int p = blockIdx.x * CB + c;
    bool act = p < g.P;
    if (list) {
        const int n = *count;
        if ((int)blockIdx.x * CB >= n) return;  // the whole block: before any barrier
        act = p < n;
        p = act ? list[p] : 0;
    }
    const size_t S = g.S;
    const int nv = g.nv, L = nv + 1;
    double *sA = rts, *sB = rts + L * CB, *sC = rts + 2 * L * CB, *sD = rts + 3 * L * CB;
    const double *Alt = g.Alt, *Alth = g.Alth;
    const double Rd = g.Rd, Cp = g.Cp, Gravit = g.Gravit, A = g.A;
    const double bc = 5.677036E-8;  // the reference's Stefan-Boltzmann constant (profx_RT.h:178)
    auto Tl = [&](int lev) { return LV(f.p, lev) / (Rd * LV(f.Rho, lev)); };
    auto pint = [&](int lev) {  // interface pressure (phtemp)
        if (lev == 0) {
            const double psm = LV(f.p, 1) - LV(f.Rho, 0) * Gravit * (-Alt[0] - Alt[1]);
            return 0.5 * (LV(f.p, 0) + psm);
        }
        if (lev == nv) {
            double pp = LV(f.p, nv - 2) + (LV(f.p, nv - 1) - LV(f.p, nv - 2)) / (Alt[nv - 1] - Alt[nv - 2])
                                              * (2 * Alth[nv] - Alt[nv - 1] - Alt[nv - 2]);
            if (pp < 0) pp = 0;
            return 0.5 * (LV(f.p, nv - 1) + pp);
        }
        const double xi = Alth[lev], xim = Alt[lev - 1], xip = Alt[lev];
        const double a = (xi - xip) / (xim - xip), b = (xi - xim) / (xip - xim);
        return LV(f.p, lev - 1) * a + LV(f.p, lev) * b;
    };
    auto tint = [&](int lev) {  // interface temperature (thtemp)
        if (lev == 0) return Tl(0);
        if (lev == nv) return Tl(nv - 1);
        const double xi = Alth[lev], xim = Alt[lev - 1], xip = Alt[lev];
        const double a = (xi - xip) / (xim - xip), b = (xi - xim) / (xip - xim);
        return Tl(lev - 1) * a + Tl(lev) * b;
    };
    auto tau_sw = [&](int lev) {
        return q.n_sw * q.kappa_sw * rt_pow(LV(f.p, lev) / q.ps0, q.n_sw - 1) * LV(f.Rho, lev)
               * (Alth[lev + 1] - Alth[lev]);
    };
    auto tau_lw = [&](int lev) {
        const double t = q.kappa_lw * (1 + q.n_lw * (1 - q.f_lw) / q.f_lw * rt_pow(LV(f.p, lev) / q.ps0, q.n_lw - 1))
                         * LV(f.Rho, lev) * (Alth[lev + 1] - Alth[lev]);
        return t < 0.0 ? 0.0 : t;
    };
    auto planck = [&](double t) { return bc * t * t * t * t; };
    double coszrs = 0.0;
    if (act) {
        const double lon = g.lonlat[p], lat = g.lonlat[S + p];
        coszrs = MCOS(lat) * MCOS(lon - q.lon_sub);
        if (coszrs < 0.0) coszrs = 0.0;
    }
    const bool day = coszrs > 0.0;
    // phase 1
    if (act) {
        for (int lev = row; lev < nv; lev += NROW) {
            const double whl = LV(f.Wh, lev), wht = LV(f.Wh, lev + 1);
            const double xi = Alt[lev], xim = Alth[lev], xip = Alth[lev + 1];
            const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            if (!f.skip_w) LV(f.W, lev) = whl * intt + wht * intl;
            const double T = Tl(lev);
            LV(f.T, lev) = T;
            sA[lev * CB + c] = day ? MEXP(-(1.0 / coszrs) * tau_sw(lev)) : 0.0;
            const double tl_ = tau_lw(lev);
            sB[lev * CB + c] = MEXP(-(1. / q.diff_ang) * tl_);
            const double bt = planck(tint(lev + 1)), bm = planck(T), bb = planck(tint(lev));
            sC[lev * CB + c] = rt_source(bb, bm, bt, tl_, q.diff_ang);           // downward, layer lev
            sD[(lev + 1) * CB + c] = rt_source(bt, bm, bb, tl_, q.diff_ang);     // upward, to interface lev + 1
        }
    }
    __syncthreads();
    // phase 2
    if (act && row == 0) {
        if (day) {
            const double flux_top = q.incflx * coszrs * (1.0 - q.albedo);
            const double tau = (q.kappa_sw / Gravit) * pint(nv);
            double fd = flux_top * MEXP(-(1.0 / coszrs) * tau);
            sA[nv * CB + c] = fd;
            for (int lev = nv; lev >= 1; lev--) {
                fd = fd * sA[(lev - 1) * CB + c];
                sA[(lev - 1) * CB + c] = fd;
            }
        } else {
            sA[nv * CB + c] = 0.0;
        }
        double fl = 0.0;
        sC[nv * CB + c] = 0.0;
        for (int lev = nv - 1; lev >= 0; lev--) {
            fl = sC[lev * CB + c] + fl * sB[lev * CB + c];
            sC[lev * CB + c] = fl;
        }
        double fu_lo = planck(q.Tint);
        fu_lo += sC[c];
        sD[c] = fu_lo;
        for (int lev = 1; lev <= nv; lev++) {
            const double fu = sD[lev * CB + c] + fu_lo * sB[(lev - 1) * CB + c];
            sD[lev * CB + c] = fu;
            fu_lo = fu;
        }
        insol[p] = day ? q.incflx * coszrs : 0.0;
        const double rscale = g.deep ? (A + Alth[nv]) / A : 1.0;
        asr[p] = sA[nv * CB + c] * g.areasT[p] * MPOW(rscale, 2);
        olr[p] = sD[nv * CB + c] * g.areasT[p] * MPOW(rscale, 2);
    }
    __syncthreads();
    // phase 3
    if (!act) return;
    for (int l = row; l < nv; l += NROW) {
        double dtemp = 0.0;
        if (day) dtemp = -((0.0 - sA[l * CB + c]) - (0.0 - sA[(l + 1) * CB + c])) / ((Alth[l] - Alth[l + 1]));
        dtemp = dtemp
                - ((sD[l * CB + c] - sC[l * CB + c]) - (sD[(l + 1) * CB + c] - sC[(l + 1) * CB + c]))
                      / ((Alth[l] - Alth[l + 1]));
        const double pr = LV(f.p, l);
        if (pr + Rd / (Cp - Rd) * dtemp * time_step < 0) dtemp = -pr / time_step;
        LV(f.Qheat, l) = dtemp;  // Qheat is cleared at the start of ProfX, then the module adds
        // Compute_pressure (profx_auxiliary.h:95-108) closes ProfX
        LV(f.p, l) = Tl(l) * Rd * LV(f.Rho, l);
    }
}

// ------------------------------------------------------ RT surrogate glue
// Normalisation of the surrogate network (see thor_rt_surrogate_params).
struct SurPar {
    const float *in_mean, *in_std, *in_lo, *in_hi;  // [2 nv + 1]
    const float *out_mean, *out_std;                 // [nv]
    double lon_sub;
};

// Per column: Recompute_W and T = p / (rho Rd) as ProfX does, the network's
// normalised inputs [T_k, ln p_k, cos z] as BF16 row p of X, and the flag of
// a column whose inputs leave the training envelope.
// 64 columns x 4 level rows per block: W from Wh and T at every level (as
// ProfX does), the normalised inputs [T_k, ln p_k, cos z] of each column staged
// in LDS and written out as contiguous BF16 rows; a column is outside the
// training envelope if any of its normalised inputs is.
// Normalised input of the surrogate, rounded to BF16: the difference from the
// mean formed in FP64 and rounded to FP32, times the FP32 reciprocal of the
// scale (the network reads 8 significant bits; FP64 keeps the difference
// exact where the scale is small against the mean). `out` is raised when the
// value leaves the training envelope [lo, hi]. ln p is the FP64 logarithm
// (an FP32 one, an ulp or more from the host's, would move a few per cent of
// the normalised inputs across a BF16 rounding boundary).
__device__ __forceinline__ uint16_t sur_input(double x, float mean, float inv_std, float lo, float hi, bool &out) {
    const float xn = (float)(x - (double)mean) * inv_std;
    out |= !(xn >= lo && xn <= hi);
    return mlp::f2bf_dev(xn);
}
__device__ __forceinline__ double sur_lnp(double p) { return MLOG(p); }

// The envelope flags of a surrogate call alone, one thread per column, with
// the inputs formed as k_sur_persist forms them (T = p / (Rd rho), ln p,
// cos z): out-of-envelope columns are marked in mask and appended to list, so
// the exact RT can take them on another stream while k_sur_persist (given
// the mask) runs the network on the others.
__global__ __launch_bounds__(256) void k_sur_flags(Geo g, Fields f, SurPar q, int *mask, int *list, int *cnt,
                                                   int *cnt_next, int *nflag) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) *cnt_next = 0;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    bool out = false;
    for (int lev = 0; lev < nv; lev++) {
        const double pr = LV(f.p, lev);
        const double T = pr / (g.Rd * LV(f.Rho, lev));
        sur_input(T, q.in_mean[lev], 1.0f / q.in_std[lev], q.in_lo[lev], q.in_hi[lev], out);
        sur_input(sur_lnp(pr), q.in_mean[nv + lev], 1.0f / q.in_std[nv + lev], q.in_lo[nv + lev], q.in_hi[nv + lev],
                  out);
    }
    double coszrs = MCOS(g.lonlat[S + p]) * MCOS(g.lonlat[p] - q.lon_sub);
    if (coszrs < 0.0) coszrs = 0.0;
    sur_input(coszrs, q.in_mean[2 * nv], 1.0f / q.in_std[2 * nv], q.in_lo[2 * nv], q.in_hi[2 * nv], out);
    mask[p] = out;
    if (out) {
        atomicAdd(nflag, 1);
        list[atomicAdd(cnt, 1)] = p;
    }
}

__global__ __launch_bounds__(256) void k_sur_features(Geo g, Fields f, SurPar q, uint16_t *X, int *mask,
                                                      int *list, int *cnt, int *cnt_next, int *nflag) {
    constexpr int CB = 64, NROW = 4;
    extern __shared__ uint16_t sx[];  // [CB][nin]
    __shared__ int sout[CB];
    const int c = threadIdx.x % CB, row = threadIdx.x / CB;
    const int p0 = blockIdx.x * CB, p = p0 + c;
    const size_t S = g.S;
    const int nv = g.nv, nin = 2 * nv + 1;
    const double *Alt = g.Alt, *Alth = g.Alth;
    if (threadIdx.x < CB) sout[threadIdx.x] = 0;
    __syncthreads();
    bool out = false;
    auto put = [&](int j, double x) {
        sx[c * nin + j] = sur_input(x, q.in_mean[j], 1.0f / q.in_std[j], q.in_lo[j], q.in_hi[j], out);
    };
    if (p < g.P) {
        for (int lev = row; lev < nv; lev += NROW) {
            const double whl = LV(f.Wh, lev), wht = LV(f.Wh, lev + 1);
            const double xi = Alt[lev], xim = Alth[lev], xip = Alth[lev + 1];
            const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
            if (!f.skip_w) LV(f.W, lev) = whl * intt + wht * intl;
            const double pr = LV(f.p, lev);
            const double T = pr / (g.Rd * LV(f.Rho, lev));
            LV(f.T, lev) = T;
            put(lev, T);
            put(nv + lev, sur_lnp(pr));
        }
        if (row == 0) {
            double coszrs = MCOS(g.lonlat[S + p]) * MCOS(g.lonlat[p] - q.lon_sub);
            if (coszrs < 0.0) coszrs = 0.0;
            put(2 * nv, coszrs);
        }
        if (out) atomicOr(&sout[c], 1);
    }
    __syncthreads();
    const int ncols = min(CB, g.P - p0);
    for (int i = threadIdx.x; i < ncols * nin; i += 256) X[(size_t)p0 * nin + i] = sx[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) *cnt_next = 0;
    if ((int)threadIdx.x < ncols) {
        const int o = sout[threadIdx.x];
        mask[p0 + threadIdx.x] = o;
        if (o) {
            atomicAdd(nflag, 1);
            list[atomicAdd(cnt, 1)] = p0 + threadIdx.x;
        }
    }
}

// Qheat from the network's outputs Y[p][lev] (with the RT's positivity clip),
// for the columns inside the envelope; with set_p, also Compute_pressure
// p = T Rd rho for them (the exact RT does it for the others).
// The whole surrogate ProfX, one block of 8 waves per CU looping over groups
// of 32 columns, with the network's weights in the waves' registers for the
// whole kernel (mlp::load_wave_weights). Per group (32 columns x 16 level rows
// of threads, up to KT1 levels each, loaded one group ahead): ProfX's W from Wh
// and T, the normalised inputs [T_k, ln p_k, cos z] staged in LDS as BF16
// with the envelope flag of each column, the three MFMA layers (mlp3_rows),
// and Qheat = y out_std + out_mean with the reference's pressure clip for the
// columns inside the envelope; set_p: also p = rho Rd T, which closes ProfX
// (not on judge calls, where the exact RT follows). Columns outside the
// envelope are flagged in mask and appended to list (count in *cnt), for the
// exact RT to take them; *cnt_next, the next call's counter, is cleared. The
// normalisation and the level interpolation weights are staged in LDS once.
// The same values as k_sur_features + the inference kernel + k_sur_apply.
// Needs nv <= 63 (n_in <= 128, the network's 64 padded outputs). KT1 = kpad / 32;
// NL = ceil(nv / 16) levels per thread (NL <= KT1).
template <int KT1, int NL>
__global__ __launch_bounds__(mlp::kInferThreads, 1) void k_sur_persist(Geo g, Fields f, SurPar q,
                                                                       mlp::InferWeights w, int *mask, int *list,
                                                                       int *cnt, int *cnt_next, int *nflag,
                                                                       double time_step, int set_p,
                                                                       const int *pre) {
    constexpr int R = 32, LDH = 256 + 8, NROW = mlp::kInferThreads / R;
    __shared__ __align__(16) uint16_t bufA[R * LDH], bufB[R * LDH];
    __shared__ int sout[R];
    __shared__ float sin_[4][128], sout_[2][64];  // in_mean, 1 / in_std, in_lo, in_hi; out_mean, out_std
    __shared__ double sint[2][64];                // ProfX's W interpolation weights intt, intl per level
    const int tid = threadIdx.x, r = tid & (R - 1), row = tid / R;
    const size_t S = g.S;
    const int nv = g.nv, nin = 2 * nv + 1, ldx = w.kpad + 8;
    const double Rd = g.Rd, Cp = g.Cp;
    if (blockIdx.x == 0 && tid == 0) *cnt_next = 0;
    for (int j = tid; j < nin; j += mlp::kInferThreads) {
        sin_[0][j] = q.in_mean[j];
        sin_[1][j] = 1.0f / q.in_std[j];
        sin_[2][j] = q.in_lo[j];
        sin_[3][j] = q.in_hi[j];
    }
    for (int lev = tid; lev < nv; lev += mlp::kInferThreads) {
        sout_[0][lev] = q.out_mean[lev];
        sout_[1][lev] = q.out_std[lev];
        const double xi = g.Alt[lev], xim = g.Alth[lev], xip = g.Alth[lev + 1];
        sint[0][lev] = (xi - xip) / (xim - xip);
        sint[1][lev] = (xi - xim) / (xip - xim);
    }
    mlp::WaveWeights<KT1> wr;
    mlp::load_wave_weights<KT1>(w, wr);
    const int ngroups = (g.P + R - 1) / R;
    // this thread's levels row, row + NROW, ... of group gi: Wh at both
    // interfaces, p and rho, loaded one group ahead (the next group's loads
    // are in flight while this group's network runs)
    double whl[NL], wht[NL], pr[NL], rho[NL], T[NL];
    auto load_group = [&](int gi) {
        const int p = gi * R + r;
#pragma unroll
        for (int k = 0; k < NL; k++) {
            const int lev = row + k * NROW;
            const bool on = gi < ngroups && p < g.P && lev < nv;
            const int lv = on ? lev : 0, pp = on ? p : 0;
            whl[k] = f.Wh[(size_t)lv * S + pp];
            wht[k] = f.Wh[(size_t)(lv + 1) * S + pp];
            pr[k] = f.p[(size_t)lv * S + pp];
            rho[k] = f.Rho[(size_t)lv * S + pp];
        }
    };
    load_group(blockIdx.x);
    for (int gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const int p = gi * R + r;
        // pre: the flags of k_sur_flags; a flagged column is left to the exact RT
        // entirely (no reads used, no writes)
        const bool valid = p < g.P && !(pre && pre[p]);
        if (tid < R) sout[tid] = 0;
        __syncthreads();  // sout cleared (and, first time round, the LDS tables written)
        bool out = false;
        auto put = [&](int j, double x) {
            bufA[r * ldx + j] = sur_input(x, sin_[0][j], sin_[1][j], sin_[2][j], sin_[3][j], out);
        };
#pragma unroll
        for (int k = 0; k < NL; k++) {
            const int lev = row + k * NROW;
            T[k] = 0.0;
            if (lev >= nv) continue;
            if (valid) {
                if (!f.skip_w) LV(f.W, lev) = whl[k] * sint[0][lev] + wht[k] * sint[1][lev];
                T[k] = pr[k] / (Rd * rho[k]);
                LV(f.T, lev) = T[k];
                put(lev, T[k]);
                put(nv + lev, sur_lnp(pr[k]));
            } else {
                bufA[r * ldx + lev] = bufA[r * ldx + nv + lev] = 0;
            }
        }
        if (row == 0) {
            if (valid) {
                double coszrs = MCOS(g.lonlat[S + p]) * MCOS(g.lonlat[p] - q.lon_sub);
                if (coszrs < 0.0) coszrs = 0.0;
                put(2 * nv, coszrs);
            } else {
                bufA[r * ldx + 2 * nv] = 0;
            }
            for (int k = nin; k < w.kpad; k++) bufA[r * ldx + k] = 0;
        }
        if (out) atomicOr(&sout[r], 1);
        __syncthreads();
        // this group's epilogue keeps p, rho and T; the next group's inputs go in flight
        double pr_c[NL], rho_c[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) {
            pr_c[k] = pr[k];
            rho_c[k] = rho[k];
        }
        load_group(gi + gridDim.x);
        float y[2][4];
        mlp::mlp3_rows<KT1>(bufA, bufB, w.kpad, wr, y);
        // bufB is free once layer 2 is done: stage the outputs there as [R][64] FP32
        float *sY = reinterpret_cast<float *>(bufB);
        {
            const int wave = tid >> 6, lane = tid & 63, li = lane & 15, lq = (lane >> 4) * 4;
            if (wave < 4) {
#pragma unroll
                for (int ti = 0; ti < 2; ti++)
#pragma unroll
                    for (int e = 0; e < 4; e++) sY[(16 * ti + li) * 64 + 16 * wave + lq + e] = y[ti][e];
            }
        }
        __syncthreads();
        if (valid) {
            const int o = sout[r];
            if (!o) {
#pragma unroll
                for (int k = 0; k < NL; k++) {
                    const int lev = row + k * NROW;
                    if (lev >= nv) continue;
                    double Q = (double)sY[r * 64 + lev] * sout_[1][lev] + sout_[0][lev];
                    if (pr_c[k] + Rd / (Cp - Rd) * Q * time_step < 0) Q = -pr_c[k] / time_step;
                    LV(f.Qheat, lev) = Q;
                    if (set_p) LV(f.p, lev) = T[k] * Rd * rho_c[k];
                }
            }
            if (row == 0 && !pre) {
                mask[p] = o;
                if (o) {
                    atomicAdd(nflag, 1);
                    list[atomicAdd(cnt, 1)] = p;
                }
            }
        }
        __syncthreads();  // sY and bufA are read before the next group writes them
    }
}

// 64 columns x 4 level rows per block: the block's rows of Y (64 nv
// contiguous floats) are read coalesced into LDS (row stride nv + 1 against
// bank conflicts), then each thread forms Qheat (and p) for its column at
// every fourth level, with the level-planar stores coalesced across columns.
__global__ __launch_bounds__(256) void k_sur_apply(Geo g, Fields f, SurPar q, const float *Y, const int *mask,
                                                   double time_step, int set_p) {
    constexpr int CB = 64, NROW = 4;
    extern __shared__ float sYa[];  // [CB][nv + 1]
    const int nv = g.nv, ld = nv + 1;
    const int p0 = blockIdx.x * CB, ncols = min(CB, g.P - p0);
    for (int i = threadIdx.x; i < ncols * nv; i += 256) sYa[(i / nv) * ld + i % nv] = Y[(size_t)p0 * nv + i];
    __syncthreads();
    const int c = threadIdx.x % CB, row = threadIdx.x / CB, p = p0 + c;
    if (c >= ncols || mask[p]) return;
    const size_t S = g.S;
    const double Rd = g.Rd, Cp = g.Cp;
    for (int lev = row; lev < nv; lev += NROW) {
        double Q = (double)sYa[c * ld + lev] * q.out_std[lev] + q.out_mean[lev];
        const double pr = LV(f.p, lev);
        if (pr + Rd / (Cp - Rd) * Q * time_step < 0) Q = -pr / time_step;
        LV(f.Qheat, lev) = Q;
        if (set_p) LV(f.p, lev) = LV(f.T, lev) * Rd * LV(f.Rho, lev);
    }
}

// Judge: squared and absolute differences between the exact RT heating and
// the surrogate's (saved in qsur) over the columns inside the envelope, one
// partial sum / max per block.
__global__ __launch_bounds__(256) void k_sur_error(Geo g, const double *Qexact, const double *qsur, const int *mask,
                                                   long n, double *part_sq, double *part_max, double *part_cnt) {
    __shared__ double s_sq[256], s_mx[256], s_ct[256];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double sq = 0.0, mx = 0.0, ct = 0.0;
    if (i < n) {
        const int p = (int)(i % g.S);
        if (p < g.P && !mask[p]) {
            const double d = Qexact[i] - qsur[i];
            sq = d * d;
            mx = fabs(d);
            ct = 1.0;
        }
    }
    s_sq[threadIdx.x] = sq;
    s_mx[threadIdx.x] = mx;
    s_ct[threadIdx.x] = ct;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            s_sq[threadIdx.x] += s_sq[threadIdx.x + k];
            s_mx[threadIdx.x] = fmax(s_mx[threadIdx.x], s_mx[threadIdx.x + k]);
            s_ct[threadIdx.x] += s_ct[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_sq[blockIdx.x] = s_sq[0];
        part_max[blockIdx.x] = s_mx[0];
        part_cnt[blockIdx.x] = s_ct[0];
    }
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

// ===================================================================== sponge
// THOR's RayleighSponge with raysp_calc_mode = imp (profx_driver.cu:92-195),
// at the start of ProfX: (1) Recompute_W and each (point, level)'s u, v, w
// divided by its latitude bin's point count into the bin's slot (zonal_uv /
// zonal_w, profx_sponge.cu:50-102); (2) per (component, bin, level) the
// 1024-thread tree sum of gpu_sum_on_device_sponge (reduction_add.h:193-282)
// over 2048 slots per block, and the block sums added in order; (3) the
// implicit relaxation of sponge_layer (profx_sponge.cu:138-437) per (point,
// level); (4) Wh at the interior interfaces from the new W.
struct SpongePar {
    double ns, Ruv, Rw;
    int nlat, dmu, dmw, max_count, nb;  // nb = blocks of 2048 slots per row
    int dmt;                            // temperature sponge (RayleighSpongeT)
    double RT;
    int mode;                           // raysp_calc_mode: 0 imp, 1 exp1, 3 exp3
    const double *trig;                 // [4][S]: sin, cos of the longitude (in [0, 2 pi)) and of the latitude
    const double *kf;                   // [nv]: sin^2(pi/2 (n - ns) / (1 - ns)) above ns, else 0
    const int *cnt;                     // [nlat]: points per latitude bin
    __host__ __device__ int ncomp() const { return dmt ? 4 : 3; }  // zonal-mean components (u, v, w, T)
};

// The per-point and per-level factors of the sponge, formed once with the
// expressions of the per-cell formulas (so the results are unchanged).
__global__ __launch_bounds__(256) void k_sponge_tables(Geo g, SpongePar sp, double *trig, double *kf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t S = g.S;
    const int nv = g.nv;
    if (i < g.P) {
        const double lon = g.lonlat[i] < 0 ? g.lonlat[i] + 2.0 * M_PI : g.lonlat[i], lat = g.lonlat[S + i];
        trig[i] = MSIN(lon);
        trig[S + i] = MCOS(lon);
        trig[2 * S + i] = MSIN(lat);
        trig[3 * S + i] = MCOS(lat);
    }
    if (i < nv) {
        const double n = g.Alt[i] / g.Alth[nv];
        kf[i] = n >= sp.ns ? MPOW(MSIN(0.5 * M_PI * (n - sp.ns) * (1.0) / (1.0 - sp.ns)), 2.0) : 0.0;
    }
}

__device__ __forceinline__ void sponge_uv(const Geo &g, const Fields &f, const SpongePar &sp, int p, int lev,
                                          double &u, double &v) {
    const size_t S = g.S;
    const int nv = g.nv;
    const double sl = sp.trig[p], cl = sp.trig[S + p], sa = sp.trig[2 * S + p], ca = sp.trig[3 * S + p];
    const double rho = LV(f.Rho, lev);
    const double m0 = VEC(f.Mh, 0, lev, p), m1 = VEC(f.Mh, 1, lev, p), m2 = VEC(f.Mh, 2, lev, p);
    u = (m0 * (-sl) + m1 * (cl) + m2 * 0) / rho;
    v = (m0 * (-sa * cl) + m1 * (-sa * sl) + m2 * (ca)) / rho;
}

// ztab: [3][S] latitude bin, points in the bin, index within the bin;
// tmp: [ncomp][nlat][nv][max_count] (u, v, w, T), zero beyond each bin's count.
__global__ __launch_bounds__(256) void k_sponge_zonal(Geo g, Fields f, SpongePar sp, const int *ztab, double *tmp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x, lev = blockIdx.y;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    const double *Alt = g.Alt, *Alth = g.Alth;
    double w;
    if (sp.mode == 0) {  // Recompute_W first (profx_driver.cu:92)
        const double whl = LV(f.Wh, lev), wht = LV(f.Wh, lev + 1);
        const double xi = Alt[lev], xim = Alth[lev], xip = Alth[lev + 1];
        const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
        w = whl * intt + wht * intl;
        LV(f.W, lev) = w;
    } else {  // explicit modes: the stage's W (the caller passes Wk as f.W)
        w = LV(f.W, lev);
    }
    if (!sp.dmu && !sp.dmw && !sp.dmt) return;
    const int cnt = ztab[S + p];
    const size_t T = (size_t)sp.nlat * nv * sp.max_count;
    const size_t at = ((size_t)ztab[p] * nv + lev) * sp.max_count + ztab[2 * S + p];
    if (sp.dmu || sp.dmw) {
        double u, v;
        sponge_uv(g, f, sp, p, lev, u, v);
        tmp[at] = u / cnt;
        tmp[T + at] = v / cnt;
        tmp[2 * T + at] = (w / LV(f.Rho, lev)) / cnt;
    }
    if (sp.dmt) tmp[3 * T + at] = (LV(f.p, lev) / g.Rd / LV(f.Rho, lev)) / cnt;  // zonal_temp
}

// One block per row (component, bin, level): the row's 2048-slot blocks in
// turn, each summed by the 1024-thread tree into out[row][b]. Slots at or
// beyond the bin's count hold zero and are not read; a slot block with none
// below the count has the sum +0.
__global__ __launch_bounds__(1024) void k_sponge_reduce(SpongePar sp, int nv, const double *tmp, double *out) {
    __shared__ double ds[2048];
    const int t = threadIdx.x, lev = blockIdx.x % nv, cb = blockIdx.x / nv;  // cb = component * nlat + bin
    const long L = sp.max_count, n = min(L, (long)sp.cnt[cb % sp.nlat]);
    const double *row = tmp + ((size_t)cb * nv + lev) * L;
    double *o = out + ((size_t)cb * nv + lev) * sp.nb;
    for (int b = 0; b < sp.nb; b++) {
        const long base = 2048L * b;
        if (base >= n) {
            if (t == 0) o[b] = 0.0;
            continue;
        }
        const long i = base + 2 * t;
        ds[2 * t] = i < n ? row[i] : 0.0;
        ds[2 * t + 1] = i + 1 < n ? row[i + 1] : 0.0;
        for (int stride = 1024; stride > 0; stride /= 2) {
            __syncthreads();
            if (t < stride) ds[t] += ds[t + stride];
        }
        __syncthreads();
        if (t == 0) o[b] = ds[0];
        __syncthreads();
    }
}

// vbar[bin][lev][component, stride 4] = 0 + the block sums in order (the
// host loops of gpu_sum_on_device_sponge and gpu_sum_on_device).
__global__ void k_sponge_vbar(SpongePar sp, int nv, const double *out, double *vbar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (component, bin, lev)
    if (i >= sp.ncomp() * sp.nlat * nv) return;
    const int comp = i / (sp.nlat * nv), rem = i % (sp.nlat * nv);
    double s = 0.0;
    for (int b = 0; b < sp.nb; b++) s += out[(size_t)i * sp.nb + b];
    vbar[(size_t)rem * 4 + comp] = s;
}

// rv: the shrinking factor Rv_fac (1 without shrink_sponge); each rate is
// Rv_fac R sin^2(..), as sponge_layer forms it.
__global__ __launch_bounds__(256) void k_sponge_apply(Geo g, Fields f, SpongePar sp, const int *ztab,
                                                     const double *vbar, double dt, double rv) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x, lev = blockIdx.y;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv, nlat = sp.nlat;
    const int intlat = ztab[p];
    const double des_lat = M_PI / nlat;
    const double lat1 = -M_PI / 2 + 0.5 * des_lat + des_lat * (intlat - 1);
    const double lat2 = -M_PI / 2 + 0.5 * des_lat + des_lat * (intlat);
    const double lat3 = -M_PI / 2 + 0.5 * des_lat + des_lat * (intlat + 1);
    double u, v;
    sponge_uv(g, f, sp, p, lev, u, v);
    const double lat = g.lonlat[S + p];
    auto mean_at = [&](int c) {
        if (lat <= -M_PI / 2 + 0.5 * des_lat) return vbar[((size_t)intlat * nv + lev) * 4 + c];
        if (lat >= -M_PI / 2 + 0.5 * des_lat + des_lat * (nlat - 1)) return vbar[((size_t)intlat * nv + lev) * 4 + c];
        if (lat >= lat1 && lat < lat2) {
            const double intt = (lat - lat2) / (lat1 - lat2), intl = (lat - lat1) / (lat2 - lat1);
            return vbar[((size_t)(intlat - 1) * nv + lev) * 4 + c] * intt + vbar[((size_t)intlat * nv + lev) * 4 + c] * intl;
        }
        const double intt = (lat - lat3) / (lat2 - lat3), intl = (lat - lat2) / (lat3 - lat2);
        return vbar[((size_t)intlat * nv + lev) * 4 + c] * intt + vbar[((size_t)(intlat + 1) * nv + lev) * 4 + c] * intl;
    };
    const double rho = LV(f.Rho, lev);
    const double w = LV(f.W, lev) / rho;
    const double kuv = rv * sp.Ruv * sp.kf[lev], kw = rv * sp.Rw * sp.kf[lev];
    double unew, vnew, wnew;
    if (sp.dmu) {
        unew = (u / dt + kuv * mean_at(0)) / (1.0 / dt + kuv);
        vnew = (v / dt + kuv * mean_at(1)) / (1.0 / dt + kuv);
    } else {
        unew = (u / dt) / (1.0 / dt + kuv);
        vnew = (v / dt) / (1.0 / dt + kuv);
    }
    if (sp.dmw) wnew = (w / dt + kw * mean_at(2)) / (1.0 / dt + kw);
    else wnew = (w / dt) / (1.0 / dt + kw);
    const double sl = sp.trig[p], cl = sp.trig[S + p], sa = sp.trig[2 * S + p], ca = sp.trig[3 * S + p];
    const double vx = unew * (-sl) + vnew * (-sa * cl);
    const double vy = unew * (cl) + vnew * (-sa * sl);
    const double vz = vnew * (ca);
    VEC(f.Mh, 0, lev, p) = vx * rho;
    VEC(f.Mh, 1, lev, p) = vy * rho;
    VEC(f.Mh, 2, lev, p) = vz * rho;
    LV(f.W, lev) = wnew * rho;
    if (sp.dmt) {  // temperature towards its zonal mean, through the pressure
        const double kT = rv * sp.RT * sp.kf[lev], temperature = LV(f.p, lev) / g.Rd / rho;
        LV(f.p, lev) = rho * g.Rd * (temperature / dt + kT * mean_at(3)) / (1.0 / dt + kT);
    }
}

// The zonal mean of component c (0 u, 1 v, 2 w, 3 T) at point p, level lev,
// as k_sponge_apply forms it.
__device__ __forceinline__ double sponge_mean(const Geo &g, const SpongePar &sp, const int *ztab, const double *vbar,
                                              int p, int lev, int c) {
    const size_t S = g.S;
    const int nv = g.nv, nlat = sp.nlat, intlat = ztab[p];
    const double des_lat = M_PI / nlat, lat = g.lonlat[S + p];
    const double lat1 = -M_PI / 2 + 0.5 * des_lat + des_lat * (intlat - 1);
    const double lat2 = -M_PI / 2 + 0.5 * des_lat + des_lat * (intlat);
    const double lat3 = -M_PI / 2 + 0.5 * des_lat + des_lat * (intlat + 1);
    if (lat <= -M_PI / 2 + 0.5 * des_lat) return vbar[((size_t)intlat * nv + lev) * 4 + c];
    if (lat >= -M_PI / 2 + 0.5 * des_lat + des_lat * (nlat - 1)) return vbar[((size_t)intlat * nv + lev) * 4 + c];
    if (lat >= lat1 && lat < lat2) {
        const double intt = (lat - lat2) / (lat1 - lat2), intl = (lat - lat1) / (lat2 - lat1);
        return vbar[((size_t)(intlat - 1) * nv + lev) * 4 + c] * intt + vbar[((size_t)intlat * nv + lev) * 4 + c] * intl;
    }
    const double intt = (lat - lat3) / (lat2 - lat3), intl = (lat - lat2) / (lat3 - lat2);
    return vbar[((size_t)intlat * nv + lev) * 4 + c] * intt + vbar[((size_t)(intlat + 1) * nv + lev) * 4 + c] * intl;
}

// Explicit sponge (sponge_layer with implicit = false, profx_sponge.cu:255-292,
// 388-404): f.Mh, f.Rho, f.W, f.p are the stage values Mhk, Rhok, Wk,
// pressurek; the drag times rho goes into dMh and dW (cleared first, as THOR's
// memsets do: 0 + x), the heating Rd rho dT is added to Qheat.
__global__ __launch_bounds__(256) void k_sponge_drag(Geo g, Fields f, SpongePar sp, const int *ztab,
                                                    const double *vbar, double *dMh, double *dW, double *Qheat,
                                                    double rv) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x, lev = blockIdx.y;
    if (p >= g.P) return;
    const size_t S = g.S;
    const int nv = g.nv;
    double u, v;
    sponge_uv(g, f, sp, p, lev, u, v);
    const double rho = LV(f.Rho, lev), w = LV(f.W, lev) / rho;
    const double kuv = rv * sp.Ruv * sp.kf[lev], kw = rv * sp.Rw * sp.kf[lev];
    double du, dv;
    if (sp.dmu) {
        du = -kuv * (u - sponge_mean(g, sp, ztab, vbar, p, lev, 0));
        dv = -kuv * (v - sponge_mean(g, sp, ztab, vbar, p, lev, 1));
    } else {
        du = -kuv * u;
        dv = -kuv * v;
    }
    const double sl = sp.trig[p], cl = sp.trig[S + p], sa = sp.trig[2 * S + p], ca = sp.trig[3 * S + p];
    const double vx = du * (-sl) + dv * (-sa * cl);
    const double vy = du * (cl) + dv * (-sa * sl);
    const double vz = dv * (ca);
    VEC(dMh, 0, lev, p) = 0.0 + vx * rho;
    VEC(dMh, 1, lev, p) = 0.0 + vy * rho;
    VEC(dMh, 2, lev, p) = 0.0 + vz * rho;
    const double dw = sp.dmw ? -kw * (w - sponge_mean(g, sp, ztab, vbar, p, lev, 2)) : -kw * (w);
    LV(dW, lev) = 0.0 + dw * rho;
    if (sp.dmt) {
        const double kT = rv * sp.RT * sp.kf[lev], temperature = LV(f.p, lev) / g.Rd / rho;
        const double dT = -kT * (temperature - sponge_mean(g, sp, ztab, vbar, p, lev, 3));
        LV(Qheat, lev) += g.Rd * rho * dT;
    }
}

// dWh from the explicit drag dW: interpolated at the interior interfaces,
// zero at the bottom and the top (blockIdx.y = interface 0..nv).
__global__ __launch_bounds__(256) void k_sponge_dwh(Geo g, const double *dW, double *dWh) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x, lev = blockIdx.y;
    if (p >= g.P) return;
    const size_t S = g.S;
    if (lev == 0 || lev == g.nv) {
        LV(dWh, lev) = 0.0;
        return;
    }
    const double xi = g.Alth[lev], xim = g.Alt[lev - 1], xip = g.Alt[lev];
    const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
    LV(dWh, lev) = LV(dW, lev - 1) * intt + LV(dW, lev) * intl;
}

// Wh at interior interface lev (1..nv-1) from the relaxed W (blockIdx.y = lev - 1).
__global__ __launch_bounds__(256) void k_sponge_wh(Geo g, Fields f) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x, lev = blockIdx.y + 1;
    if (p >= g.P) return;
    const size_t S = g.S;
    const double xi = g.Alth[lev], xim = g.Alt[lev - 1], xip = g.Alt[lev];
    const double intt = (xi - xip) / (xim - xip), intl = (xi - xim) / (xip - xim);
    LV(f.Wh, lev) = LV(f.W, lev - 1) * intt + LV(f.W, lev) * intl;
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
