// Device-side data structures of the dynamical core.
//
// Layout in HBM (level-planar, structure of arrays): every per-cell field is
// stored as planes of S doubles, S = P rounded up to a multiple of 64, with
// points in THOR's rhombus-major order (16x16 tiles of 256 consecutive
// points, poles at P-2 and P-1):
//   scalar     f[lev * S + p]            lev in [0, nv)
//   interface  f[lev * S + p]            lev in [0, nv]
//   vector     f[(c * nv + lev) * S + p] c in {x, y, z}
// A wave of 64 lanes therefore touches 64 consecutive doubles of one level.
// Geometry is stored the same way: coefficient planes [ncoef][S].
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace thor {

constexpr int kTile = 256;  // points per rhombus tile = threads per stencil block
constexpr int kAlign = 64;

struct Geo {
    int P, S, nv, nr;
    int tiles_per_xcd;  // ceil((nr + 1) / 8): tiles (incl. the pole block) per XCD
    const int *ring;  // [6][S]: pt1..pt6 (pentagons/poles: ring[4] = ring[5])
    const int *nfc;   // [S]: 6 hexagon, 5 pentagon or pole, 0 padding
    const int *halo;  // [nr][66]: global index of each tile's halo points (-1: missing corner)
    const double *div, *grad;                  // [21][S]  (j * 3 + k)
    const double *areas, *noa, *nti, *nte;     // [18][S]
    const double *areasTr;                     // [6][S]
    const double *fr;                          // [3][S]
    const double *lonlat;                      // [2][S]
    const double *areasT;                      // [S]
    const double *Alt, *Alth, *Kdh4, *Kdhz;    // [nv], [nv+1], [nv], [nv]
    const double *Kdv6;                         // [nv] vertical hyperdiffusion strength (VertHyDiff)
    const double *Kdh2;                         // [nv] diffusive-sponge strength (DiffSponge)
    double ssp;                                 // its sign, (-1)^(order_diff_sponge / 2 + 1)
    double dsp_scale;                           // 1 / the scale of the pass output it reads (mixed form)
    // Mixed precision only: the hyperdiffusion Laplacian and the divergence-
    // damping operator as per-point weights on the 7 stencil slots (0 = self,
    // 1..6 = ring), for rscale = 1:
    //   L(a)       = rs^2 sum_s wlap[s] (a_s - a_0)
    //   DivM(a)_c  = rs^2 sum_{s,k} wdd[c*21 + s*3 + k] a_{s,k}
    //                + rs dz sum_s udd[c*7 + s] (wl_s - wt_s)
    const float *wlap;  // [7][S]
    const float *wdd;   // [63][S]
    const float *udd;   // [21][S]
    // Mixed precision only: per-level FP32 constants [4][nv]: rscale, rscale^2,
    // 1 / (Alth[lev] - Alth[lev + 1]) and Kdhz.
    const float *lvc;
    // Mixed precision only: per-level FP64 constants of the chunked slow modes
    // [6][nv]: r2p, r2m, r2l, rscale, dr2dz, (A / (A + Alt))^2 (k_slow_levels).
    const double *slv;
    double A, Omega, Gravit, P_Ref, Rd, Cp;
    int deep, nonhydro, ghv;
    int energy;  // total-energy thermodynamic equation (FP64 form only)
};

// Buffers of one large step. "k" = value at the current RK stage start,
// "s" = acoustic deviation, plain = state y(n).
struct Fields {
    double *Mh, *Wh, *W, *Rho, *p;        // state y(n)
    double *Mhk, *Whk, *Wk, *Rhok, *pk;   // stage values read by the slow part
    double *Mhk_out, *Whk_out, *Wk_out;   // UpdateRK2 destinations
    double *Rhok_out, *pk_out;
    double *Mhs, *Whs, *Rhos, *ps;        // deviations
    double *T, *h, *pt, *gtil;            // [nv]
    double *hh, *pth, *gtilh;             // [nv+1]
    // Energy equation: pt and pth hold the flux variable h + ep + ek of the
    // density/pressure update instead of the potential temperature, and Etot
    // the total energy, advanced by every acoustic substep.
    double *Etot;                         // [nv] (energy equation only)
    double *Adv;                          // [3][nv]
    double *diff;                         // [6][nv]
    double *diffmh;                       // [3][nv]
    double *diffw, *diffrh, *diffpr;      // [nv]
    double *divg;                         // [3][nv]
    double *SlowMh, *SlowWh, *SlowRho, *Slowp;
    double *Sp, *Sd;                      // implicit-solve sources [nv]
    double *hflx_r, *hflx_pt;             // horizontal flux divergences of the density / pressure update [nv]
    double *hflx_ptk;                     // mixed path: the stage part rs sum_k div_k(Mhk_k pt) of hflx_pt [nv]
    double *tAA, *tCC, *tTT;              // Thomas factorisation of the stage [nv]
    double *Qheat, *dMh, *dWh;            // ProfX tendencies
    unsigned int *flags;                  // bit 0: NaN pressure, bit 1: NaN temperature
    int skip_w;                           // ProfX: W already set (by the sponge), do not recompute it from Wh
    int thomas_rows;                      // mixed path: both sweeps of the vertical solve on all rows of the block
    // Mixed precision: FP32 storage of the momentum deviation, the advection,
    // hyperdiffusion and divergence-damping tendencies and the slow momentum
    // tendency (same layouts as their FP64 counterparts above).
    float *mMhs, *mAdv, *mdiff, *mdiffmh, *mdiffw, *mdiffrh, *mdiffpr, *mdivg, *mSlowMh;
    double *diffx;  // [6][nv]: second hyperdiffusion array for HyDiffOrder >= 6 (passes alternate)
    double *diffmv; // [3][nv]: vertical momentum hyperdiffusion (FP64 form with VertHyDiff; else null)
    double *dsp;    // [6][nv]: the hyperdiffusion pass output the diffusive sponge reads (else null)
    float *mdsp;
    float *mdiffx;
    float *mhpt;  // [2][nv]: h and pt of the stage rounded to FP32, for the substep flux divergences
};

// Storage of the mode-dependent fields: FP64 in the FP64 path, FP32 in the
// mixed-precision path.
template <bool MX> struct FS;
template <> struct FS<false> {
    using R = double;
    static __host__ __device__ double *Mhs(const Fields &f) { return f.Mhs; }
    static __host__ __device__ double *Adv(const Fields &f) { return f.Adv; }
    static __host__ __device__ double *diff(const Fields &f) { return f.diff; }
    static __host__ __device__ double *diffx(const Fields &f) { return f.diffx; }
    static __host__ __device__ void set_diff(Fields &f, double *a) { f.diff = a; }
    static __host__ __device__ const double *dsp(const Fields &f) { return f.dsp; }
    static __host__ __device__ void set_dsp(Fields &f, double *a) { f.dsp = a; }
    static __host__ __device__ double *diffmh(const Fields &f) { return f.diffmh; }
    static __host__ __device__ double *diffw(const Fields &f) { return f.diffw; }
    static __host__ __device__ double *diffrh(const Fields &f) { return f.diffrh; }
    static __host__ __device__ double *diffpr(const Fields &f) { return f.diffpr; }
    static __host__ __device__ double *divg(const Fields &f) { return f.divg; }
    static __host__ __device__ double *SlowMh(const Fields &f) { return f.SlowMh; }
};
template <> struct FS<true> {
    using R = float;
    static __host__ __device__ float *Mhs(const Fields &f) { return f.mMhs; }
    static __host__ __device__ float *Adv(const Fields &f) { return f.mAdv; }
    static __host__ __device__ float *diff(const Fields &f) { return f.mdiff; }
    static __host__ __device__ float *diffx(const Fields &f) { return f.mdiffx; }
    static __host__ __device__ void set_diff(Fields &f, float *a) { f.mdiff = a; }
    static __host__ __device__ const float *dsp(const Fields &f) { return f.mdsp; }
    static __host__ __device__ void set_dsp(Fields &f, float *a) { f.mdsp = a; }
    static __host__ __device__ float *diffmh(const Fields &f) { return f.mdiffmh; }
    static __host__ __device__ float *diffw(const Fields &f) { return f.mdiffw; }
    static __host__ __device__ float *diffrh(const Fields &f) { return f.mdiffrh; }
    static __host__ __device__ float *diffpr(const Fields &f) { return f.mdiffpr; }
    static __host__ __device__ float *divg(const Fields &f) { return f.mdivg; }
    static __host__ __device__ float *SlowMh(const Fields &f) { return f.mSlowMh; }
};

enum : unsigned { FLAG_NAN_P = 1u, FLAG_NAN_T = 2u };

} // namespace thor
