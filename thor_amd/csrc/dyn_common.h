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
    double A, Omega, Gravit, P_Ref, Rd, Cp;
    int deep, nonhydro, ghv;
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
    double *Adv;                          // [3][nv]
    double *diff;                         // [6][nv]
    double *diffmh;                       // [3][nv]
    double *diffw, *diffrh, *diffpr;      // [nv]
    double *divg;                         // [3][nv]
    double *SlowMh, *SlowWh, *SlowRho, *Slowp;
    double *Sp, *Sd;                      // implicit-solve sources [nv]
    double *tAA, *tCC, *tTT;              // Thomas factorisation of the stage [nv]
    double *Qheat, *dMh, *dWh;            // ProfX tendencies
    unsigned int *flags;                  // bit 0: NaN pressure, bit 1: NaN temperature
};

enum : unsigned { FLAG_NAN_P = 1u, FLAG_NAN_T = 2u };

} // namespace thor
