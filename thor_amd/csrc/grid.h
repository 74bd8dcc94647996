// Icosahedral grid and discrete-operator builder (host side).
//
// Produces the same arrays, in the same per-point layout, that THOR's Icogrid
// hands to its dynamical core (src/headers/grid.h:63-129, src/ESP/grid.cu:85-302):
//   point_xyz[P][3], point_xyzq[P][6][3], point_local[P][6], maps[nr][18][18],
//   func_r[P][3], lonlat[P][2], areasT[P], areasTr[P][6], areas[P][6][3],
//   nvec/nvecoa/nvecti/nvecte[P][6][3], div[P][7][3], grad[P][7][3],
//   Altitude[nv], Altitudeh[nv+1].
//
// Point order is rhombus-major: fc*nfaces*256 + ky*kxl*256 + kx*256 + y*16 + x,
// with the north pole at P-2 and the south pole at P-1 (grid.cu:631-656).
//
// Neighbour discovery is topological (lattice adjacency, O(P)) instead of the
// reference's all-pairs distance search (grid.cu:663-724); the resulting
// neighbour order is the rhombus order of grid.cu:1495-1621 (pt1..pt6) and the
// pole order of grid.cu:919-1005.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace thor {

struct IcoGrid {
    int glevel    = 0;
    int nl_region = 16;
    int nfaces    = 1;  // sub-rhombi per main rhombus
    int kxl       = 1;  // sqrt(nfaces)
    int nr        = 0;  // number of 16x16 rhombi
    int point_num = 0;  // P
    int nv        = 0;
    double A      = 0.0;
    double top    = 0.0;
    bool spring   = true;
    double beta   = 1.15;
    int spring_iterations = 0;

    int pent_ind[12];
    std::vector<double> xyz, xyzq, func_r, lonlat;
    std::vector<int> point_local, maps;
    std::vector<double> areasT, areasTr, areas;
    std::vector<double> nvec, nvecoa, nvecti, nvecte, mvec;
    std::vector<double> div, grad;
    std::vector<double> Altitude, Altitudeh;

    bool is_pent(int i) const;
};

// Builds the horizontal structure (points, neighbours, maps, spring dynamics,
// q-points) on the unit sphere. Expensive at glevel >= 6 (spring dynamics);
// cache with save_horizontal/load_horizontal.
void build_horizontal(IcoGrid &g, int glevel, bool spring, double beta);

// Scales to radius A, builds areas, control vectors, div/grad operators,
// lon/lat and the uniform vertical grid (grid.cu:193-302).
void finish_grid(IcoGrid &g, double A, double top, int nv);

// Spring dynamics of build_horizontal on the current HIP device (grid_gpu.hip):
// false (and the reason in *err) when there is no device or a HIP call fails.
bool spring_dynamics_device(IcoGrid &g, std::string *err);

bool save_horizontal(const IcoGrid &g, const std::string &path);
bool load_horizontal(IcoGrid &g, const std::string &path, int glevel, bool spring, double beta);

} // namespace thor
