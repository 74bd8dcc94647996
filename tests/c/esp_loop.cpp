// A THOR-style driver loop over the C ABI (include/thor_dyn.h), as a
// maintainer would bind it inside ESP: the caller owns the model state and
// the forcing arrays in device memory, in THOR's layout (the ESP members
// Mh_d, Wh_d, W_d, Rho_d, pressure_d and profx_Qheat_d, profx_dMh_d,
// profx_dWh_d, src/headers/esp.h:224-371), and calls the step once per time
// step where esp.cu:1395-1399 calls ESP::Thor. The forcing is the caller's
// own (a fixed heating, momentum and vertical-momentum tendency formed on the
// host, standing in for THOR's ProfX modules).
//
// Usage: esp_loop <grid cache dir> <precision 0|1> <steps> <output file>
// Output: the forcing arrays, then the state after the loop, each as
// (uint64 count, count doubles) in the order Qheat, dMh, dWh, Mh, Wh, W, Rho,
// pressure.
#include "thor_dyn.h"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK_HIP(x)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) {                                                                        \
            std::fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_));     \
            return 3;                                                                                  \
        }                                                                                              \
    } while (0)
#define CHECK_THOR(x)                                                                                  \
    do {                                                                                               \
        thor_status s_ = (x);                                                                          \
        if (s_ != THOR_OK) {                                                                           \
            std::fprintf(stderr, "%s:%d %s: status %d: %s\n", __FILE__, __LINE__, #x, s_, thor_last_error()); \
            return 4;                                                                                  \
        }                                                                                              \
    } while (0)

static bool put(std::FILE *f, const std::vector<double> &v) {
    const uint64_t n = v.size();
    return std::fwrite(&n, sizeof n, 1, f) == 1 && std::fwrite(v.data(), sizeof(double), n, f) == n;
}

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s <grid cache dir> <precision 0|1> <steps> <output file>\n", argv[0]);
        return 2;
    }
    const char *cache = argv[1];
    const int precision = std::atoi(argv[2]), nsteps = std::atoi(argv[3]);

    thor_grid *grid = nullptr;
    CHECK_THOR(thor_grid_build(4, 20, 6.371e6, 36000.0, 1, 1.15, cache, &grid));
    thor_grid_desc gd;
    CHECK_THOR(thor_grid_describe(grid, &gd));
    thor_params prm;
    thor_params_default_earth(&prm);
    prm.glevel = 4;
    prm.timestep = 600.0;
    prm.n_substeps = 6;
    prm.precision = precision;
    const int P = gd.point_num, nv = gd.nv;
    const size_t n1 = (size_t)P * nv, ni = (size_t)P * (nv + 1), n3 = 3 * n1;

    // initial state on the host, then in the caller's device arrays
    std::vector<double> Mh(n3), Wh(ni), W(n1), Rho(n1), pr(n1);
    CHECK_THOR(thor_ic_rest(&gd, &prm, 1016, 1e-3, 1.0, Mh.data(), Wh.data(), W.data(), Rho.data(), pr.data()));
    double *Mh_d, *Wh_d, *W_d, *Rho_d, *pressure_d, *Qheat_d, *dMh_d, *dWh_d;
    CHECK_HIP(hipMalloc((void **)&Mh_d, n3 * sizeof(double)));
    CHECK_HIP(hipMalloc((void **)&Wh_d, ni * sizeof(double)));
    CHECK_HIP(hipMalloc((void **)&W_d, n1 * sizeof(double)));
    CHECK_HIP(hipMalloc((void **)&Rho_d, n1 * sizeof(double)));
    CHECK_HIP(hipMalloc((void **)&pressure_d, n1 * sizeof(double)));
    CHECK_HIP(hipMalloc((void **)&Qheat_d, n1 * sizeof(double)));
    CHECK_HIP(hipMalloc((void **)&dMh_d, n3 * sizeof(double)));
    CHECK_HIP(hipMalloc((void **)&dWh_d, ni * sizeof(double)));
    CHECK_HIP(hipMemcpy(Mh_d, Mh.data(), n3 * sizeof(double), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(Wh_d, Wh.data(), ni * sizeof(double), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(W_d, W.data(), n1 * sizeof(double), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(Rho_d, Rho.data(), n1 * sizeof(double), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(pressure_d, pr.data(), n1 * sizeof(double), hipMemcpyHostToDevice));

    // the caller's forcing: heating peaked at the equator and decaying with
    // height, an eastward momentum tendency, a small interior dWh
    std::vector<double> Q(n1), dMh(n3), dWh(ni, 0.0);
    for (int p = 0; p < P; p++) {
        const double lon = gd.lonlat[2 * p], lat = gd.lonlat[2 * p + 1];
        const double ex = -std::sin(lon), ey = std::cos(lon);  // eastward unit vector (no z part)
        for (int lev = 0; lev < nv; lev++) {
            const double z = gd.Altitude[lev];
            Q[(size_t)p * nv + lev] = 2e-3 * std::cos(lat) * std::exp(-z / 8000.0);
            const double f = 1e-6 * std::cos(lat) * Rho[(size_t)p * nv + lev];
            dMh[((size_t)p * nv + lev) * 3 + 0] = f * ex;
            dMh[((size_t)p * nv + lev) * 3 + 1] = f * ey;
            dMh[((size_t)p * nv + lev) * 3 + 2] = 0.0;
        }
        for (int lev = 1; lev < nv; lev++)
            dWh[(size_t)p * (nv + 1) + lev] = 1e-8 * std::sin(M_PI * lev / nv) * std::sin(lat);
    }
    CHECK_HIP(hipMemcpy(Qheat_d, Q.data(), n1 * sizeof(double), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(dMh_d, dMh.data(), n3 * sizeof(double), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(dWh_d, dWh.data(), ni * sizeof(double), hipMemcpyHostToDevice));

    thor_dyn *dyn = nullptr;
    CHECK_THOR(thor_dyn_create(&gd, &prm, &dyn));
    thor_state state = {1, Mh_d, Wh_d, W_d, Rho_d, pressure_d};
    const thor_forcing forcing = {Qheat_d, dMh_d, dWh_d};
    // the main loop: the step works in place on the caller's arrays, ordered
    // against the caller's (default) stream
    for (int step = 1; step <= nsteps; step++) CHECK_THOR(thor_dyn_step(dyn, &state, &forcing, 0.0, 0, nullptr));
    CHECK_THOR(thor_dyn_sync(dyn));
    CHECK_HIP(hipMemcpy(Mh.data(), Mh_d, n3 * sizeof(double), hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(Wh.data(), Wh_d, ni * sizeof(double), hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(W.data(), W_d, n1 * sizeof(double), hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(Rho.data(), Rho_d, n1 * sizeof(double), hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(pr.data(), pressure_d, n1 * sizeof(double), hipMemcpyDeviceToHost));
    thor_dyn_destroy(dyn);
    for (double *ptr : {Mh_d, Wh_d, W_d, Rho_d, pressure_d, Qheat_d, dMh_d, dWh_d}) CHECK_HIP(hipFree(ptr));
    thor_grid_destroy(grid);

    std::FILE *f = std::fopen(argv[4], "wb");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", argv[4]);
        return 5;
    }
    bool ok = true;
    for (const std::vector<double> *v : {&Q, &dMh, &dWh, &Mh, &Wh, &W, &Rho, &pr}) ok = ok && put(f, *v);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return 5;
    std::printf("esp_loop: %d steps of g4 L20 (precision %d) written to %s\n", nsteps, precision, argv[4]);
    return 0;
}
