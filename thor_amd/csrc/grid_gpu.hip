// Spring dynamics of the icosahedral grid on the GPU (Tomita et al. 2001, as in
// THOR's Icogrid::spring_dynamics, grid.cu:1670-1815): the same forces,
// explicit update, fixed pentagons and stopping rule as the host loop in
// grid.cpp, one kernel per iteration over all points.
//
// Positions and velocities are double-buffered (iteration it reads buffer
// it % 2 and writes the other), so each launch sees the previous iteration's
// positions as the host loop's Pold. The largest net force of iteration it
// (pentagons excluded) goes to slot[it] by an atomic max on its bits
// (non-negative doubles order like their bit patterns); iteration it + 1
// returns at once when slot[it] is below the threshold, so the positions
// freeze at the converging iteration exactly as the host loop's break. The
// host launches the iterations in batches and reads the slots back between
// batches to find that iteration.
//
// The arithmetic is the host loop's; acos and sqrt are the device's, so the
// positions can differ from the host loop's in the last bits (the system is
// strongly damped: the difference does not grow).
#include "grid.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace thor {
namespace {

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator*(D3 a, double b) { return {a.x * b, a.y * b, a.z * b}; }
__device__ __forceinline__ D3 operator*(double b, D3 a) { return {b * a.x, b * a.y, b * a.z}; }
__device__ __forceinline__ D3 operator/(D3 a, double b) { return {a.x / b, a.y / b, a.z / b}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double length(D3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ D3 ld3(const double *p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void st3(double *p, int i, D3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

struct SpringArgs {
    const int *nb;       // [P][6] neighbours (point_local)
    const int *pent;     // [P] 1 for the 12 pentagons (fixed)
    const double *fixed; // [P][3] the pentagons' positions
    double *xyz[2], *vel[2];
    unsigned long long *slot;  // [lim] bits of each iteration's largest net force
    int P;
    double lba, dbar, drag, tstep, cond;
};

__global__ __launch_bounds__(256) void k_spring(SpringArgs a, int it) {
    if (it > 0 && __longlong_as_double((long long)a.slot[it - 1]) < a.cond) return;  // converged earlier
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = it & 1;
    const double *X = a.xyz[b], *Vl = a.vel[b];
    double fnet = 0.0;
    if (i < a.P) {
        D3 xi = ld3(X, i), out_x, out_v;
        if (a.pent[i]) {
            out_x = ld3(a.fixed, i);
            out_v = {0.0, 0.0, 0.0};
        } else {
            D3 Fn[6];
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const D3 pj = ld3(X, a.nb[i * 6 + j]);
                const D3 W = cross(xi, pj);
                const D3 V = cross(W, xi);
                const double l = length(V);
                const double d = acos(dot(xi, pj));
                Fn[j] = ((d - a.dbar) * V) / l;
            }
            const D3 Fsum = Fn[0] + Fn[1] + Fn[2] + Fn[3] + Fn[4] + Fn[5];
            fnet = length(Fsum) / a.lba;
            const D3 v0 = ld3(Vl, i);
            const D3 F = Fsum - a.drag * v0;
            const D3 D = xi + v0 * a.tstep;
            out_x = D / length(D);
            const D3 vv = v0 + F * a.tstep;
            const double H = dot(out_x, vv);
            out_v = vv - H * out_x;
        }
        st3(a.xyz[b ^ 1], i, out_x);
        st3(a.vel[b ^ 1], i, out_v);
    }
    // block maximum of the net force, then one atomic max per block
    __shared__ double red[256];
    red[threadIdx.x] = fnet;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(&a.slot[it], (unsigned long long)__double_as_longlong(red[0]));
}

} // namespace

bool spring_dynamics_device(IcoGrid &g, std::string *err) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        if (err) *err = "no HIP device";
        return false;
    }
    const int P = g.point_num;
    const int lim = 100000;
    // pentagons: their ring's sixth entry repeats the fifth (5 distinct neighbours)
    std::vector<int> nb(g.point_local), pent(P, 0);
    std::vector<double> fixed((size_t)3 * P, 0.0);
    for (int k = 0; k < 12; k++) {
        const int p = g.pent_ind[k];
        pent[p] = 1;
        for (int c = 0; c < 3; c++) fixed[3 * p + c] = g.xyz[3 * p + c];
    }
    SpringArgs a{};
    a.P = P;
    a.lba = 2.0 * M_PI / (10.0 * std::pow(2.0, g.glevel - 1));
    a.dbar = g.beta * a.lba;
    a.drag = 1;
    a.tstep = 2.0E-2;
    a.cond = 1.0E-4;
    std::vector<void *> mem;
    bool ok = true;
    auto alloc = [&](void **p, size_t bytes) {
        if (ok && hipMalloc(p, bytes) == hipSuccess) mem.push_back(*p);
        else ok = false;
    };
    int *d_nb = nullptr, *d_pent = nullptr;
    double *d_fixed = nullptr;
    alloc((void **)&d_nb, nb.size() * sizeof(int));
    alloc((void **)&d_pent, pent.size() * sizeof(int));
    alloc((void **)&d_fixed, fixed.size() * sizeof(double));
    for (int b = 0; b < 2; b++) {
        alloc((void **)&a.xyz[b], (size_t)3 * P * sizeof(double));
        alloc((void **)&a.vel[b], (size_t)3 * P * sizeof(double));
    }
    alloc((void **)&a.slot, (size_t)lim * sizeof(unsigned long long));
    hipStream_t st = nullptr;
    ok = ok && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMemcpyAsync(d_nb, nb.data(), nb.size() * sizeof(int), hipMemcpyHostToDevice, st) == hipSuccess
         && hipMemcpyAsync(d_pent, pent.data(), pent.size() * sizeof(int), hipMemcpyHostToDevice, st) == hipSuccess
         && hipMemcpyAsync(d_fixed, fixed.data(), fixed.size() * sizeof(double), hipMemcpyHostToDevice, st)
                == hipSuccess
         && hipMemcpyAsync(a.xyz[0], g.xyz.data(), (size_t)3 * P * sizeof(double), hipMemcpyHostToDevice, st)
                == hipSuccess
         && hipMemsetAsync(a.vel[0], 0, (size_t)3 * P * sizeof(double), st) == hipSuccess
         && hipMemsetAsync(a.slot, 0, (size_t)lim * sizeof(unsigned long long), st) == hipSuccess;
    a.nb = d_nb;
    a.pent = d_pent;
    a.fixed = d_fixed;
    int conv = -1;  // the converging iteration (the host loop's final `it`)
    if (ok) {
        const unsigned grid = (unsigned)((P + 255) / 256);
        const int batch = 1000;
        std::vector<unsigned long long> s(batch);
        for (int it0 = 0; it0 < lim && conv < 0 && ok; it0 += batch) {
            const int n = std::min(batch, lim - it0);
            for (int it = it0; it < it0 + n; it++) k_spring<<<grid, 256, 0, st>>>(a, it);
            ok = hipGetLastError() == hipSuccess
                 && hipMemcpyAsync(s.data(), a.slot + it0, n * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   st) == hipSuccess
                 && hipStreamSynchronize(st) == hipSuccess;
            for (int k = 0; k < n && ok; k++) {
                double v;
                std::memcpy(&v, &s[k], sizeof v);
                if (v < a.cond) {
                    conv = it0 + k;
                    break;
                }
            }
        }
        if (conv < 0) conv = lim;  // not converged: the host loop stops at lim too
    }
    if (ok) {
        // iteration conv wrote buffer (conv + 1) % 2 (after lim iterations: lim % 2)
        const int b = conv < lim ? (conv + 1) & 1 : lim & 1;
        ok = hipMemcpyAsync(g.xyz.data(), a.xyz[b], (size_t)3 * P * sizeof(double), hipMemcpyDeviceToHost, st)
                 == hipSuccess
             && hipStreamSynchronize(st) == hipSuccess;
        g.spring_iterations = conv;
    }
    if (st) (void)hipStreamDestroy(st);
    for (void *p : mem) (void)hipFree(p);
    if (!ok && err) *err = "HIP error in the spring-dynamics kernels";
    return ok;
}

} // namespace thor
