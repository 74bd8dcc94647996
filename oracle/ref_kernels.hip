// ORACLE — test infrastructure only.
//
// Runs THOR's own kernels, compiled for gfx950 straight from the reference
// tree, so tests can check the serial restatement in thor_oracle.cpp kernel
// by kernel. The kernels come from the reference headers
//   src/headers/dyn/thor_adv_cor.h    Compute_Advec_Cori1, _Poles, Compute_Advec_Cori2
//   src/headers/dyn/thor_diff.h       Diffusion_Op, _Poles, Correct_Horizontal, vertical_diff
//   src/headers/dyn/thor_div.h        DivM_Op, _Poles
//   src/headers/dyn/thor_slowmodes.h  Compute_Slow_Modes, _Poles
//   src/headers/phy/profx_held_suarez.h  held_suarez
//   src/physics/modules/inc/thor_chemistry.h  Tracer_Eq, Tracer_Eq_Diffusion, _Poles
//                                     (kRefTracers tracers; the first diffusion
//                                     pass goes to "diff", as in chemistry.cu)
// and are launched with the geometry of src/ESP/thor_driver.cu:70-95 and
// src/ESP/profx_driver.cu:79,303-314, on device copies of named arrays in
// THOR's own layout (the oracle's names). Built by `make -C oracle ref`
// into oracle/_ref/ only where the reference tree is present; nothing in the
// product loads it.
#include <hip/hip_runtime.h>

#include "dyn/thor_adv_cor.h"
#include "dyn/thor_diff.h"
#include "dyn/thor_div.h"
#include "dyn/thor_slowmodes.h"
#include "phy/profx_held_suarez.h"
#include "thor_chemistry.h"

#include "thor_oracle.h"

#include <map>
#include <string>
#include <vector>

namespace {

constexpr int kRefTracers = 2;

struct RefKernels {
    oracle_params p{};
    int P = 0, nv = 0, nr = 0;
    std::map<std::string, double *> dev;
    std::map<std::string, long> size;
    int *maps = nullptr, *point_local = nullptr;
    bool ok = true;

    double *add(const std::string &n, long count, const double *host = nullptr) {
        double *d = nullptr;
        if (hipMalloc((void **)&d, count * sizeof(double)) != hipSuccess) {
            ok = false;
            return nullptr;
        }
        if (host) ok &= hipMemcpy(d, host, count * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
        else ok &= hipMemset(d, 0, count * sizeof(double)) == hipSuccess;
        dev[n] = d;
        size[n] = count;
        return d;
    }
    double *operator[](const char *n) { return dev.at(n); }
    ~RefKernels() {
        for (auto &kv : dev) (void)hipFree(kv.second);
        (void)hipFree(maps);
        (void)hipFree(point_local);
    }
};

int launch_status() {
    if (hipGetLastError() != hipSuccess) return -2;
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

} // namespace

extern "C" {

void *ref_create(const oracle_grid *g, const oracle_params *p) {
    auto *r = new RefKernels();
    r->p = *p;
    r->P = g->point_num;
    r->nv = g->nv;
    r->nr = g->nr;
    const long P = r->P, N = P * r->nv, Ni = P * (r->nv + 1);
    // geometry (grid.h:63-129 layouts)
    r->add("div", 21 * P, g->div);
    r->add("grad", 21 * P, g->grad);
    r->add("areas", 18 * P, g->areas);
    r->add("areasTr", 6 * P, g->areasTr);
    r->add("areasT", P, g->areasT);
    r->add("nvecoa", 18 * P, g->nvecoa);
    r->add("nvecti", 18 * P, g->nvecti);
    r->add("nvecte", 18 * P, g->nvecte);
    r->add("func_r", 3 * P, g->func_r);
    r->add("lonlat", 2 * P, g->lonlat);
    r->add("Altitude", r->nv, g->Altitude);
    r->add("Altitudeh", r->nv + 1, g->Altitudeh);
    const long nmaps = (long)r->nr * 18 * 18;
    r->ok &= hipMalloc((void **)&r->maps, nmaps * sizeof(int)) == hipSuccess;
    r->ok &= hipMalloc((void **)&r->point_local, 6 * P * sizeof(int)) == hipSuccess;
    if (r->ok) {
        r->ok &= hipMemcpy(r->maps, g->maps, nmaps * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
        r->ok &= hipMemcpy(r->point_local, g->point_local, 6 * P * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    }
    // state and work arrays, zero until set
    for (const char *n : {"W", "Rho", "pressure", "temperature", "Rd", "Cp", "profx_Qheat", "Wk", "Rhok", "pressurek",
                          "h", "pt", "gtil", "SlowRho", "Slowpressure", "diffw", "diffrh", "diffpr", "diffwv",
                          "diffrv", "diffprv"})
        r->add(n, N);
    for (const char *n : {"Wh", "profx_dWh", "Whk", "hh", "pth", "gtilh", "SlowWh"}) r->add(n, Ni);
    for (const char *n : {"Mh", "profx_dMh", "Mhk", "Adv", "v", "SlowMh", "diffmh", "diffmv", "DivM", "divg_Mh", "Mhs"})
        r->add(n, 3 * N);
    r->add("Rhos", N);
    r->add("Whs", Ni);
    for (const char *n : {"tracerk", "tracers", "difftr"}) r->add(n, kRefTracers * N);
    for (const char *n : {"diff", "diff_sponge", "boundary_flux"}) r->add(n, 6 * N);
    for (const char *n : {"Kdh4", "Kdhz", "Kdh2", "Kdv6"}) r->add(n, r->nv);
    r->add("diff2", 6 * P * (r->nv + 2));  // vertical_diff's ghost-extended columns
    if (!r->ok) {
        delete r;
        return nullptr;
    }
    return r;
}

void ref_destroy(void *h) { delete static_cast<RefKernels *>(h); }

long ref_array_size(void *h, const char *name) {
    auto *r = static_cast<RefKernels *>(h);
    auto it = r->size.find(name);
    return it == r->size.end() ? -1 : it->second;
}

int ref_set(void *h, const char *name, const double *src, long n) {
    auto *r = static_cast<RefKernels *>(h);
    auto it = r->size.find(name);
    if (it == r->size.end() || it->second != n) return -1;
    return hipMemcpy(r->dev[name], src, n * sizeof(double), hipMemcpyHostToDevice) == hipSuccess ? 0 : -2;
}

int ref_get(void *h, const char *name, double *dst, long n) {
    auto *r = static_cast<RefKernels *>(h);
    auto it = r->size.find(name);
    if (it == r->size.end() || it->second != n) return -1;
    return hipMemcpy(dst, r->dev[name], n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess ? 0 : -2;
}

// Same names and arguments as oracle_kernel().
int ref_kernel(void *h, const char *name, int a0, int a1) {
    auto &r = *static_cast<RefKernels *>(h);
    const std::string n(name);
    const int P = r.P, nv = r.nv, NTH = 256;
    const double A = r.p.A;
    const bool deep = r.p.DeepModel != 0, nonhydro = r.p.NonHydro != 0, ghv = r.p.GravHeightVar != 0;
    const dim3 NT(16, 16, 1), NB(r.nr, nv, 1), NBD(r.nr, nv, 6), NBDP(2, nv, 6), NBP(2, nv, 1);
    const dim3 NBALL1(P / NTH + 1, nv, 1);
    auto d3 = [&](const char *k) { return (double3 *)r[k]; };
    if (n == "advec_cori1") {
        Compute_Advec_Cori1<16, 16><<<NB, NT>>>(d3("Adv"), d3("v"), d3("Mhk"), d3("div"), r["Wk"], r["Rhok"],
                                                r["Altitude"], A, d3("func_r"), r.maps, 16, deep);
        Compute_Advec_Cori_Poles<6><<<2, 1>>>(r["Adv"], r["v"], r["Mhk"], r["div"], r["Wk"], r["Rhok"], r["Altitude"],
                                              A, r["func_r"], r.point_local, P, nv, deep);
    } else if (n == "advec_cori2") {
        Compute_Advec_Cori2<<<P / NTH + 1, NTH>>>(r["Adv"], r["v"], r["Whk"], r["Rhok"], r["Altitude"], r["Altitudeh"],
                                                  r["func_r"], r.p.Omega, A, nv, P, deep);
    } else if (n == "diffusion") {
        const int stepnum = a0;
        const bool last = a1 != 0;
        Diffusion_Op<16, 16><<<NBD, NT>>>(r["diffmh"], r["diffw"], r["diffrh"], r["diffpr"], r["diff"],
                                          r["diff_sponge"], r["Mhk"], r["Rhok"], r["temperature"], r["Wk"],
                                          r["areasTr"], r["areas"], r["nvecoa"], r["nvecti"], r["nvecte"],
                                          r["func_r"], r["Kdh4"], r["Altitude"], A, r["Rd"], r["Cp"], r.maps, 16,
                                          stepnum, last, deep, r.p.DiffSponge != 0, r.p.order_diff_sponge,
                                          r["Kdh2"], r["boundary_flux"], r.p.energy_equation != 0, r.p.HyDiffOrder);
        Diffusion_Op_Poles<5><<<NBDP, 1>>>(r["diffmh"], r["diffw"], r["diffrh"], r["diffpr"], r["diff"],
                                           r["diff_sponge"], r["Mhk"], r["Rhok"], r["temperature"], r["Wk"],
                                           r["func_r"], r["areasTr"], r["areas"], r["nvecoa"], r["nvecti"],
                                           r["nvecte"], r["Kdh4"], r["Altitude"], r["Altitudeh"], A, r["Rd"],
                                           r["Cp"], r.point_local, P, stepnum, last, deep, r.p.DiffSponge != 0,
                                           r.p.order_diff_sponge, r["Kdh2"], r["boundary_flux"], r.p.energy_equation != 0,
                                           r.p.HyDiffOrder);
    } else if (n == "vertical_diff") {
        const dim3 NBALL0(P / NTH + 1, 1, 6);  // thor_driver.cu:86
        vertical_diff<<<NBALL0, NTH>>>(r["diffmv"], r["diffwv"], r["diffrv"], r["diffprv"], r["diff"], r["diff2"],
                                       r["Mhk"], r["Rhok"], r["temperature"], r["Wk"], r["Kdv6"], r["Altitude"],
                                       r["Altitudeh"], A, r.p.Rd, r.p.VertHyDiffOrder, deep, P, nv);
    } else if (n == "correct_horizontal") {
        Correct_Horizontal<<<NBALL1, NTH>>>(r["diffmh"], r["diffmv"], r["func_r"], P);
    } else if (n == "divm") {
        const bool last = a0 != 0;
        DivM_Op<16, 16><<<NB, NT>>>(r["DivM"], r["divg_Mh"], r["Mhk"], r["Whk"], r["Kdhz"], r["areasTr"], r["areas"],
                                    r["nvecoa"], r["nvecti"], r["nvecte"], r["func_r"], r["Altitudeh"],
                                    r["Altitude"], A, r.maps, 16, last, deep);
        DivM_Op_Poles<5><<<NBP, 1>>>(r["DivM"], r["divg_Mh"], r["Mhk"], r["Whk"], r["Kdhz"], r["areasTr"],
                                     r["areas"], r["nvecoa"], r["nvecti"], r["nvecte"], r["func_r"], r["Altitudeh"],
                                     r["Altitude"], A, r.point_local, P, last, deep);
    } else if (n == "slow_modes") {
        Compute_Slow_Modes<16, 16><<<NB, NT>>>(
            r["SlowMh"], r["SlowWh"], r["SlowRho"], r["Slowpressure"], r["Mhk"], r["Whk"], r["Rhok"], r["Adv"],
            r["DivM"], r["diffmh"], r["diffw"], r["diffrh"], r["diffpr"], r["diffmv"], r["diffwv"], r["diffrv"],
            r["diffprv"], r["pressurek"], r["h"], r["hh"], r["gtil"], r["grad"], r["div"], r["Altitude"],
            r["Altitudeh"], A, r.p.Gravit, r["Cp"], r["Rd"], r["func_r"], r.maps, 16, deep, nonhydro,
            r["profx_dMh"], r["profx_dWh"], r["profx_Qheat"], r.p.energy_equation != 0, ghv);
        Compute_Slow_Modes_Poles<6><<<2, 1>>>(
            r["SlowMh"], r["SlowWh"], r["SlowRho"], r["Slowpressure"], r["Mhk"], r["Whk"], r["Rhok"], r["Adv"],
            r["DivM"], r["diffmh"], r["diffw"], r["diffrh"], r["diffpr"], r["diffmv"], r["diffwv"], r["diffrv"],
            r["diffprv"], r["pressurek"], r["h"], r["hh"], r["gtil"], r["grad"], r["div"], r["Altitude"],
            r["Altitudeh"], A, r.p.Gravit, r["Cp"], r["Rd"], r["func_r"], r.point_local, nv, P, deep, nonhydro,
            r["profx_dMh"], r["profx_dWh"], r["profx_Qheat"], r.p.energy_equation != 0, ghv);
    } else if (n == "tracer_diffusion") {  // a0: laststep (chemistry.cu:345-427)
        const dim3 NBTR(r.nr, nv, kRefTracers), NBTRP(2, nv, kRefTracers);
        Tracer_Eq_Diffusion<16, 16><<<NBTR, NT>>>(r["difftr"], r["diff"], r["tracerk"], r["Rhok"], r["areasTr"],
                                                  r["nvecoa"], r["nvecti"], r["nvecte"], r["Kdh4"], r["Altitude"], A,
                                                  r.maps, kRefTracers, 16, a0, deep);
        Tracer_Eq_Diffusion_Poles<5><<<NBTRP, 1>>>(r["difftr"], r["diff"], r["tracerk"], r["Rhok"], r["areasTr"],
                                                   r["nvecoa"], r["nvecti"], r["nvecte"], r["Kdh4"], r["Altitude"],
                                                   r["Altitudeh"], A, r.point_local, kRefTracers, P, a0, deep);
    } else if (n == "tracer_eq") {  // a0: the substep length in milliseconds (chemistry.cu:429-492)
        const dim3 NBTR(r.nr, nv, kRefTracers), NBPT(2, 1, kRefTracers);
        const double dt = a0 * 1e-3;
        Tracer_Eq<16, 16><<<NBTR, NT>>>(r["tracers"], r["tracerk"], r["Rhos"], r["Rhok"], r["Mhs"], r["Mhk"], r["Whs"],
                                        r["Whk"], r["difftr"], r["div"], r["Altitude"], r["Altitudeh"], A, dt, r.maps,
                                        kRefTracers, 16, deep);
        Tracer_Eq_Poles<6><<<NBPT, 1>>>(r["tracers"], r["tracerk"], r["Rhos"], r["Rhok"], r["Mhs"], r["Mhk"],
                                        r["Whs"], r["Whk"], r["difftr"], r["div"], r["Altitude"], r["Altitudeh"], A,
                                        dt, r.point_local, kRefTracers, P, nv, deep);
    } else if (n == "held_suarez") {
        held_suarez<<<NBALL1, NTH>>>(r["Mh"], r["pressure"], r["Rho"], r["temperature"], r.p.Gravit, r.p.Cp, r.p.Rd,
                                     r["Altitude"], r["Altitudeh"], r["lonlat"], r.p.timestep, P);
    } else {
        return -1;
    }
    return launch_status();
}

} // extern "C"
