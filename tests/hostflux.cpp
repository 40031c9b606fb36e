// hostflux.cpp - host build of the device flux functions, for CPU tests.
//
// Compiled by tests/test_flux_host.py with hipcc's host pass only: the
// __host__ __device__ functions of flux.hpp and indicator.hpp run on the CPU,
// so the closed-form and characteristic linear fluxes of the learned hybrid
// and the fast WENO5 flux can be checked against the CPU oracle without a GPU.
#include "indicator.hpp"

namespace {

artra::FluxModel Model(int roe, int sw)
{
    artra::FluxModel m{};
    m.gamma = 1.4;
    m.roe = roe;
    m.sw = sw;
    return m;
}

template <int S>
void Linear(int roe, int sw, int closed, const double st[6][5], double F[5])
{
    const artra::FluxModel mdl = Model(roe, sw);
    artra::FaceState fs;
    artra::face_state<S>(mdl, st[2], st[3], fs);
    if (closed) {
        double Sg[5], Dl[5];
        artra::linear_sum_diff(st, Sg, Dl);
        artra::linear_flux_llf_sd<S>(mdl.gamma, fs, Sg, Dl, F);
    } else {
        artra::linear_flux<S>(fs, st, F);
    }
}

template <int S>
void Weno(int roe, int sw, const double st[6][5], double F[5])
{
    artra::fast::convective_flux<S, 6>(Model(roe, sw), st, F);
}

}  // namespace

extern "C" {

// Linear fifth-order upwind flux at the face between stencil nodes 2 and 3:
// closed = 0 the characteristic form (every split), closed = 1 the closed
// form of the local Lax-Friedrichs split.
void host_linear_flux(int s, int roe, int sw, int closed, const double *st, double *F)
{
    const double(*w)[5] = reinterpret_cast<const double(*)[5]>(st);
    if (s == 0) Linear<0>(roe, sw, closed, w, F);
    else if (s == 1) Linear<1>(roe, sw, closed, w, F);
    else Linear<2>(roe, sw, closed, w, F);
}

// The two linear combinations of the six-node stencil as their sum and
// difference (per component), as the routed faces form them.
void host_linear_sums(const double *st, double *Sg, double *Dl)
{
    artra::linear_sum_diff(reinterpret_cast<const double(*)[5]>(st), Sg, Dl);
}

// The fast-arithmetic WENO5 flux of the plain sweeps.
void host_weno_flux(int s, int roe, int sw, const double *st, double *F)
{
    const double(*w)[5] = reinterpret_cast<const double(*)[5]>(st);
    if (s == 0) Weno<0>(roe, sw, w, F);
    else if (s == 1) Weno<1>(roe, sw, w, F);
    else Weno<2>(roe, sw, w, F);
}

}  // extern "C"
