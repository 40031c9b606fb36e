// artra_device.hip - device side of libartra_hip for gfx950 (MI355X).
//
// Kernels
//   sweep_x        directional sweep along x (fluid_dynamics.c:157-274 for
//                  s = X): one thread per face, faces of consecutive lines
//                  flattened so every 256-thread block updates 255 nodes;
//                  face fluxes meet in LDS and each node takes its left and
//                  right flux from there (LU + SolveOperator fused).
//   sweep_march    sweeps along y and z: a thread walks a column segment,
//                  carrying the stencil window and the previous face flux in
//                  registers, so each face flux is computed once; loads are
//                  coalesced along x across the wavefront.
//   bc_group       a group of TreatBoundary boxes whose nodes do not
//                  interact (boundary_treatment.c:28-192).
//   ib_layer       one ghost layer of the immersed-boundary fill from the
//                  host-precomputed stencil table (immersed_boundary.c:270-487).
//   cfl_max        per-axis maximum characteristic speed (solve.c:108-153).
//
// All FP64 arithmetic follows the reference's operation order; the file is
// compiled with -ffp-contract=off so results match the CPU bit for bit.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "flux.hpp"

extern "C" {
#include "artra_device.h"
}

using artra::FluxModel;

namespace {

thread_local char gError[512];

int Fail(const char *fmt, ...)
{
    va_list args;
    va_start(args, fmt);
    vsnprintf(gError, sizeof gError, fmt, args);
    va_end(args);
    return 1;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return Fail("%s failed: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

constexpr int kXBlock = 256;        // threads (faces) per sweep_x block
constexpr int kMarchX = 64;         // x extent of a sweep_march block (one wavefront)
constexpr int kMarchT = 4;          // transverse extent of a sweep_march block
constexpr int kMaxGroup = 8;        // boxes per bc_group launch

struct Grid {
    int nx, ny, nz;
    long N;                         // nodes per component array
    int lo[3], hi[3];               // interior range
};

struct StageArgs {
    const double *Utn;              // level read by the stencils
    const double *Uto;              // level of the a-term
    double *Utm;                    // level written
    const uint8_t *fluid;
    double a, b, r;                 // U_tm = a U_to + b (U_tn + r Phi)
    int accumulate;                 // U_tm += b r Phi (unsplit, non-x)
    int toIsTn;                     // U_to and U_tn are the same level
    double dd[3];
};

__device__ __forceinline__ int Index(const Grid &g, int i, int j, int k)
{
    return (k * g.ny + j) * g.nx + i;
}

__device__ __forceinline__ void LoadNode(const double *U, long N, int idx, double out[5])
{
#pragma unroll
    for (int c = 0; c < 5; ++c) out[c] = U[c * N + idx];
}

// Solution operator of one node (fluid_dynamics.c:243-274).  FAST evaluates
// the same expression with fused multiply-adds.
template <bool VISC, bool FAST>
__device__ __forceinline__ void UpdateNode(const StageArgs &sa, long N, int idx, const double FL[5], const double FR[5],
                                           const double FvL[5], const double FvR[5], const double Un[5])
{
    double phi[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        if (FAST) phi[c] = VISC ? (FL[c] - FR[c]) + (FvR[c] - FvL[c]) : FL[c] - FR[c];
        else phi[c] = VISC ? ((FL[c] - FR[c]) + FvR[c]) - FvL[c] : ((FL[c] - FR[c]) + 0.0) - 0.0;
    }
    if (sa.accumulate) {
        const double br = sa.b * sa.r;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            sa.Utm[c * N + idx] = FAST ? fma(br, phi[c], sa.Utm[c * N + idx]) : sa.Utm[c * N + idx] + br * phi[c];
        }
        return;
    }
    double Uo[5];
    if (sa.toIsTn) {
#pragma unroll
        for (int c = 0; c < 5; ++c) Uo[c] = Un[c];
    } else {
        LoadNode(sa.Uto, N, idx, Uo);
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        if (FAST) sa.Utm[c * N + idx] = fma(sa.a, Uo[c], sa.b * fma(sa.r, phi[c], Un[c]));
        else sa.Utm[c * N + idx] = sa.a * Uo[c] + sa.b * (Un[c] + sa.r * phi[c]);
    }
}

template <int S, int NS, bool FAST>
__device__ __forceinline__ void ConvFlux(const FluxModel &mdl, const double st[NS][5], double F[5])
{
    if (FAST) artra::fast::convective_flux<S, NS>(mdl, st, F);
    else artra::convective_flux<S, NS>(mdl, st, F);
}

template <int S, bool FAST>
__device__ __forceinline__ void ViscFlux(const FluxModel &mdl, const double dd[3], const double P[5], const double E[5],
                                         const double nb[8][5], double F[5])
{
    if (FAST) artra::fast::viscous_flux<S>(mdl, dd, P, E, nb, F);
    else artra::viscous_flux<S>(mdl, dd, P, E, nb, F);
}

// Neighbour states of the viscous face stencil at (P, E = P + e_S).
template <int S>
__device__ __forceinline__ void LoadViscousNeighbours(const Grid &g, const double *U, int idxP, double nb[8][5])
{
    constexpr int TA = (S == 0) ? 1 : 0;
    constexpr int TB = (S == 2) ? 1 : 2;
    const int stride[3] = {1, g.nx, g.nx * g.ny};
    const int e = stride[S], a = stride[TA], b = stride[TB];
    LoadNode(U, g.N, idxP - a, nb[0]);
    LoadNode(U, g.N, idxP + a, nb[1]);
    LoadNode(U, g.N, idxP - b, nb[2]);
    LoadNode(U, g.N, idxP + b, nb[3]);
    LoadNode(U, g.N, idxP + e - a, nb[4]);
    LoadNode(U, g.N, idxP + e + a, nb[5]);
    LoadNode(U, g.N, idxP + e - b, nb[6]);
    LoadNode(U, g.N, idxP + e + b, nb[7]);
}

// ------------------------------------------------------------------ sweeps
template <int NS, bool VISC, bool FAST>
__global__ __launch_bounds__(kXBlock) void sweep_x(Grid g, StageArgs sa, FluxModel mdl)
{
    constexpr int SL = (NS == 6) ? -2 : -1;
    constexpr int K0 = -SL;
    __shared__ double sF[5][kXBlock];
    __shared__ double sFv[VISC ? 5 : 1][kXBlock];
    const int faces = g.hi[0] - g.lo[0] + 1;           // faces per x line
    const int ny = g.hi[1] - g.lo[1];
    const long total = (long)faces * ny * (g.hi[2] - g.lo[2]);
    const long gface = (long)blockIdx.x * (kXBlock - 1) + threadIdx.x;
    int f = 0, j = 0, k = 0;
    double Un[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (gface < total) {
        const long line = gface / faces;
        f = (int)(gface - line * faces);
        j = g.lo[1] + (int)(line % ny);
        k = g.lo[2] + (int)(line / ny);
        const int i = g.lo[0] - 1 + f;                 // face between i and i + 1
        const int base = Index(g, i, j, k);
        double st[NS][5];
#pragma unroll
        for (int m = 0; m < NS; ++m) LoadNode(sa.Utn, g.N, base + SL + m, st[m]);
        double F[5];
        ConvFlux<0, NS, FAST>(mdl, st, F);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            sF[c][threadIdx.x] = F[c];
            Un[c] = st[K0 + 1][c];                     // node i + 1, updated by this thread
        }
        if (VISC) {
            double nb[8][5], Fv[5];
            LoadViscousNeighbours<0>(g, sa.Utn, base, nb);
            ViscFlux<0, FAST>(mdl, sa.dd, st[K0], st[K0 + 1], nb, Fv);
#pragma unroll
            for (int c = 0; c < 5; ++c) sFv[VISC ? c : 0][threadIdx.x] = Fv[c];
        }
    }
    __syncthreads();
    // node between this thread's face and the next thread's face
    if (threadIdx.x < kXBlock - 1 && gface + 1 < total && f + 1 < faces) {
        const int i = g.lo[0] + f;
        const int idx = Index(g, i, j, k);
        if (sa.fluid[idx]) {
            double FL[5], FR[5], FvL[5], FvR[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                FL[c] = sF[c][threadIdx.x];
                FR[c] = sF[c][threadIdx.x + 1];
                FvL[c] = VISC ? sFv[VISC ? c : 0][threadIdx.x] : 0.0;
                FvR[c] = VISC ? sFv[VISC ? c : 0][threadIdx.x + 1] : 0.0;
            }
            UpdateNode<VISC, FAST>(sa, g.N, idx, FL, FR, FvL, FvR, Un);
        }
    }
}

template <int S, int NS, bool VISC, bool FAST>
__global__ __launch_bounds__(kMarchX *kMarchT) void sweep_march(Grid g, StageArgs sa, FluxModel mdl, int segment)
{
    constexpr int SL = (NS == 6) ? -2 : -1;
    constexpr int K0 = -SL;
    constexpr int T2 = (S == 1) ? 2 : 1;               // the other transverse axis
    const int i = g.lo[0] + blockIdx.x * kMarchX + threadIdx.x;
    const int t2 = g.lo[T2] + blockIdx.y * kMarchT + threadIdx.y;
    if (i >= g.hi[0] || t2 >= g.hi[T2]) return;
    const int s0 = g.lo[S] + blockIdx.z * segment;
    const int s1 = min(g.hi[S], s0 + segment);
    if (s0 >= s1) return;
    const int stride = (S == 1) ? g.nx : g.nx * g.ny;
    const int idx0 = (S == 1) ? Index(g, i, s0, t2) : Index(g, i, t2, s0);
    // window for the face between s0 - 1 and s0
    double st[NS][5];
#pragma unroll
    for (int m = 0; m < NS; ++m) LoadNode(sa.Utn, g.N, idx0 + (SL - 1 + m) * stride, st[m]);
    double FL[5], FvL[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    ConvFlux<S, NS, FAST>(mdl, st, FL);
    if (VISC) {
        double nb[8][5];
        LoadViscousNeighbours<S>(g, sa.Utn, idx0 - stride, nb);
        ViscFlux<S, FAST>(mdl, sa.dd, st[K0], st[K0 + 1], nb, FvL);
    }
    int idx = idx0;
    for (int s = s0; s < s1; ++s, idx += stride) {
#pragma unroll
        for (int m = 0; m < NS - 1; ++m) {
#pragma unroll
            for (int c = 0; c < 5; ++c) st[m][c] = st[m + 1][c];
        }
        LoadNode(sa.Utn, g.N, idx + (SL + NS - 1) * stride, st[NS - 1]);
        double FR[5], FvR[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        ConvFlux<S, NS, FAST>(mdl, st, FR);
        if (VISC) {
            double nb[8][5];
            LoadViscousNeighbours<S>(g, sa.Utn, idx, nb);
            ViscFlux<S, FAST>(mdl, sa.dd, st[K0], st[K0 + 1], nb, FvR);
        }
        if (sa.fluid[idx]) UpdateNode<VISC, FAST>(sa, g.N, idx, FL, FR, FvL, FvR, st[K0]);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            FL[c] = FR[c];
            FvL[c] = FvR[c];
        }
    }
}

// ---------------------------------------------------------------- boundary
struct BcBox {
    int lo[3], hi[3];
    int face, layer;
    long offset;                    // first flat index of this box in the group
};

struct BcGroup {
    BcBox box[kMaxGroup];
    int count;
    long total;
};

struct BcFace {
    int type;
    int N[3];
    double given[6];
};

struct BcArgs {
    double *U;                      // level base
    Grid g;
    int m[3];
    double gamma, gasR;
    BcFace face[7];
};

__device__ __forceinline__ void Primitive(double gamma, double gasR, const double U[5], double Uo[6])
{
    Uo[0] = U[0];
    Uo[1] = U[1] / U[0];
    Uo[2] = U[2] / U[0];
    Uo[3] = U[3] / U[0];
    Uo[4] = (U[4] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]) * (gamma - 1.0);
    Uo[5] = Uo[4] / (Uo[0] * gasR);
}

__device__ __forceinline__ void Conservative(double gamma, const double Uo[6], double U[5])
{
    U[0] = Uo[0];
    U[1] = Uo[0] * Uo[1];
    U[2] = Uo[0] * Uo[2];
    U[3] = Uo[0] * Uo[3];
    U[4] = 0.5 * Uo[0] * (Uo[1] * Uo[1] + Uo[2] * Uo[2] + Uo[3] * Uo[3]) + Uo[4] / (gamma - 1.0);
}

__device__ __forceinline__ void StoreNode(double *U, long N, int idx, const double v[5])
{
#pragma unroll
    for (int c = 0; c < 5; ++c) U[c * N + idx] = v[c];
}

__global__ __launch_bounds__(256) void bc_group(BcArgs ba, BcGroup grp)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= grp.total) return;
    int b = 0;
    while (b + 1 < grp.count && t >= grp.box[b + 1].offset) ++b;
    const BcBox &bx = grp.box[b];
    const long local = t - bx.offset;
    const int ex = bx.hi[0] - bx.lo[0], ey = bx.hi[1] - bx.lo[1];
    const int i = bx.lo[0] + (int)(local % ex);
    const int j = bx.lo[1] + (int)((local / ex) % ey);
    const int k = bx.lo[2] + (int)(local / ((long)ex * ey));
    const BcFace &fc = ba.face[bx.face];
    const Grid &g = ba.g;
    const int idx = Index(g, i, j, k);
    const int r = bx.layer;
    const int *N = fc.N;
    double out[5];
    if (r != 0) {
        if (fc.type == 2 || fc.type == 3) {            // slip / no-slip wall: mirror about the wall
            double UO[5], UI[5], UoO[6], UoI[6], UoG[6];
            LoadNode(ba.U, g.N, Index(g, i - r * N[0], j - r * N[1], k - r * N[2]), UO);
            Primitive(ba.gamma, ba.gasR, UO, UoO);
            LoadNode(ba.U, g.N, Index(g, i - 2 * r * N[0], j - 2 * r * N[1], k - 2 * r * N[2]), UI);
            Primitive(ba.gamma, ba.gasR, UI, UoI);
            UoG[1] = UoO[1] + UoO[1] - UoI[1];
            UoG[2] = UoO[2] + UoO[2] - UoI[2];
            UoG[3] = UoO[3] + UoO[3] - UoI[3];
            UoG[4] = UoI[4];
            UoG[5] = UoI[5];
            UoG[0] = UoG[4] / (UoG[5] * ba.gasR);
            Conservative(ba.gamma, UoG, out);
        } else if (fc.type == 4) {                     // periodic: copy from one period inward
            LoadNode(ba.U, g.N, Index(g, i - ba.m[0] * N[0], j - ba.m[1] * N[1], k - ba.m[2] * N[2]), out);
        } else {                                       // copy the inward neighbour
            LoadNode(ba.U, g.N, Index(g, i - N[0], j - N[1], k - N[2]), out);
        }
        StoreNode(ba.U, g.N, idx, out);
        return;
    }
    switch (fc.type) {
        case 0:                                        // inflow: prescribed state
            Conservative(ba.gamma, fc.given, out);
            break;
        case 1:                                        // outflow: zero gradient
            LoadNode(ba.U, g.N, Index(g, i - N[0], j - N[1], k - N[2]), out);
            break;
        case 2:
        case 3: {                                      // walls
            double Uh[5], Uoh[6], UoO[6];
            LoadNode(ba.U, g.N, Index(g, i - N[0], j - N[1], k - N[2]), Uh);
            Primitive(ba.gamma, ba.gasR, Uh, Uoh);
            if (fc.type == 2) {
                UoO[1] = (!N[0]) * Uoh[1];
                UoO[2] = (!N[1]) * Uoh[2];
                UoO[3] = (!N[2]) * Uoh[3];
            } else {
                UoO[1] = 0.0;
                UoO[2] = 0.0;
                UoO[3] = 0.0;
            }
            UoO[4] = Uoh[4];
            UoO[5] = (0.0 > fc.given[5]) ? Uoh[5] : fc.given[5];
            UoO[0] = UoO[4] / (UoO[5] * ba.gasR);
            Conservative(ba.gamma, UoO, out);
            break;
        }
        default:
            return;
    }
    StoreNode(ba.U, g.N, idx, out);
}

// Resolved boundary fill: each entry copies one interior node or writes the
// inflow state of a face (boundary_treatment.c:66-185 replayed on the host).
struct GatherArgs {
    double *U;
    long N;
    double gamma;
    double given[7][6];
};

__global__ __launch_bounds__(256) void bc_gather(GatherArgs ga, const int *__restrict__ dst,
                                                 const int *__restrict__ src, long count)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int d = dst[t], q = src[t];
    double out[5];
    if (q >= 0) {
        LoadNode(ga.U, ga.N, q, out);
    } else {
        Conservative(ga.gamma, ga.given[-3 - q], out);
    }
    StoreNode(ga.U, ga.N, d, out);
}

// ------------------------------------------------------- immersed boundary
struct GhostDev {
    int node, kind, first, count;
    int noslip, isothermal;
    double wallT, wO;
    double Vs[3], N[3], Ta[3], Tb[3];
};

__global__ __launch_bounds__(128) void ib_layer(double *U, long N, const GhostDev *ghosts, int begin, int end,
                                                const int *snode, const double *sw, double gamma, double gasR)
{
    const int gidx = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (gidx >= end) return;
    const GhostDev gh = ghosts[gidx];
    double Uo[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double wsum = 0.0;
    for (int q = gh.first; q < gh.first + gh.count; ++q) {
        double Uh[5], Uoh[6];
        LoadNode(U, N, snode[q], Uh);
        Primitive(gamma, gasR, Uh, Uoh);
        const double w = sw[q];
#pragma unroll
        for (int c = 0; c < 6; ++c) Uo[c] = Uo[c] + Uoh[c] * w;
        wsum = wsum + w;
    }
    double UoG[6];
    if (gh.kind == 0) {
        // boundary point from the pre-estimate, then the boundary point joins the stencil
        const double weight = 1.0 / wsum;
        double UoO[6];
        UoO[0] = 0.0;
        if (gh.noslip) {
            UoO[1] = gh.Vs[0];
            UoO[2] = gh.Vs[1];
            UoO[3] = gh.Vs[2];
        } else {
            const double V[3] = {Uo[1] * weight, Uo[2] * weight, Uo[3] * weight};
            const double rx = gh.Vs[0] * gh.N[0] + gh.Vs[1] * gh.N[1] + gh.Vs[2] * gh.N[2];
            const double ry = V[0] * gh.Ta[0] + V[1] * gh.Ta[1] + V[2] * gh.Ta[2];
            const double rz = V[0] * gh.Tb[0] + V[1] * gh.Tb[1] + V[2] * gh.Tb[2];
#pragma unroll
            for (int s = 0; s < 3; ++s) UoO[1 + s] = gh.N[s] * rx + gh.Ta[s] * ry + gh.Tb[s] * rz;
        }
        UoO[4] = Uo[4] * weight;
        UoO[5] = gh.isothermal ? gh.wallT : Uo[5] * weight;
#pragma unroll
        for (int c = 0; c < 6; ++c) Uo[c] = Uo[c] + UoO[c] * gh.wO;
        wsum = wsum + gh.wO;
#pragma unroll
        for (int c = 0; c < 6; ++c) Uo[c] = Uo[c] / wsum;
        // method of images across the wall
        UoG[1] = UoO[1] + UoO[1] - Uo[1];
        UoG[2] = UoO[2] + UoO[2] - Uo[2];
        UoG[3] = UoO[3] + UoO[3] - Uo[3];
        UoG[4] = Uo[4];
        UoG[5] = Uo[5];
    } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) UoG[c] = Uo[c] / wsum;
    }
    UoG[0] = UoG[4] / (UoG[5] * gasR);
    double out[5];
    Conservative(gamma, UoG, out);
    StoreNode(U, N, gh.node, out);
}

// ------------------------------------------------------------- time step
__global__ __launch_bounds__(256) void cfl_max(const double *U, const uint8_t *fluid, Grid g, double gamma,
                                               double gasR, unsigned long long *vmax)
{
    const int ex = g.hi[0] - g.lo[0], ey = g.hi[1] - g.lo[1];
    const long total = (long)ex * ey * (g.hi[2] - g.lo[2]);
    double m[3] = {0.0, 0.0, 0.0};
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int i = g.lo[0] + (int)(t % ex);
        const int j = g.lo[1] + (int)((t / ex) % ey);
        const int k = g.lo[2] + (int)(t / ((long)ex * ey));
        const int idx = Index(g, i, j, k);
        if (!fluid[idx]) continue;
        double Un[5], Uo[6];
        LoadNode(U, g.N, idx, Un);
        Primitive(gamma, gasR, Un, Uo);
        const double c = sqrt(gamma * gasR * Uo[5]);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const double V = fabs(Uo[s + 1]) + c;
            if (m[s] < V) m[s] = V;
        }
    }
    __shared__ double red[3][256];
#pragma unroll
    for (int s = 0; s < 3; ++s) red[s][threadIdx.x] = m[s];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const double o = red[s][threadIdx.x + h];
                if (red[s][threadIdx.x] < o) red[s][threadIdx.x] = o;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // speeds are >= 0, so their bit patterns order like the values
#pragma unroll
        for (int s = 0; s < 3; ++s) atomicMax(vmax + s, (unsigned long long)__double_as_longlong(red[s][0]));
    }
}

// Plain streaming copy: the reference point for HBM counters and bandwidth.
__global__ __launch_bounds__(256) void stream_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) dst[t] = src[t];
}

__global__ void fill_u64(unsigned long long *p, int n, unsigned long long v)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

}  // namespace

// ------------------------------------------------------------- device state
struct ArtraDevice {
    ArtraDevParams par;
    Grid g;
    double *U = nullptr;            // [3][5][N]
    uint8_t *fluid = nullptr;
    unsigned long long *vmax = nullptr;
    hipStream_t stream = nullptr;
    std::vector<BcGroup> groups[4];          // mode 0: full fill; 1 + s: faces of axis s only
    BcArgs bc{};
    int gatherResolved = 0;
    int useGather = 1;
    int *gDst[4] = {nullptr, nullptr, nullptr, nullptr};
    int *gSrc[4] = {nullptr, nullptr, nullptr, nullptr};
    long gCount[4] = {0, 0, 0, 0};
    GhostDev *ghosts = nullptr;
    int *snode = nullptr;
    double *sw = nullptr;
    std::vector<int> layerStart;
    FluxModel mdl{};
    int segment = 32;
    int exact = 0;                  // 1: reference operation order (bit-faithful), 0: fast arithmetic
    int partialFill = 1;            // axis-only fills between the stages of inviscid split sweeps
    // timing
    bool timing = false;
    double ms[ARTRA_T_COUNT] = {0.0, 0.0, 0.0, 0.0};
    long launches[ARTRA_T_COUNT] = {0, 0, 0, 0};
    std::vector<hipEvent_t> evPool;
    std::vector<std::pair<int, int>> pending;   // (category, first event index)
    size_t evUsed = 0;
};

namespace {

double *Level(ArtraDevice *dev, int level)
{
    return dev->U + (size_t)level * 5 * dev->g.N;
}

int Begin(ArtraDevice *dev, int cat, int *slot)
{
    *slot = -1;
    if (!dev->timing) return 0;
    if (dev->evUsed + 2 > dev->evPool.size()) {
        const size_t grow = dev->evPool.size() ? dev->evPool.size() : 256;
        for (size_t q = 0; q < grow; ++q) {
            hipEvent_t ev;
            HIP_TRY(hipEventCreate(&ev));
            dev->evPool.push_back(ev);
        }
    }
    *slot = (int)dev->evUsed;
    dev->evUsed += 2;
    HIP_TRY(hipEventRecord(dev->evPool[*slot], dev->stream));
    dev->pending.emplace_back(cat, *slot);
    return 0;
}

int End(ArtraDevice *dev, int slot)
{
    if (slot < 0) return 0;
    HIP_TRY(hipEventRecord(dev->evPool[slot + 1], dev->stream));
    return 0;
}

int Harvest(ArtraDevice *dev)
{
    if (dev->pending.empty()) return 0;
    HIP_TRY(hipStreamSynchronize(dev->stream));
    for (const auto &p : dev->pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, dev->evPool[p.second], dev->evPool[p.second + 1]));
        dev->ms[p.first] += ms;
        dev->launches[p.first] += 1;
    }
    dev->pending.clear();
    dev->evUsed = 0;
    return 0;
}

template <int NS, bool VISC, bool FAST>
void LaunchSweep(ArtraDevice *dev, int s, const StageArgs &sa)
{
    const Grid &g = dev->g;
    if (s == 0) {
        const long faces = (long)(g.hi[0] - g.lo[0] + 1) * (g.hi[1] - g.lo[1]) * (g.hi[2] - g.lo[2]);
        const long blocks = (faces - 1 + (kXBlock - 2)) / (kXBlock - 1);
        if (blocks <= 0) return;
        hipLaunchKernelGGL((sweep_x<NS, VISC, FAST>), dim3((unsigned)blocks), dim3(kXBlock), 0, dev->stream, g, sa,
                           dev->mdl);
        return;
    }
    const int T2 = (s == 1) ? 2 : 1;
    const int ex = g.hi[0] - g.lo[0], et = g.hi[T2] - g.lo[T2], es = g.hi[s] - g.lo[s];
    const int seg = dev->segment;
    if (ex <= 0 || et <= 0 || es <= 0) return;
    dim3 grid((ex + kMarchX - 1) / kMarchX, (et + kMarchT - 1) / kMarchT, (es + seg - 1) / seg);
    dim3 block(kMarchX, kMarchT);
    if (s == 1) {
        hipLaunchKernelGGL((sweep_march<1, NS, VISC, FAST>), grid, block, 0, dev->stream, g, sa, dev->mdl, seg);
    } else {
        hipLaunchKernelGGL((sweep_march<2, NS, VISC, FAST>), grid, block, 0, dev->stream, g, sa, dev->mdl, seg);
    }
}

int Sweep(ArtraDevice *dev, double dt, double a, double b, int to, int tn, int tm, int p)
{
    const int s0 = (p == 3) ? 0 : p, s1 = (p == 3) ? 3 : p + 1;
    for (int s = s0; s < s1; ++s) {
        StageArgs sa;
        sa.Utn = Level(dev, tn);
        sa.Uto = Level(dev, to);
        sa.Utm = Level(dev, tm);
        sa.fluid = dev->fluid;
        sa.a = a;
        sa.b = b;
        sa.r = dt * dev->par.dd[s];
        sa.accumulate = (dev->par.multidim == 1 && s != 0);
        sa.toIsTn = (to == tn);
        for (int q = 0; q < 3; ++q) sa.dd[q] = dev->par.dd[q];
        int slot;
        if (Begin(dev, ARTRA_T_SWEEP, &slot)) return 1;
        const bool visc = dev->mdl.refMu > 0.0;
        const bool fast = !dev->exact;
        if (dev->par.sScheme == 1) {
            if (visc) {
                if (fast) LaunchSweep<6, true, true>(dev, s, sa); else LaunchSweep<6, true, false>(dev, s, sa);
            } else {
                if (fast) LaunchSweep<6, false, true>(dev, s, sa); else LaunchSweep<6, false, false>(dev, s, sa);
            }
        } else {
            if (visc) {
                if (fast) LaunchSweep<4, true, true>(dev, s, sa); else LaunchSweep<4, true, false>(dev, s, sa);
            } else {
                if (fast) LaunchSweep<4, false, true>(dev, s, sa); else LaunchSweep<4, false, false>(dev, s, sa);
            }
        }
        HIP_TRY(hipGetLastError());
        if (End(dev, slot)) return 1;
    }
    return 0;
}

// mode 0: the whole external boundary; 1 + s: only the ghosts a sweep along
// axis s reads (valid between the stages of an inviscid split sweep).
int TreatBoundary(ArtraDevice *dev, int level, int mode = 0)
{
    double *U = Level(dev, level);
    // immersed boundary first, layer by layer
    const int layers = (int)dev->layerStart.size() - 1;
    for (int r = 0; r < layers; ++r) {
        const int b = dev->layerStart[r], e = dev->layerStart[r + 1];
        if (e <= b) continue;
        int slot;
        if (Begin(dev, ARTRA_T_IB, &slot)) return 1;
        hipLaunchKernelGGL(ib_layer, dim3((e - b + 127) / 128), dim3(128), 0, dev->stream, U, dev->g.N, dev->ghosts, b,
                           e, dev->snode, dev->sw, dev->par.gamma, dev->par.gasR);
        HIP_TRY(hipGetLastError());
        if (End(dev, slot)) return 1;
    }
    if (dev->gatherResolved && dev->useGather) {
        const long count = dev->gCount[mode];
        if (count <= 0) return 0;
        GatherArgs ga;
        ga.U = U;
        ga.N = dev->g.N;
        ga.gamma = dev->par.gamma;
        for (int p = 0; p < 7; ++p)
            for (int q = 0; q < 6; ++q) ga.given[p][q] = dev->par.varBC[p][q];
        int slot;
        if (Begin(dev, ARTRA_T_BC, &slot)) return 1;
        hipLaunchKernelGGL(bc_gather, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, dev->stream, ga,
                           dev->gDst[mode], dev->gSrc[mode], count);
        HIP_TRY(hipGetLastError());
        return End(dev, slot);
    }
    BcArgs ba = dev->bc;
    ba.U = U;
    for (const BcGroup &grp : dev->groups[mode]) {
        if (grp.total <= 0) continue;
        int slot;
        if (Begin(dev, ARTRA_T_BC, &slot)) return 1;
        hipLaunchKernelGGL(bc_group, dim3((unsigned)((grp.total + 255) / 256)), dim3(256), 0, dev->stream, ba, grp);
        HIP_TRY(hipGetLastError());
        if (End(dev, slot)) return 1;
    }
    return 0;
}

// Runge-Kutta stages of one directional operator (fluid_dynamics.c:121-148).
// Between the stages of an inviscid sweep along one axis only that axis's
// ghosts are read, so only they are filled; the last stage fills everything,
// exactly as TreatBoundary does after every stage.
int Stages(ArtraDevice *dev, double dt, int p)
{
    const int mid = (p < 3 && !(dev->mdl.refMu > 0.0) && dev->partialFill) ? 1 + p : 0;
    if (dev->par.tScheme == 0) {
        if (Sweep(dev, dt, 0.0, 1.0, 0, 0, 1, p) || TreatBoundary(dev, 1, mid)) return 1;
        if (Sweep(dev, dt, 1.0 / 2.0, 1.0 / 2.0, 0, 1, 0, p) || TreatBoundary(dev, 0)) return 1;
        return 0;
    }
    if (Sweep(dev, dt, 0.0, 1.0, 0, 0, 1, p) || TreatBoundary(dev, 1, mid)) return 1;
    if (Sweep(dev, dt, 3.0 / 4.0, 1.0 / 4.0, 0, 1, 2, p) || TreatBoundary(dev, 2, mid)) return 1;
    if (Sweep(dev, dt, 1.0 / 3.0, 2.0 / 3.0, 0, 2, 0, p) || TreatBoundary(dev, 0)) return 1;
    return 0;
}

}  // namespace

extern "C" {

const char *artra_dev_error(void)
{
    return gError;
}

int artra_dev_create(const ArtraDevParams *par, const uint8_t *fluid, ArtraDevice **out)
{
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return Fail("no HIP device is visible");
    ArtraDevice *dev = new ArtraDevice();
    dev->par = *par;
    Grid &g = dev->g;
    g.nx = par->n[0];
    g.ny = par->n[1];
    g.nz = par->n[2];
    g.N = (long)g.nx * g.ny * g.nz;
    for (int s = 0; s < 3; ++s) {
        g.lo[s] = par->pin[s][0];
        g.hi[s] = par->pin[s][1];
    }
    dev->mdl.gamma = par->gamma;
    dev->mdl.cv = par->cv;
    dev->mdl.refMu = par->refMu;
    dev->mdl.refT = par->refT;
    dev->mdl.roe = par->jacobMean;
    dev->mdl.sw = par->fluxSplit;
    const size_t fieldBytes = (size_t)3 * 5 * g.N * sizeof(double);
    hipError_t e = hipStreamCreateWithFlags(&dev->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&dev->U, fieldBytes);
    if (e == hipSuccess) e = hipMalloc(&dev->fluid, (size_t)g.N);
    if (e == hipSuccess) e = hipMalloc(&dev->vmax, 4 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemcpy(dev->fluid, fluid, (size_t)g.N, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        artra_dev_destroy(dev);
        return Fail("device allocation of %.1f MB failed: %s", fieldBytes / 1e6, hipGetErrorString(e));
    }
    BcArgs &ba = dev->bc;
    ba.g = g;
    for (int s = 0; s < 3; ++s) ba.m[s] = par->m[s];
    ba.gamma = par->gamma;
    ba.gasR = par->gasR;
    for (int p = 0; p < 7; ++p) {
        ba.face[p].type = par->typeBC[p];
        for (int s = 0; s < 3; ++s) ba.face[p].N[s] = par->normal[p][s];
        for (int q = 0; q < 6; ++q) ba.face[p].given[q] = par->varBC[p][q];
    }
    dev->layerStart.assign(1, 0);
    *out = dev;
    return 0;
}

void artra_dev_destroy(ArtraDevice *dev)
{
    if (!dev) return;
    if (dev->stream) hipStreamSynchronize(dev->stream);
    for (hipEvent_t ev : dev->evPool) hipEventDestroy(ev);
    hipFree(dev->U);
    hipFree(dev->fluid);
    hipFree(dev->vmax);
    hipFree(dev->ghosts);
    hipFree(dev->snode);
    hipFree(dev->sw);
    for (int m = 0; m < 4; ++m) {
        hipFree(dev->gDst[m]);
        hipFree(dev->gSrc[m]);
    }
    if (dev->stream) hipStreamDestroy(dev->stream);
    delete dev;
}

int artra_dev_set_boundary(ArtraDevice *dev, int mode, const ArtraBcBox *boxes, int nbox, const int *group_start,
                           int ngroup)
{
    if (mode < 0 || mode > 3) return Fail("bad boundary mode %d", mode);
    std::vector<BcGroup> &groups = dev->groups[mode];
    groups.clear();
    for (int q = 0; q < ngroup; ++q) {
        BcGroup grp{};
        long offset = 0;
        for (int b = group_start[q]; b < group_start[q + 1]; ++b) {
            if (grp.count == kMaxGroup) {
                grp.total = offset;
                groups.push_back(grp);
                grp = BcGroup{};
                offset = 0;
            }
            BcBox &bx = grp.box[grp.count++];
            long vol = 1;
            for (int s = 0; s < 3; ++s) {
                bx.lo[s] = boxes[b].box[s][0];
                bx.hi[s] = boxes[b].box[s][1];
                vol *= bx.hi[s] - bx.lo[s];
            }
            bx.face = boxes[b].face;
            bx.layer = boxes[b].layer;
            bx.offset = offset;
            offset += vol;
        }
        grp.total = offset;
        if (grp.count) groups.push_back(grp);
    }
    (void)nbox;
    return 0;
}

int artra_dev_set_gather(ArtraDevice *dev, const ArtraBcGather *gt)
{
    for (int m = 0; m < 4; ++m) {
        hipFree(dev->gDst[m]);
        hipFree(dev->gSrc[m]);
        dev->gDst[m] = dev->gSrc[m] = nullptr;
        dev->gCount[m] = 0;
    }
    dev->gatherResolved = 0;
    if (!gt->resolved) return 0;
    for (int m = 0; m < 4; ++m) {
        const long n = (m == 0) ? gt->count : gt->axis_count[m - 1];
        const int *d = (m == 0) ? gt->dst : gt->axis_dst[m - 1];
        const int *q = (m == 0) ? gt->src : gt->axis_src[m - 1];
        dev->gCount[m] = n;
        if (n == 0) continue;
        HIP_TRY(hipMalloc(&dev->gDst[m], n * sizeof(int)));
        HIP_TRY(hipMalloc(&dev->gSrc[m], n * sizeof(int)));
        HIP_TRY(hipMemcpy(dev->gDst[m], d, n * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(dev->gSrc[m], q, n * sizeof(int), hipMemcpyHostToDevice));
    }
    dev->gatherResolved = 1;
    return 0;
}

int artra_dev_set_fill_options(ArtraDevice *dev, int use_gather, int partial)
{
    dev->useGather = (use_gather != 0);
    dev->partialFill = (partial != 0);
    return 0;
}

int artra_dev_set_ghosts(ArtraDevice *dev, const ArtraGhostTable *t)
{
    hipFree(dev->ghosts);
    hipFree(dev->snode);
    hipFree(dev->sw);
    dev->ghosts = nullptr;
    dev->snode = nullptr;
    dev->sw = nullptr;
    dev->layerStart.assign(t->layer_start, t->layer_start + t->layers + 1);
    if (t->nghost == 0) return 0;
    std::vector<GhostDev> host(t->nghost);
    for (int q = 0; q < t->nghost; ++q) {
        const ArtraGhost &s = t->ghosts[q];
        GhostDev &d = host[q];
        d.node = s.node;
        d.kind = s.kind;
        d.first = s.first;
        d.count = s.count;
        d.noslip = s.noslip;
        d.isothermal = s.isothermal;
        d.wallT = s.wallT;
        d.wO = s.wO;
        for (int k = 0; k < 3; ++k) {
            d.Vs[k] = s.Vs[k];
            d.N[k] = s.N[k];
            d.Ta[k] = s.Ta[k];
            d.Tb[k] = s.Tb[k];
        }
    }
    HIP_TRY(hipMalloc(&dev->ghosts, host.size() * sizeof(GhostDev)));
    HIP_TRY(hipMalloc(&dev->snode, (size_t)t->nstencil * sizeof(int)));
    HIP_TRY(hipMalloc(&dev->sw, (size_t)t->nstencil * sizeof(double)));
    HIP_TRY(hipMemcpy(dev->ghosts, host.data(), host.size() * sizeof(GhostDev), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dev->snode, t->stencil_node, (size_t)t->nstencil * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dev->sw, t->stencil_w, (size_t)t->nstencil * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int artra_dev_upload(ArtraDevice *dev, int level, const double *soa)
{
    HIP_TRY(hipMemcpyAsync(Level(dev, level), soa, (size_t)5 * dev->g.N * sizeof(double), hipMemcpyHostToDevice,
                           dev->stream));
    HIP_TRY(hipStreamSynchronize(dev->stream));
    return 0;
}

int artra_dev_download(ArtraDevice *dev, int level, double *soa)
{
    HIP_TRY(hipMemcpyAsync(soa, Level(dev, level), (size_t)5 * dev->g.N * sizeof(double), hipMemcpyDeviceToHost,
                           dev->stream));
    HIP_TRY(hipStreamSynchronize(dev->stream));
    return 0;
}

int artra_dev_poison(ArtraDevice *dev)
{
    HIP_TRY(hipMemsetAsync(dev->U, 1, (size_t)3 * 5 * dev->g.N * sizeof(double), dev->stream));
    HIP_TRY(hipStreamSynchronize(dev->stream));
    return 0;
}

int artra_dev_treat_boundary(ArtraDevice *dev, int level)
{
    if (level < 0 || level > 2) return Fail("bad time level %d", level);
    return TreatBoundary(dev, level);
}

int artra_dev_max_speed(ArtraDevice *dev, const double floor[3], double vmax[3])
{
    unsigned long long init[3];
    for (int s = 0; s < 3; ++s) {
        const double f = (floor[s] > 0.0) ? floor[s] : 0.0;
        memcpy(init + s, &f, sizeof f);
    }
    int slot;
    if (Begin(dev, ARTRA_T_CFL, &slot)) return 1;
    for (int s = 0; s < 3; ++s) {
        hipLaunchKernelGGL(fill_u64, dim3(1), dim3(1), 0, dev->stream, dev->vmax + s, 1, init[s]);
    }
    const Grid &g = dev->g;
    const long total = (long)(g.hi[0] - g.lo[0]) * (g.hi[1] - g.lo[1]) * (g.hi[2] - g.lo[2]);
    long blocks = (total + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(cfl_max, dim3((unsigned)blocks), dim3(256), 0, dev->stream, Level(dev, 0), dev->fluid, g,
                       dev->par.gamma, dev->par.gasR, dev->vmax);
    HIP_TRY(hipGetLastError());
    if (End(dev, slot)) return 1;
    unsigned long long host[3];
    HIP_TRY(hipMemcpyAsync(host, dev->vmax, sizeof host, hipMemcpyDeviceToHost, dev->stream));
    HIP_TRY(hipStreamSynchronize(dev->stream));
    for (int s = 0; s < 3; ++s) memcpy(vmax + s, host + s, sizeof(double));
    return Harvest(dev);
}

int artra_dev_evolve(ArtraDevice *dev, double dt)
{
    const double h = 0.5 * dt;
    if (dev->par.multidim == 1) {
        if (Stages(dev, h, 3) || Stages(dev, h, 3)) return 1;
        return 0;
    }
    static const int order[9][6] = {{2, 1, 0, 0, 1, 2},    {2, 1, 1, 2, -1, -1},   {2, 0, 0, 2, -1, -1},
                                    {1, 0, 0, 1, -1, -1},  {-1, -1, -1, -1, -1, -1}, {2, 2, -1, -1, -1, -1},
                                    {-1, -1, -1, -1, -1, -1}, {1, 1, -1, -1, -1, -1}, {0, 0, -1, -1, -1, -1}};
    const int c = dev->par.collapse;
    if (c < 0 || c > 8 || order[c][0] < 0) return Fail("unsupported collapse flag %d", c);
    for (int q = 0; q < 6 && order[c][q] >= 0; ++q) {
        if (Stages(dev, h, order[c][q])) return 1;
    }
    if (dev->timing && dev->evUsed > 4096) return Harvest(dev);
    return 0;
}

int artra_dev_sync(ArtraDevice *dev)
{
    HIP_TRY(hipStreamSynchronize(dev->stream));
    return Harvest(dev);
}

int artra_dev_set_timing(ArtraDevice *dev, int on)
{
    if (Harvest(dev)) return 1;
    dev->timing = (on != 0);
    for (int q = 0; q < ARTRA_T_COUNT; ++q) {
        dev->ms[q] = 0.0;
        dev->launches[q] = 0;
    }
    return 0;
}

int artra_dev_get_timing(ArtraDevice *dev, double ms[ARTRA_T_COUNT], long launches[ARTRA_T_COUNT])
{
    if (Harvest(dev)) return 1;
    for (int q = 0; q < ARTRA_T_COUNT; ++q) {
        ms[q] = dev->ms[q];
        launches[q] = dev->launches[q];
    }
    return 0;
}

int artra_dev_copy_probe(long bytes, int iters, double *gbps)
{
    const long n = bytes / (long)sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    hipEvent_t e0, e1;
    HIP_TRY(hipMalloc(&a, n * sizeof(double2)));
    HIP_TRY(hipMalloc(&b, n * sizeof(double2)));
    HIP_TRY(hipMemset(a, 0, n * sizeof(double2)));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int cus = 256;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0);
    const dim3 grid(cus * 8), block(256);
    hipLaunchKernelGGL(stream_copy, grid, block, 0, 0, a, b, n);
    HIP_TRY(hipEventRecord(e0, 0));
    for (int q = 0; q < iters; ++q) hipLaunchKernelGGL(stream_copy, grid, block, 0, 0, a, b, n);
    HIP_TRY(hipEventRecord(e1, 0));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    *gbps = 2.0 * n * sizeof(double2) * iters / (ms * 1e-3) / 1e9;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return 0;
}

int artra_dev_set_exact(ArtraDevice *dev, int exact)
{
    dev->exact = (exact != 0);
    return 0;
}

int artra_dev_set_tuning(ArtraDevice *dev, int march_segment, int xblock)
{
    (void)xblock;
    if (march_segment > 0) dev->segment = march_segment;
    return 0;
}

}  // extern "C"
