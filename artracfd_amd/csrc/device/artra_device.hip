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
//   ib_fill        the immersed-boundary fill of one ghost layer (or of all
//                  layers when none reads another) from the host-precomputed
//                  stencil table, a 64-lane block per ghost node
//                  (immersed_boundary.c:270-487).
//   cfl_max        per-axis maximum characteristic speed (solve.c:108-153).
//
// All FP64 arithmetic follows the reference's operation order; the file is
// compiled with -ffp-contract=off so results match the CPU bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "flux.hpp"
#include "indicator.hpp"

extern "C" {
#include "artra_device.h"
}
#include "geomcore.h"

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
constexpr int kListBlocks = 768;    // persistent blocks of rk_line_list (3 per CU)
constexpr int kMarchChunk = ARTRA_DEV_MARCH_CHUNK;   // steps per device-stepped march call (the record's length)

// bytes of a device march's scratch with a record of cap steps (see MarchScratch)
constexpr int kMarchClock = ARTRA_DEV_MARCH_CLOCK;     // doubles of a device march's clock
constexpr size_t MarchScratchBytes(int cap)
{
    return (size_t)(4 + kMarchClock + 3 + cap) * sizeof(double);
}

// Scratch of the allocation-free march (artra_dev_march_noalloc): speed
// maxima, clock, step and a record of up to kNoAllocSteps steps, laid out
// as MarchScratch.  Every GPU the library runs on holds its own copy until
// the process ends; calls that overlap on one GPU share it and race.
constexpr int kNoAllocSteps = ARTRA_DEV_NOALLOC_STEPS;
static __device__ __align__(256) double scratch[4 + kMarchClock + 3 + kNoAllocSteps];
static_assert(sizeof scratch == MarchScratchBytes(kNoAllocSteps), "scratch layout");

constexpr int kMaxDevices = 64;
std::mutex gScratchLock;
void *gScratch[kMaxDevices];        // per GPU: the device address of its copy of scratch

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
    const double *step;             // when set, the operator's step on the device: r = *step * dds
    double dds;
};

// The stage's r from the device-resident step when there is one; false when
// that step is zero (a march that has reached its end time: nothing moves).
__device__ __forceinline__ bool ResolveStep(StageArgs &sa)
{
    if (!sa.step) return true;
    const double h = *sa.step;
    if (h == 0.0) return false;
    sa.r = h * sa.dds;
    return true;
}

__device__ __forceinline__ int Index(const Grid &g, int i, int j, int k)
{
    return (k * g.ny + j) * g.nx + i;
}

// The field is node-major: the five conserved components of a node are
// contiguous (40 bytes), so a node is one or two cache lines whichever
// direction a line runs.
__device__ __forceinline__ double *NodeAt(double *U, long idx)
{
    return U + idx * 5;
}

__device__ __forceinline__ const double *NodeAt(const double *U, long idx)
{
    return U + idx * 5;
}

__device__ __forceinline__ void LoadNode(const double *U, long idx, double out[5])
{
    const double *p = NodeAt(U, idx);
#pragma unroll
    for (int c = 0; c < 5; ++c) out[c] = p[c];
}

// Solution operator of one node (fluid_dynamics.c:243-274).  FAST evaluates
// the same expression with fused multiply-adds.
template <bool VISC, bool FAST>
__device__ __forceinline__ void UpdateNode(const StageArgs &sa, long idx, const double FL[5], const double FR[5],
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
        double *um = NodeAt(sa.Utm, idx);
#pragma unroll
        for (int c = 0; c < 5; ++c) um[c] = FAST ? fma(br, phi[c], um[c]) : um[c] + br * phi[c];
        return;
    }
    double Uo[5];
    if (sa.toIsTn) {
#pragma unroll
        for (int c = 0; c < 5; ++c) Uo[c] = Un[c];
    } else {
        LoadNode(sa.Uto, idx, Uo);
    }
    double *um = NodeAt(sa.Utm, idx);
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        if (FAST) um[c] = fma(sa.a, Uo[c], sa.b * fma(sa.r, phi[c], Un[c]));
        else um[c] = sa.a * Uo[c] + sa.b * (Un[c] + sa.r * phi[c]);
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
    LoadNode(U, idxP - a, nb[0]);
    LoadNode(U, idxP + a, nb[1]);
    LoadNode(U, idxP - b, nb[2]);
    LoadNode(U, idxP + b, nb[3]);
    LoadNode(U, idxP + e - a, nb[4]);
    LoadNode(U, idxP + e + a, nb[5]);
    LoadNode(U, idxP + e - b, nb[6]);
    LoadNode(U, idxP + e + b, nb[7]);
}

// ------------------------------------------------------------------ sweeps
template <int NS, bool VISC, bool FAST>
__global__ __launch_bounds__(kXBlock) void sweep_x(Grid g, StageArgs sa, FluxModel mdl)
{
    if (!ResolveStep(sa)) return;
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
        for (int m = 0; m < NS; ++m) LoadNode(sa.Utn, base + SL + m, st[m]);
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
            UpdateNode<VISC, FAST>(sa, idx, FL, FR, FvL, FvR, Un);
        }
    }
}

template <int S, int NS, bool VISC, bool FAST>
__global__ __launch_bounds__(kMarchX *kMarchT) __attribute__((amdgpu_waves_per_eu(2))) void sweep_march(Grid g, StageArgs sa, FluxModel mdl, int segment)
{
    if (!ResolveStep(sa)) return;
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
    for (int m = 0; m < NS; ++m) LoadNode(sa.Utn, idx0 + (SL - 1 + m) * stride, st[m]);
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
        LoadNode(sa.Utn, idx + (SL + NS - 1) * stride, st[NS - 1]);
        double FR[5], FvR[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (VISC) {
            // the viscous flux first: its neighbour values are dead before
            // the characteristic projection needs its registers
            double nb[8][5];
            LoadViscousNeighbours<S>(g, sa.Utn, idx, nb);
            ViscFlux<S, FAST>(mdl, sa.dd, st[K0], st[K0 + 1], nb, FvR);
        }
        ConvFlux<S, NS, FAST>(mdl, st, FR);
        if (sa.fluid[idx]) UpdateNode<VISC, FAST>(sa, idx, FL, FR, FvL, FvR, st[K0]);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            FL[c] = FR[c];
            FvL[c] = FvR[c];
        }
    }
}

// One Runge-Kutta stage of the gravity source operator (LLLU's PHI branch,
// fluid_dynamics.c:177-214, with Phi from source_term.c:21-39): every
// interior fluid node takes U_tm = a U_to + b (U_tn + dt Phi(U_tn)),
// Phi = (0, rho g, rho g . u); sa.r is the stage's dt itself.  The update
// never accumulates (the reference passes OPTSPLIT for this operator).
template <bool FAST>
__global__ __launch_bounds__(256) void source_stage(Grid g, StageArgs sa, double gx, double gy, double gz)
{
    if (!ResolveStep(sa)) return;
    const int ex = g.hi[0] - g.lo[0], ey = g.hi[1] - g.lo[1];
    const long total = (long)ex * ey * (g.hi[2] - g.lo[2]);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int i = g.lo[0] + (int)(t % ex);
    const int j = g.lo[1] + (int)((t / ex) % ey);
    const int k = g.lo[2] + (int)(t / ((long)ex * ey));
    const int idx = Index(g, i, j, k);
    if (!sa.fluid[idx]) return;
    double Un[5], Uo[5], phi[5];
    LoadNode(sa.Utn, idx, Un);
    if (sa.toIsTn) {
#pragma unroll
        for (int c = 0; c < 5; ++c) Uo[c] = Un[c];
    } else {
        LoadNode(sa.Uto, idx, Uo);
    }
    const double V[3] = {Un[1] / Un[0], Un[2] / Un[0], Un[3] / Un[0]};
    const double fb[3] = {Un[0] * gx, Un[0] * gy, Un[0] * gz};
    phi[0] = 0.0;
    phi[1] = fb[0];
    phi[2] = fb[1];
    phi[3] = fb[2];
    phi[4] = FAST ? fma(fb[0], V[0], fma(fb[1], V[1], fb[2] * V[2])) : fb[0] * V[0] + fb[1] * V[1] + fb[2] * V[2];
    double *um = NodeAt(sa.Utm, idx);
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        if (FAST) um[c] = fma(sa.a, Uo[c], sa.b * fma(sa.r, phi[c], Un[c]));
        else um[c] = sa.a * Uo[c] + sa.b * (Un[c] + sa.r * phi[c]);
    }
}

// Features and labels of every face along S of level TO (training data).
// Face q of a line is the face between nodes lo-1+q and lo+q along S.
template <int S>
__global__ __launch_bounds__(256) void sample_faces(Grid g, const double *U, FluxModel mdl, float *feat, float *label)
{
    constexpr int A = (S == 0) ? 1 : 0;
    constexpr int B = (S == 2) ? 1 : 2;
    const int faces = g.hi[S] - g.lo[S] + 1;
    const int na = g.hi[A] - g.lo[A];
    const long total = (long)faces * na * (g.hi[B] - g.lo[B]);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int q = (int)(t % faces);
    const long line = t / faces;
    int c[3];
    c[S] = g.lo[S] - 1 + q;
    c[A] = g.lo[A] + (int)(line % na);
    c[B] = g.lo[B] + (int)(line / na);
    const int stride = (S == 0) ? 1 : (S == 1) ? g.nx : g.nx * g.ny;
    const int base = Index(g, c[0], c[1], c[2]);
    double st[6][5];
#pragma unroll
    for (int m = 0; m < 6; ++m) LoadNode(U, base + (m - 2) * stride, st[m]);
    float x[artra::kFeat];
    artra::face_features<6>(mdl.gamma, st, x);
#pragma unroll
    for (int k = 0; k < artra::kFeat; ++k) feat[t * artra::kFeat + k] = x[k];
    artra::FaceState fs;
    artra::face_state<S>(mdl, st[2], st[3], fs);
    double F[5], lab = 0.0;
    artra::weno_flux<S>(fs, st, F, &lab);
    label[t] = (float)lab;
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

__device__ __forceinline__ void StoreNode(double *U, long idx, const double v[5])
{
    double *p = NodeAt(U, idx);
#pragma unroll
    for (int c = 0; c < 5; ++c) p[c] = v[c];
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
            LoadNode(ba.U, Index(g, i - r * N[0], j - r * N[1], k - r * N[2]), UO);
            Primitive(ba.gamma, ba.gasR, UO, UoO);
            LoadNode(ba.U, Index(g, i - 2 * r * N[0], j - 2 * r * N[1], k - 2 * r * N[2]), UI);
            Primitive(ba.gamma, ba.gasR, UI, UoI);
            UoG[1] = UoO[1] + UoO[1] - UoI[1];
            UoG[2] = UoO[2] + UoO[2] - UoI[2];
            UoG[3] = UoO[3] + UoO[3] - UoI[3];
            UoG[4] = UoI[4];
            UoG[5] = UoI[5];
            UoG[0] = UoG[4] / (UoG[5] * ba.gasR);
            Conservative(ba.gamma, UoG, out);
        } else if (fc.type == 4) {                     // periodic: copy from one period inward
            LoadNode(ba.U, Index(g, i - ba.m[0] * N[0], j - ba.m[1] * N[1], k - ba.m[2] * N[2]), out);
        } else {                                       // copy the inward neighbour
            LoadNode(ba.U, Index(g, i - N[0], j - N[1], k - N[2]), out);
        }
        StoreNode(ba.U, idx, out);
        return;
    }
    switch (fc.type) {
        case 0:                                        // inflow: prescribed state
            Conservative(ba.gamma, fc.given, out);
            break;
        case 1:                                        // outflow: zero gradient
            LoadNode(ba.U, Index(g, i - N[0], j - N[1], k - N[2]), out);
            break;
        case 2:
        case 3: {                                      // walls
            double Uh[5], Uoh[6], UoO[6];
            LoadNode(ba.U, Index(g, i - N[0], j - N[1], k - N[2]), Uh);
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
    StoreNode(ba.U, idx, out);
}

// Resolved boundary fill: each entry copies one interior node or writes the
// inflow state of a face (boundary_treatment.c:66-185 replayed on the host).
struct GatherArgs {
    double *U;
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
        LoadNode(ga.U, q, out);
    } else {
        Conservative(ga.gamma, ga.given[-3 - q], out);
    }
    StoreNode(ga.U, d, out);
}

// ------------------------------------------------- fused directional sweep
// All Runge-Kutta stages of one split directional operator for one grid line
// per block: the line (ghosts included) is read from HBM once, the stages run
// on it in LDS - face fluxes into an LDS face buffer, node updates in place,
// the ghost ends refilled between stages by per-line copy rules - and the
// interior is written back once.  Valid where a stage along S reads nothing
// but its own line: inviscid, split, no immersed body, and a between-stage
// fill that only copies along the line (derived from the resolved gather).
// With HYB the learned indicator routes each face: smooth faces take the
// linear fifth-order upwind flux, troubled ones the full WENO5 flux.
constexpr int kLineRules = 12;
constexpr int kRefillPerNode = 4;   // ghosts a line's node refills between stages, at most

// The ghosts a node feeds, kRefillPerNode line positions packed two to a
// register (16 bits each, 0xffff for none; line positions are below 2^16).
struct Feeds {
    unsigned w[kRefillPerNode / 2];
};

// (every slot is written through a compile-time index, so the set stays in
// registers)
__device__ __forceinline__ Feeds FeedsOf(const int *rules, int nrules, int node)
{
    static_assert(kRefillPerNode == 4, "two packed registers");
    unsigned w0 = 0xffffffffu, w1 = 0xffffffffu;
    int nf = 0;
    for (int e = 0; e < nrules; ++e) {
        if (rules[kLineRules + e] != node || nf >= kRefillPerNode) continue;
        const unsigned pos = (unsigned)rules[e];
        const unsigned sh = 16u * (nf & 1);
        const unsigned mask = ~(0xffffu << sh);
        if (nf < 2) w0 = (w0 & mask) | (pos << sh);
        else w1 = (w1 & mask) | (pos << sh);
        ++nf;
    }
    Feeds f;
    f.w[0] = w0;
    f.w[1] = w1;
    return f;
}

__device__ __forceinline__ int FeedAt(const Feeds &f, int e)
{
    const unsigned v = (f.w[e >> 1] >> (16u * (e & 1))) & 0xffffu;
    return (v == 0xffffu) ? -1 : (int)v;
}
// The face tallies of the indicator are spread over slots on separate
// 128-byte lines (block b adds into slot b mod kCountSlots): one counter
// pair taking an atomic from every block serialises tens of thousands of
// atomics per launch on one L2 line.
constexpr int kCountSlots = 256;
constexpr int kCountStride = 16;                 // unsigned long long per slot
constexpr int kTallies = ARTRA_HYBRID_TALLIES;   // faces linear, classified, WENO5, flagged troubled
constexpr size_t kCountBytes = (size_t)kCountSlots * kCountStride * sizeof(unsigned long long);

struct LineArgs {
    double *U;                      // level TO: read at the start, written at the end
    int n;                          // nodes along the line, ghosts included
    int lo, hi;                     // interior range along the line
    int n1;                         // interior extent of the first transverse axis
    int lo1, lo2;                   // interior starts of the transverse axes
    int stride, stride1, stride2;   // node-index strides along S and the transverse axes
    int stages;
    double a[3], b[3];
    double r;                       // dt / ds
    const double *dt;               // when set, the step's dt on the device (r = *dt * dd)
    double dd;                      // 1 / ds
    double gamma;
    int nrules;
    int rdst[kLineRules], rsrc[kLineRules];   // src >= 0: line position, src < 0: constant state -1 - src
    double given[2][6];             // primitive states of constant (inflow) rules
};

struct LineHybrid {
    const uint16_t *W1, *W2;        // bf16 weights
    const float *bias;              // b1[32], b2[32], w3[32]
    float b3, threshold;
    unsigned long long *counts;     // kCountSlots slots of kCountStride: [0] faces on the linear flux, [1] faces
                                    // classified, [2] faces of overflowed lines
    float smoothVariation;          // faces varying less than this (relative) skip the network; 0: none
};

// Blocks are dealt to the XCDs round robin; this bijection gives each XCD a
// contiguous run of lines, so lines that share cache lines (adjacent x for
// y and z lines) meet in the same L2.
__device__ __forceinline__ long XcdLine(long b, long nb)
{
    const long q = nb / 8, rem = nb % 8;
    const long x = b % 8, i = b / 8;
    return (x < rem) ? x * (q + 1) + i : rem * (q + 1) + (x - rem) * q + i;
}

// A value every lane of the wavefront holds alike, moved to scalar
// registers (it then holds no vector registers through the stage loop).
__device__ __forceinline__ double Uniform(double x)
{
    const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)bits);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(bits >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// LDS of one line: the states [n][5] and a work area holding the face
// fluxes [F][5], the operator's initial state of the interior nodes
// [F - 1][5] and the fill rules with their constant states (kRuleBytes);
// with the indicator (hyb) also the FP32 density and pressure of every node
// [2][n] (the network's features).
__host__ __device__ __forceinline__ size_t LineStateBytes(int n)
{
    return ((size_t)5 * n * sizeof(double) + 15) / 16 * 16;
}

// The fill rules in LDS: destinations and sources [2][kLineRules] (ints),
// then the conservative state of every constant-state rule [kLineRules][5].
constexpr size_t kRuleBytes = (size_t)2 * kLineRules * sizeof(int) + (size_t)kLineRules * 5 * sizeof(double);

// The rules and the constant states of one line's fill into LDS (threads tt
// of T); the caller's barrier publishes them.
__device__ __forceinline__ void LoadRules(const LineArgs &la, int *rules, int tt, int T)
{
    double *cst = (double *)(rules + 2 * kLineRules);
    for (int e = tt; e < la.nrules; e += T) {
        rules[e] = la.rdst[e];
        rules[kLineRules + e] = la.rsrc[e];
        if (la.rsrc[e] < 0) {
            double v[5];
            Conservative(la.gamma, la.given[-1 - la.rsrc[e]], v);
#pragma unroll
            for (int c = 0; c < 5; ++c) cst[e * 5 + c] = v[c];
        }
    }
}

__host__ __device__ __forceinline__ size_t LineWorkPlainBytes(int F)
{
    return ((size_t)(5 * F + 5 * (F - 1)) * sizeof(double) + kRuleBytes + 15) / 16 * 16;
}

__host__ __device__ __forceinline__ size_t LineWorkBytes(int n, int F, bool hyb)
{
    return LineWorkPlainBytes(F) + (hyb ? ((size_t)2 * n * sizeof(float) + 15) / 16 * 16 : 0);
}

__host__ __device__ __forceinline__ size_t LineLdsBytes(int n, int F, bool hyb)
{
    return LineStateBytes(n) + LineWorkBytes(n, F, hyb);
}

// The line's node index of interior position (line % n1, line / n1) across S.
__device__ __forceinline__ long LineBase(const LineArgs &la, long line)
{
    const int a1 = (int)(line % la.n1), a2 = (int)(line / la.n1);
    return (long)(la.lo1 + a1) * la.stride1 + (long)(la.lo2 + a2) * la.stride2;
}

// The linear fifth-order face flux of the learned hybrid at the face between
// nodes f and f + 1 of a line held node-major in LDS: the six-node linear
// combinations, the face state (arithmetic mean of the two nodes' velocities
// and total enthalpies, or the model's average) and the closed-form local
// Lax-Friedrichs flux (the characteristic form for Steger-Warming).  The
// one definition every hybrid kernel uses, so that a face's flux does not
// depend on the kernel that computed it.  MEAN compiles only the default
// model (arithmetic mean, local Lax-Friedrichs), whose registers fit more
// wavefronts; otherwise the model is read at run time.
template <int S, bool MEAN = false>
__device__ __forceinline__ void LinearFaceFlux(const FluxModel &mdl, const double *st, int f, double Fx[5])
{
    double Sg[5], Dl[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        // one component's six stencil values at a time (fewer live registers)
        double w[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) w[j] = st[(f - 2 + j) * 5 + c];
        artra::linear_sum_diff_1(w, Sg[c], Dl[c]);
    }
    double UL[5], UR[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        UL[c] = st[f * 5 + c];
        UR[c] = st[(f + 1) * 5 + c];
    }
    artra::FaceState fs;
    if (MEAN || (!mdl.roe && !mdl.sw)) {
        double nL[4], nR[4];
        artra::node_velocity_enthalpy(mdl.gamma, UL, nL);
        artra::node_velocity_enthalpy(mdl.gamma, UR, nR);
        artra::face_state_mean<S>(mdl.gamma, nL, nR, fs);
        artra::linear_flux_llf_sd<S>(mdl.gamma, fs, Sg, Dl, Fx);
    } else {
        artra::face_state<S>(mdl, UL, UR, fs);
        if (mdl.sw) {
            double Up[5], Um[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                Up[c] = 0.5 * (Sg[c] + Dl[c]);
                Um[c] = 0.5 * (Sg[c] - Dl[c]);
            }
            artra::linear_flux_from<S>(fs, Up, Um, Fx);
        } else {
            artra::linear_flux_llf_sd<S>(mdl.gamma, fs, Sg, Dl, Fx);
        }
    }
}

// Whether the face q of a line (between nodes lo - 1 + q and lo + q) is
// troubled: its density and pressure stencils vary by more than the
// smoothness gate allows and the network's logit exceeds the threshold.
// Every lane of the wavefront must call it (the network runs on whole
// wavefronts); valid = false for lanes past the last face.  A wavefront
// whose faces are all near-uniform skips the network.
__device__ __forceinline__ bool ClassifyFace(const LineHybrid &hy, const float *nodeRP, int n, int f, bool valid,
                                             const artra::MlpWave &frags, const float *bias)
{
    const bool smooth = hy.smoothVariation > 0.0f &&
                        artra::near_uniform(nodeRP + f - 2, nodeRP + n + f - 2, hy.smoothVariation);
    bool troubled = false;
    if (__ballot(valid && !smooth)) {
        float x[artra::kFeat];
        artra::features_from(nodeRP + f - 2, nodeRP + n + f - 2, x);
        const float z = artra::mlp_logit_wave(x, frags, bias, hy.b3);
        troubled = valid && !smooth && z > hy.threshold;
    }
    return troubled;
}

// All stages of the directional operator on one line, run by the T threads
// tt = 0..T-1 (a multiple of the wavefront).  Every thread of the block must
// call it with the same loop trip counts (same n, lo, hi, stages), because
// the stages are separated by block barriers.  With write = false the line
// is computed but not stored (padding slots of a batched block).
//
// With HYB, the indicator classifies the faces once per operator (at its
// first RK stage, before the stage loop: the solution moves less than a
// cell within one operator, and the network's fragments are then dead in
// the flux passes) and the routing is per wavefront: a wavefront whose 64
// faces are all smooth computes the linear flux, one holding a troubled
// face computes the full WENO5 flux on all its faces (a branch the whole
// wavefront takes, never both branches), so a line crossed by a shock pays
// full WENO5 only on the wavefronts around the shock and no wavefront pays
// more than the plain kernel.
template <int S, int NS, bool FAST, bool HYB, int K>
__device__ __forceinline__ void LineBody(const LineArgs &la, const FluxModel &mdl, const LineHybrid &hy, long line,
                                         int tt, int T, char *mem, bool write, unsigned long long *tally,
                                         const float *bias, const artra::MlpWave &frags)
{
    constexpr int SL = (NS == 6) ? 2 : 1;           // stencil nodes left of the face's left node
    const int n = la.n, F = la.hi - la.lo + 1;
    double *st = (double *)mem;                     // [n][5]
    char *work = mem + LineStateBytes(n);
    double *fb = (double *)work;                    // [F][5]
    double *u0 = fb + 5 * F;                        // [F - 1][5]
    int *rules = (int *)(u0 + 5 * (F - 1));         // [2][kLineRules]: destinations, sources
    const double *cst = (const double *)(rules + 2 * kLineRules);   // [kLineRules][5] constant states
    float *nodeRP = (float *)(work + LineWorkPlainBytes(F));   // [2][n] (HYB)
    const long base = LineBase(la, line);
    for (int p = tt; p < n; p += T) {
        double v[5];
        LoadNode(la.U, base + (long)p * la.stride, v);
#pragma unroll
        for (int c = 0; c < 5; ++c) st[p * 5 + c] = v[c];
        if (p >= la.lo && p < la.hi) {
#pragma unroll
            for (int c = 0; c < 5; ++c) u0[(p - la.lo) * 5 + c] = v[c];
        }
        if (HYB) artra::node_rho_p(mdl.gamma, v, nodeRP[p], nodeRP[n + p]);
    }
    // the line's fill rules into LDS: read per lane from the argument
    // block once here, where the field loads hide the latency
    LoadRules(la, rules, tt, T);
    // the nodes this thread updates (every interior node of a fused line is
    // a fluid node: the fused path runs only where all are)
    bool own[K];
#pragma unroll
    for (int k = 0; k < K; ++k) own[k] = la.lo + tt + T * k < la.hi;
    __syncthreads();
    // between-stage refill duties: the ghosts this thread's nodes feed, and
    // (threads tt < nrules) one constant-state rule
    Feeds feed[K];
#pragma unroll
    for (int k = 0; k < K; ++k) feed[k] = FeedsOf(rules, la.nrules, la.lo + tt + T * k);
    const int constRule = (tt < la.nrules && rules[kLineRules + tt] < 0) ? tt : -1;
    const int M = (F + T - 1) / T;
    const double r = Uniform(la.dt ? *la.dt * la.dd : la.r);
    if (la.dt && *la.dt == 0.0) write = false;      // a batched case past its end time
    // With HYB the faces are dealt to the wavefronts rotated by a whole
    // wavefront per line: the troubled faces of a shock crossing every line
    // at the same place then fall on a different wavefront (and SIMD) from
    // line to line instead of loading one SIMD with all of them.
    const int QT = M * T;
    const int rot = HYB ? (int)(line % (QT / 64)) * 64 : 0;
    // bit m: this wavefront's faces of pass m take the full WENO5 flux
    unsigned wenoPasses = HYB ? 0u : ~0u;
    if (HYB) {
        // classification (every thread runs the same trip count)
        unsigned long long wenoFaces = 0, faces = 0, flagged = 0;
        for (int m = 0; m < M; ++m) {
            const int q = (tt + T * m + rot) % QT;
            const bool valid = q < F;
            const int f = la.lo - 1 + (valid ? q : F - 1);
            const bool troubled = ClassifyFace(hy, nodeRP, n, f, valid, frags, bias);
            const unsigned long long here = __popcll(__ballot(valid));
            const unsigned long long marked = __popcll(__ballot(troubled));
            if (marked) {
                wenoPasses |= 1u << m;
                wenoFaces += here;
                flagged += marked;
            }
            faces += here;
        }
        wenoPasses = __builtin_amdgcn_readfirstlane(wenoPasses);
        // tallies per wavefront: faces on the linear flux, faces classified,
        // faces on the full WENO5 flux, faces the indicator flagged troubled
        if (__lane_id() == 0 && write) {
            atomicAdd(&tally[0], faces - wenoFaces);
            atomicAdd(&tally[1], faces);
            if (wenoFaces) {
                atomicAdd(&tally[2], wenoFaces);
                atomicAdd(&tally[3], flagged);
            }
        }
    }
    for (int s = 0; s < la.stages; ++s) {
        for (int m = 0; m < M; ++m) {
            const int q = HYB ? (tt + T * m + rot) % QT : tt + T * m;
            if (q >= F) {
                if (HYB) continue;
                break;
            }
            const int f = la.lo - 1 + q;   // face between f and f + 1
            double Fx[5];
            if (HYB && !((wenoPasses >> m) & 1u)) {
                LinearFaceFlux<S>(mdl, st, f, Fx);
            } else if (FAST && NS == 6) {
                artra::weno_flux_line<S>(mdl, st, f, Fx);
            } else {
                double w[NS][5];
#pragma unroll
                for (int j = 0; j < NS; ++j)
#pragma unroll
                    for (int c = 0; c < 5; ++c) w[j][c] = st[(f - SL + j) * 5 + c];
                ConvFlux<S, NS, FAST>(mdl, w, Fx);
            }
#pragma unroll
            for (int c = 0; c < 5; ++c) fb[q * 5 + c] = Fx[c];
        }
        __syncthreads();
        // node updates (fluid_dynamics.c:243-274); each thread reads and
        // writes only its own nodes and, before another stage, refills the
        // ghosts they feed (the line's between-stage fill)
        const double a = la.a[s], b = la.b[s];
        const bool refill = s + 1 < la.stages;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (!own[k]) continue;
            const int i = la.lo + tt + T * k;
            const int q = i - la.lo;
            double v[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const double FL = fb[q * 5 + c], FR = fb[(q + 1) * 5 + c];
                const double phi = FAST ? FL - FR : ((FL - FR) + 0.0) - 0.0;
                const double Un = st[i * 5 + c];
                const double Uo = (s == 0) ? Un : u0[q * 5 + c];
                v[c] = FAST ? fma(a, Uo, b * fma(r, phi, Un)) : a * Uo + b * (Un + r * phi);
                st[i * 5 + c] = v[c];
            }
#pragma unroll
            for (int e = 0; e < kRefillPerNode; ++e) {
                const int d = FeedAt(feed[k], e);
                if (!refill || d < 0) continue;
#pragma unroll
                for (int c = 0; c < 5; ++c) st[d * 5 + c] = v[c];
            }
        }
        if (refill && constRule >= 0) {
#pragma unroll
            for (int c = 0; c < 5; ++c) st[rules[constRule] * 5 + c] = cst[constRule * 5 + c];
        }
        __syncthreads();
    }
    if (!write) return;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (!own[k]) continue;
        const int i = la.lo + tt + T * k;
#pragma unroll
        for (int c = 0; c < 5; ++c) NodeAt(la.U, base + (long)i * la.stride)[c] = st[i * 5 + c];
    }
}

// One line per block.
template <int S, int NS, bool FAST, bool HYB, int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HYB ? 3 : FAST ? 3 : 2))) void rk_line(LineArgs la, FluxModel mdl,
                                                                                        LineHybrid hy)
{
    if (la.dt && *la.dt == 0.0) return;             // a device-stepped march past its end time
    extern __shared__ double lds[];
    __shared__ unsigned long long tally[kTallies];
    __shared__ float bias[HYB ? 3 * artra::kHidden : 1];
    artra::MlpWave frags;
    if (HYB) {
        for (int e = threadIdx.x; e < 3 * artra::kHidden; e += blockDim.x) bias[e] = hy.bias[e];
        if (threadIdx.x < kTallies) tally[threadIdx.x] = 0ull;
        artra::mlp_wave_fragments(hy.W1, hy.W2, frags);
    }
    const long line = XcdLine(blockIdx.x, gridDim.x);
    LineBody<S, NS, FAST, HYB, K>(la, mdl, hy, line, threadIdx.x, blockDim.x, (char *)lds, true, tally, bias, frags);
    if (HYB && threadIdx.x == 0) {
        unsigned long long *slot = hy.counts + (blockIdx.x % kCountSlots) * kCountStride;
        atomicAdd(slot, tally[0]);
        atomicAdd(slot + 1, tally[1]);
        atomicAdd(slot + 2, tally[2]);
        atomicAdd(slot + 3, tally[3]);
    }
}

// The learned hybrid's lean pass: one line per block, T >= interior nodes
// (one node and at most two faces per thread).  The line is loaded and its
// faces classified exactly as LineBody classifies them (ClassifyFace); a
// line with any troubled face is not computed here but appended to defer[]
// (rk_line_list runs it again with the full hybrid body), every other line
// runs all RK stages with the linear fifth-order flux only (LinearFaceFlux):
// no WENO5 code and no network fragments past the classification, so it
// holds fewer registers and runs more lines per compute unit.  Its faces
// are those LineBody computes for the same line, bit for bit.
template <int S, bool MEAN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MEAN ? 6 : 4))) void rk_lean(
    LineArgs la, FluxModel mdl, LineHybrid hy, int *defer, int *ndefer, unsigned long long *deferTotal)
{
    if (la.dt && *la.dt == 0.0) return;             // a device-stepped march past its end time
    extern __shared__ double lds[];
    __shared__ float bias[3 * artra::kHidden];
    __shared__ int troubledAny;
    const int tt = threadIdx.x, T = blockDim.x;
    const int n = la.n, F = la.hi - la.lo + 1;
    double *st = lds;                                   // [n][5]
    double *fb = st + 5 * n;                            // [F][5]
    int *rules = (int *)(fb + 5 * F);                   // [2][kLineRules], constant states
    const double *cst = (const double *)(rules + 2 * kLineRules);
    float *nodeRP = (float *)((char *)rules + kRuleBytes);   // [2][n]
    const long line = XcdLine(blockIdx.x, gridDim.x);
    const int a1 = (int)(line % la.n1), a2 = (int)(line / la.n1);
    const long base = (long)(la.lo1 + a1) * la.stride1 + (long)(la.lo2 + a2) * la.stride2;
    artra::MlpWave frags;
    artra::mlp_wave_fragments(hy.W1, hy.W2, frags);
    for (int e = tt; e < 3 * artra::kHidden; e += T) bias[e] = hy.bias[e];
    for (int p = tt; p < n; p += T) {
        double v[5];
        LoadNode(la.U, base + (long)p * la.stride, v);
#pragma unroll
        for (int c = 0; c < 5; ++c) st[p * 5 + c] = v[c];
        artra::node_rho_p(mdl.gamma, v, nodeRP[p], nodeRP[n + p]);
    }
    LoadRules(la, rules, tt, T);
    const int i = la.lo + tt;                       // this thread's node
    const bool own = i < la.hi;
    if (tt == 0) troubledAny = 0;
    __syncthreads();
    // classification: the faces q = tt and q = tt + T (same trip count on
    // every thread, whole wavefronts run the network)
    const int M = (F + T - 1) / T;
    for (int m = 0; m < M; ++m) {
        const int q = tt + T * m;
        const bool valid = q < F;
        const int f = la.lo - 1 + (valid ? q : F - 1);
        const bool troubled = ClassifyFace(hy, nodeRP, n, f, valid, frags, bias);
        if (__ballot(troubled) && __lane_id() == 0) troubledAny = 1;
    }
    __syncthreads();
    if (troubledAny) {
        if (tt == 0) {
            defer[atomicAdd(ndefer, 1)] = (int)line;
            atomicAdd(deferTotal, 1ull);
        }
        return;
    }
    if (tt == 0) {
        unsigned long long *slot = hy.counts + (blockIdx.x % kCountSlots) * kCountStride;
        atomicAdd(slot, (unsigned long long)F);
        atomicAdd(slot + 1, (unsigned long long)F);
    }
    double U0[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) U0[c] = st[(own ? i : la.lo) * 5 + c];
    const double r = Uniform(la.dt ? *la.dt * la.dd : la.r);
    // between-stage refill duties (as in LineBody)
    const Feeds feed = FeedsOf(rules, la.nrules, i);
    const int constRule = (tt < la.nrules && rules[kLineRules + tt] < 0) ? tt : -1;
    for (int s = 0; s < la.stages; ++s) {
        for (int m = 0; m < M; ++m) {
            const int q = tt + T * m;
            if (q >= F) break;
            double Fx[5];
            LinearFaceFlux<S, MEAN>(mdl, st, la.lo - 1 + q, Fx);
#pragma unroll
            for (int c = 0; c < 5; ++c) fb[q * 5 + c] = Fx[c];
        }
        __syncthreads();
        // node update (fluid_dynamics.c:243-274) and, before another
        // stage, the refill of the ghosts this node feeds
        const bool refill = s + 1 < la.stages;
        if (own) {
            const double a = la.a[s], b = la.b[s];
            const int q = i - la.lo;
            double v[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const double phi = fb[q * 5 + c] - fb[(q + 1) * 5 + c];
                const double Un = st[i * 5 + c];
                const double Uo = (s == 0) ? Un : U0[c];
                v[c] = fma(a, Uo, b * fma(r, phi, Un));
                st[i * 5 + c] = v[c];
            }
#pragma unroll
            for (int e = 0; e < kRefillPerNode; ++e) {
                const int d = FeedAt(feed, e);
                if (!refill || d < 0) continue;
#pragma unroll
                for (int c = 0; c < 5; ++c) st[d * 5 + c] = v[c];
            }
        }
        if (refill && constRule >= 0) {
#pragma unroll
            for (int c = 0; c < 5; ++c) st[rules[constRule] * 5 + c] = cst[constRule * 5 + c];
        }
        __syncthreads();
    }
    if (own) {
#pragma unroll
        for (int c = 0; c < 5; ++c) NodeAt(la.U, base + (long)i * la.stride)[c] = st[i * 5 + c];
    }
}

// LDS of rk_lean: states [n][5], face fluxes [F][5], fill rules and
// constant states, FP32 density and pressure [2][n].
__host__ __device__ __forceinline__ size_t LeanLdsBytes(int n, int F)
{
    return (size_t)(5 * n + 5 * F) * sizeof(double) + kRuleBytes + (size_t)2 * n * sizeof(float);
}

// The full hybrid kernel over a list of lines (those rk_lean deferred):
// blocks walk the list, whose length is read on the device.
template <int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void rk_line_list(LineArgs la,
                                                                                          FluxModel mdl,
                                                                                          LineHybrid hy,
                                                                                          const int *list,
                                                                                          const int *count)
{
    if (la.dt && *la.dt == 0.0) return;
    const int total = *count;
    if ((int)blockIdx.x >= total) return;
    extern __shared__ double lds[];
    __shared__ unsigned long long tally[kTallies];
    __shared__ float bias[3 * artra::kHidden];
    for (int e = threadIdx.x; e < 3 * artra::kHidden; e += blockDim.x) bias[e] = hy.bias[e];
    if (threadIdx.x < kTallies) tally[threadIdx.x] = 0ull;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        // the network's fragments are loaded per line (cache hits) and die
        // after its classification; an opaque memory effect keeps the loads
        // from being hoisted out of the loop, where they would hold
        // registers through every stage
        asm volatile("" ::: "memory");
        artra::MlpWave frags;
        artra::mlp_wave_fragments(hy.W1, hy.W2, frags);
        LineBody<S, 6, true, true, 1>(la, mdl, hy, list[w], threadIdx.x, blockDim.x, (char *)lds, true, tally, bias,
                                      frags);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        unsigned long long *slot = hy.counts + (blockIdx.x % kCountSlots) * kCountStride;
        atomicAdd(slot, tally[0]);
        atomicAdd(slot + 1, tally[1]);
        atomicAdd(slot + 2, tally[2]);
        atomicAdd(slot + 3, tally[3]);
    }
}

// Many independent cases in one persistent launch: blocks of 256 threads
// walk a tile list; a tile is P = 256 / T lines of one case, each run by T
// threads (the tile shape picked per grid size).  Lines past the case's end
// in its last tile are computed on a copy of its last line and not stored.
struct BatchTile {
    int c;                          // case
    int T;                          // threads per line
    int count;                      // lines in this tile (<= 256 / T)
    int pad;
    long first;                     // first line of the tile
};

template <int S, int NS, bool HYB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void rk_batch(const BatchTile *tiles,
                                                                                         int ntiles,
                                                                                         const LineArgs *args,
                                                                                         FluxModel mdl, LineHybrid hy)
{
    extern __shared__ double lds[];
    __shared__ unsigned long long tally[kTallies];
    __shared__ float bias[HYB ? 3 * artra::kHidden : 1];
    if (HYB) {
        for (int e = threadIdx.x; e < 3 * artra::kHidden; e += blockDim.x) bias[e] = hy.bias[e];
        if (threadIdx.x < kTallies) tally[threadIdx.x] = 0ull;
    }
    for (long w = blockIdx.x; w < ntiles; w += gridDim.x) {
        const BatchTile bt = tiles[w];
        const LineArgs &la = args[bt.c];
        // loaded per tile (L2 hits, issued ahead of the field loads) so that
        // they hold no registers past the tile's classification; an opaque
        // memory effect keeps the loads from being hoisted out of the loop
        if (HYB) asm volatile("" ::: "memory");
        artra::MlpWave frags;
        if (HYB) artra::mlp_wave_fragments(hy.W1, hy.W2, frags);
        const int slot = threadIdx.x / bt.T, tt = threadIdx.x % bt.T;
        const bool write = slot < bt.count;
        const long line = bt.first + (write ? slot : bt.count - 1);
        char *mem = (char *)lds + (size_t)slot * LineLdsBytes(la.n, la.hi - la.lo + 1, HYB);
        LineBody<S, NS, true, HYB, 1>(la, mdl, hy, line, tt, bt.T, mem, write, tally, bias, frags);
        __syncthreads();
    }
    if (HYB && threadIdx.x == 0) {
        unsigned long long *slot = hy.counts + (blockIdx.x % kCountSlots) * kCountStride;
        atomicAdd(slot, tally[0]);
        atomicAdd(slot + 1, tally[1]);
        atomicAdd(slot + 2, tally[2]);
        atomicAdd(slot + 3, tally[3]);
    }
}

// ------------------------------------------------------- immersed boundary
struct GhostDev {
    int node, kind, first, count;
    int noslip, isothermal;
    double wallT, wO;
    double Vs[3], N[3], Ta[3], Tb[3];
};

// Node flags of a box, packed k-j-i, written into the grid's flag array.
struct FlagBox {
    int lo[3], n[3];
};

__global__ __launch_bounds__(256) void scatter_flags(uint8_t *fluid, const uint8_t *packed, Grid g, FlagBox b,
                                                     long count)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int i = b.lo[0] + (int)(t % b.n[0]);
    const int j = b.lo[1] + (int)((t / b.n[0]) % b.n[1]);
    const int k = b.lo[2] + (int)(t / ((long)b.n[0] * b.n[1]));
    fluid[((long)k * g.ny + j) * g.nx + i] = packed[t];
}

// ------------------------------------------ moving-body classification
// (the grid geometry and search path: ArtraGeoSetup, artra_device.h)
typedef ArtraGeoSetup GeoSetup;

// A body's arrays on the device (triangulated bodies; the connectivity is
// uploaded once, the positions and normals after every move).
struct GeoBodyDev {
    int faceN = 0, edgeN = 0, vertN = 0;
    int *f = nullptr, *e = nullptr;
    double *v = nullptr, *Nf = nullptr, *Ne = nullptr, *Nv = nullptr;   // parts of block
    double *block = nullptr;
    double *stage = nullptr;        // pinned host copy of block for the upload
    double4 *sph = nullptr;         // each facet's bounding sphere (centre, radius)
    double4 *bsph = nullptr;        // each 32-facet block's bounding sphere
    bool uploaded = false;
};

// A node box [lo, hi) and its node count.
struct GeoBox {
    int lo[3], n[3];
    long count;
};

__device__ __forceinline__ long GeoNode(const Grid &g, const GeoBox &b, long t, int &i, int &j, int &k)
{
    i = b.lo[0] + (int)(t % b.n[0]);
    j = b.lo[1] + (int)((t / b.n[0]) % b.n[1]);
    k = b.lo[2] + (int)(t / ((long)b.n[0] * b.n[1]));
    return ((long)k * g.ny + j) * g.nx + i;
}

__device__ __forceinline__ void GeoPoint(const GeoSetup &gs, int i, int j, int k, double p[3])
{
    p[0] = gs.smin[0] + (i - gs.ng[0]) * gs.d[0];
    p[1] = gs.smin[1] + (j - gs.ng[1]) * gs.d[1];
    p[2] = gs.smin[2] + (k - gs.ng[2]) * gs.d[2];
}

// Pass 1 (immersed_boundary.c:72-113): keep the domain field in gst, clear
// the fluid nodes' face links, release the interfacial nodes of moving bodies.
__global__ __launch_bounds__(256) void geo_reset(Grid g, GeoBox b, int *did, int *fid, const int *lid, int *gst,
                                                 const int *moving, int nbody)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.count) return;
    int i, j, k;
    const long q = GeoNode(g, b, t, i, j, k);
    const int gid = did[q];
    gst[q] = gid;
    if (0 == gid) {
        fid[q] = 0;
        return;
    }
    if (gid <= nbody && !moving[gid - 1]) return;
    if (0 < lid[q]) did[q] = 0;
}

// Pass 2 for a sphere (immersed_boundary.c:128-182): nodes of its box not yet
// classified that lie in or on it.
__global__ __launch_bounds__(256) void geo_sphere(Grid g, GeoSetup gs, GeoBox b, int *did, int *fid, double Ox,
                                                  double Oy, double Oz, double r, int body)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.count) return;
    int i, j, k;
    const long q = GeoNode(g, b, t, i, j, k);
    if (0 != did[q]) return;
    double p[3];
    GeoPoint(gs, i, j, k, p);
    const double dv[3] = {p[0] - Ox, p[1] - Oy, p[2] - Oz};
    if (r * r >= GcDot(dv, dv)) {
        did[q] = body + 1;
        fid[q] = 0;
    }
}

// Pass 2 for a triangulated body (computational_geometry.c:463-506): one
// wavefront per node; the lanes measure the facets that bounding spheres
// (of 32-facet blocks, then of facets) cannot rule out, the lanes'
// (distance, facet) pairs reduce to the lexicographic minimum (the first
// facet at the global minimum, as the serial scan), and the pseudo-normal
// at that facet decides.
// Bounding spheres for the closest-facet search below: per facet its
// centroid and radius, per block of 32 consecutive facets (a half-wavefront
// here) the mean of their centroids and the farthest reach of their spheres
// from it.
__global__ __launch_bounds__(256) void geo_spheres(ArtraPolyView pv, double4 *sph, double4 *bsph)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = n < pv.faceN;
    double c[3] = {0.0, 0.0, 0.0}, r2 = 0.0;
    if (on) {
        const double *v0 = pv.v[pv.f[n][0]], *v1 = pv.v[pv.f[n][1]], *v2 = pv.v[pv.f[n][2]];
        for (int s = 0; s < 3; ++s) c[s] = (v0[s] + v1[s] + v2[s]) / 3.0;
        const double *vs[3] = {v0, v1, v2};
        for (int q = 0; q < 3; ++q) {
            const double d[3] = {vs[q][0] - c[0], vs[q][1] - c[1], vs[q][2] - c[2]};
            r2 = fmax(r2, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        }
        sph[n] = make_double4(c[0], c[1], c[2], sqrt(r2));
    }
    const double r = sqrt(r2);
    // the block's mean centroid and reach over its half-wavefront
    double m[3] = {c[0], c[1], c[2]};
    int cnt = on ? 1 : 0;
    for (int off = 16; off > 0; off >>= 1) {
        for (int s = 0; s < 3; ++s) m[s] += __shfl_xor(m[s], off, 32);
        cnt += __shfl_xor(cnt, off, 32);
    }
    if (cnt == 0) return;                           // the whole half-wavefront past the facets
    for (int s = 0; s < 3; ++s) m[s] /= (double)cnt;
    double reach = 0.0;
    if (on) {
        const double d[3] = {c[0] - m[0], c[1] - m[1], c[2] - m[2]};
        reach = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + r;
    }
    for (int off = 16; off > 0; off >>= 1) reach = fmax(reach, __shfl_xor(reach, off, 32));
    if ((n & 31) == 0) bsph[n >> 5] = make_double4(m[0], m[1], m[2], reach);
}

// Pass 2 for a triangulated body (computational_geometry.c:463-506): one
// wavefront per node; the lanes measure the facets that bounding spheres
// (of 32-facet blocks, then of facets) cannot rule out, the lanes'
// (distance, facet) pairs reduce to the lexicographic minimum (the first
// facet at the global minimum, as the serial scan), and the pseudo-normal
// at that facet decides.
// A bounding sphere per facet (centroid, radius) for the closest-facet
// search below.
__global__ __launch_bounds__(256) void geo_spheres(ArtraPolyView pv, double4 *sph)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= pv.faceN) return;
    const double *v0 = pv.v[pv.f[n][0]], *v1 = pv.v[pv.f[n][1]], *v2 = pv.v[pv.f[n][2]];
    double c[3], r2 = 0.0;
    for (int s = 0; s < 3; ++s) c[s] = (v0[s] + v1[s] + v2[s]) / 3.0;
    const double *vs[3] = {v0, v1, v2};
    for (int q = 0; q < 3; ++q) {
        const double d[3] = {vs[q][0] - c[0], vs[q][1] - c[1], vs[q][2] - c[2]};
        r2 = fmax(r2, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    }
    sph[n] = make_double4(c[0], c[1], c[2], sqrt(r2));
}

// A bounding sphere per block of 32 consecutive facets: the mean of their
// centroids and the farthest reach of their spheres from it.
__global__ __launch_bounds__(64) void geo_blocks(int faceN, const double4 *sph, double4 *bsph)
{
    const int bk = blockIdx.x * blockDim.x + threadIdx.x;
    const int first = bk * 32, last = min(first + 32, faceN);
    if (first >= faceN) return;
    double c[3] = {0.0, 0.0, 0.0};
    for (int n = first; n < last; ++n) {
        c[0] += sph[n].x;
        c[1] += sph[n].y;
        c[2] += sph[n].z;
    }
    for (int s = 0; s < 3; ++s) c[s] /= (double)(last - first);
    double r = 0.0;
    for (int n = first; n < last; ++n) {
        const double d[3] = {sph[n].x - c[0], sph[n].y - c[1], sph[n].z - c[2]};
        r = fmax(r, sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + sph[n].w);
    }
    bsph[bk] = make_double4(c[0], c[1], c[2], r);
}

__global__ __launch_bounds__(256) void geo_mesh(Grid g, GeoSetup gs, GeoBox b, int *did, int *fid, ArtraPolyView pv,
                                                const double4 *sph, const double4 *bsph, int body)
{
    const long t = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (t >= b.count) return;                       // whole wavefronts leave together
    const int lane = __lane_id();
    int i, j, k;
    const long q = GeoNode(g, b, t, i, j, k);
    if (0 != did[q]) return;                        // the same node on every lane
    double p[3];
    GeoPoint(gs, i, j, k, p);
    // a bound first: the squared distance U to the nearest of the facet
    // blocks' first centroids is at least the minimum; a block or facet
    // whose bounding sphere lies farther than U (with a margin far above
    // rounding) cannot hold the minimum and is skipped; the others are
    // measured, each lane keeping the lexicographic minimum of (distance,
    // facet), which the serial scan's first facet at the minimum is
    const int nblk = (pv.faceN + 31) >> 5;
    double U = DBL_MAX, R2 = 0.0;
    for (int bk = lane; bk < nblk; bk += 64) {
        const double4 c = sph[bk * 32];
        const double d[3] = {p[0] - c.x, p[1] - c.y, p[2] - c.z};
        U = fmin(U, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        R2 = fmax(R2, bsph[bk].w * bsph[bk].w);
    }
    for (int off = 32; off > 0; off >>= 1) {
        U = fmin(U, __shfl_xor(U, off));
        R2 = fmax(R2, __shfl_xor(R2, off));
    }
    const double thr = U + 1e-6 * (U + R2);
    double best = FLT_MAX;                          // the serial scan's start: facet 0 at FLT_MAX
    int cid = 0;
    for (int base = 0; base < nblk; base += 64) {
        bool live = false;
        if (base + lane < nblk) {
            const double4 c = bsph[base + lane];
            const double d[3] = {p[0] - c.x, p[1] - c.y, p[2] - c.z};
            const double dc = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            const double gap = dc - c.w - 1e-9 * (dc + c.w);
            live = !(gap > 0.0 && gap * gap > thr);
        }
        unsigned long long m = __ballot(live);
        while (m) {                                 // two blocks at a time, a half-wavefront each
            const int b0 = __ffsll((long long)m) - 1;
            m &= m - 1;
            int b1 = -1;
            if (m) {
                b1 = __ffsll((long long)m) - 1;
                m &= m - 1;
            }
            const int mb = (lane < 32) ? b0 : b1;
            const int n = (base + mb) * 32 + (lane & 31);
            if (mb < 0 || n >= pv.faceN) continue;
            const double4 c = sph[n];
            const double d[3] = {p[0] - c.x, p[1] - c.y, p[2] - c.z};
            const double dc = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            const double gap = dc - c.w - 1e-9 * (dc + c.w);
            if (gap > 0.0 && gap * gap > thr) continue;
            const double d2 = GcFacetDistance2(p, n, &pv);
            if (d2 < best || (d2 == best && n < cid)) {
                best = d2;
                cid = n;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oc = __shfl_xor(cid, off);
        if (ob < best || (ob == best && oc < cid)) {
            best = ob;
            cid = oc;
        }
    }
    if (lane == 0 && GcInsideOfFacet(p, cid, &pv)) {
        did[q] = body + 1;
        fid[q] = cid;
    }
}

// The first search-path layer (1..gl) holding a node whose did equals (same)
// or differs from (!same) `did`, or 0 (immersed_boundary.c:227-264).
__device__ int GeoInterState(const Grid &g, const GeoSetup &gs, const int *didv, int k, int j, int i, int did,
                             bool same)
{
    for (int n = 0; n < gs.pathSep[0]; ++n) {
        const int kh = k + gs.path[n][2], jh = j + gs.path[n][1], ih = i + gs.path[n][0];
        if (kh < g.lo[2] || kh >= g.hi[2] || jh < g.lo[1] || jh >= g.hi[1] || ih < g.lo[0] || ih >= g.hi[0]) continue;
        const int other = didv[((long)kh * g.ny + jh) * g.nx + ih];
        if ((same && other == did) || (!same && other != did)) {
            for (int r = 1; r <= gs.gl; ++r) {
                if (gs.pathSep[r] > n) return r;
            }
        }
    }
    return 0;
}

// Pass 3's interfacial states (immersed_boundary.c:183-225, which read did
// only), the sweeps' fluid flags, and the region's flags packed for the host:
// packed[0..4][t] = did, fid, lid, gst and the previous domain field.
__global__ __launch_bounds__(256) void geo_interface(Grid g, GeoSetup gs, GeoBox b, const int *did, const int *fid,
                                                     int *lid, int *gst, uint8_t *fluid, int *packed)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.count) return;
    int i, j, k;
    const long q = GeoNode(g, b, t, i, j, k);
    const int dq = did[q];
    int l = 0, s = 0;
    if (0 != dq) {
        l = GeoInterState(g, gs, did, k, j, i, dq, false);
        if (0 < l) s = GeoInterState(g, gs, did, k, j, i, 0, true);
    }
    packed[t] = dq;
    packed[b.count + t] = fid[q];
    packed[2 * b.count + t] = l;
    packed[3 * b.count + t] = s;
    packed[4 * b.count + t] = gst[q];               // the previous domain field (pass 1)
    lid[q] = l;
    gst[q] = s;
    fluid[q] = (0 == dq);
}

// The ghost table on the device (immersed_boundary.c:270-487 via
// artra_build_ghost_table): a wavefront per 64 nodes of a body's box; a
// ghost of layer r takes the next entry of the table and kGhostSlot stencil
// entries, found by the same search (radius 2 around the image point's node
// over fluid nodes for an image-point ghost, radius 1 around itself over
// the previous layer's ghosts of the body otherwise, growing to 3).
constexpr int kGhostSlot = 343;
constexpr int kGeoCounters = 16;    // ghost-table build counters: taken, overflow, layer ends

struct GhostBuildArgs {
    Grid g;
    GeoSetup gs;
    double dd[3];
    double tinyL;
    int ibmLayer;
    int cap;                        // ghost entries available
};

struct GhostBody {
    ArtraPolyView pv;
    double O[3], r, V[3], W[3], cf, T;
    int body;
};

__device__ __forceinline__ double GbInverse(double tiny, double d2)
{
    if (tiny > d2) d2 = tiny;
    return 1.0 / d2;
}

__device__ __forceinline__ double GbDist2(const double a[3], const double b[3])
{
    const double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    return GcDot(d, d);
}

__global__ __launch_bounds__(256) void ib_build(GhostBuildArgs a, GeoBox b, int r, GhostBody body, const int *did,
                                                const int *fid, const int *gst, GhostDev *ghosts, int *snode,
                                                double *sw, int *counter, int *overflow)
{
    // each wavefront covers 64 nodes of the box and searches their ghosts'
    // stencils one at a time, the lanes scanning a search box together:
    // candidate c of the (2 rr + 1)^3 box is (kh, jh, ih) in k-j-i order,
    // and the found nodes take consecutive entries in that order (a ballot
    // prefix), as the serial search appends them
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const Grid &g = a.g;
    const int lane = __lane_id();
    int i = 0, j = 0, k = 0;
    bool mine = false;
    if (t < b.count) {
        const long q = GeoNode(g, b, t, i, j, k);
        mine = (r == gst[q] && body.body + 1 == did[q]);
    }
    // each ghost's search centre on its own lane, all at once: the image
    // point (through the body's surface) or the node itself
    const bool image = (a.ibmLayer >= r);
    double c[3] = {0.0, 0.0, 0.0}, pO[3] = {0.0, 0.0, 0.0}, pI[3] = {0.0, 0.0, 0.0}, N[3] = {0.0, 0.0, 0.0};
    int nc[3] = {i, j, k};
    if (mine) {
        double pG[3];
        GeoPoint(a.gs, i, j, k, pG);
        if (image) {
            GcGeometricData(pG, fid[((long)k * g.ny + j) * g.nx + i], &body.pv, body.O, body.r, pO, pI, N);
            for (int s = 0; s < 3; ++s) {
                c[s] = pI[s];
                nc[s] = GcMapNode(pI[s], a.gs.smin[s], a.dd[s], a.gs.ng[s]);
            }
        } else {
            for (int s = 0; s < 3; ++s) c[s] = pG[s];
        }
    }
    const int h = image ? 2 : 1, want = image ? 0 : body.body + 1, layer = image ? 0 : r - 1;
    unsigned long long todo = __ballot(mine);
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // the wavefront's ghosts take consecutive entries from one reservation
    int e = 0;
    if (lane == 0 && todo) e = atomicAdd(counter, __popcll(todo));
    e = __shfl(e, 0) - 1;
    int myEntry = -1, myCount = 0;
    while (todo) {
        const int src = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        ++e;
        if (e >= a.cap) {
            if (lane == 0) atomicOr(overflow, 1);
            continue;
        }
        double cs[3];
        int ns[3];
        for (int s = 0; s < 3; ++s) {
            cs[s] = __shfl(c[s], src);
            ns[s] = __shfl(nc[s], src);
        }
        const int first = e * kGhostSlot;
        int count = 0;
        bool lost = false;
        for (int rr = h; 0 == count; ++rr) {
            if (rr > 3) {
                lost = true;
                break;
            }
            const int w = 2 * rr + 1, cand = w * w * w;
            // every chunk's flags are loaded before the first ballot, so the
            // loads of a search overlap
            constexpr int kChunks = (kGhostSlot + 63) / 64;
            int qv[kChunks];
            bool hv[kChunks];
#pragma unroll
            for (int ch = 0; ch < kChunks; ++ch) {
                const int cc = ch * 64 + lane;
                hv[ch] = false;
                qv[ch] = -1;
                if (cc < cand) {
                    const int x = ns[0] + cc % w - rr, y = ns[1] + (cc / w) % w - rr, z = ns[2] + cc / (w * w) - rr;
                    if (z >= g.lo[2] && z < g.hi[2] && y >= g.lo[1] && y < g.hi[1] && x >= g.lo[0] && x < g.hi[0])
                        qv[ch] = (int)(((long)z * g.ny + y) * g.nx + x);
                }
                if (qv[ch] >= 0) {
                    const int dq = did[qv[ch]], gq2 = gst[qv[ch]];
                    hv[ch] = (want == dq) && (layer == 0 || layer == gq2);
                }
            }
#pragma unroll
            for (int ch = 0; ch < kChunks; ++ch) {
                if (ch * 64 >= cand) break;
                const unsigned long long hits = __ballot(hv[ch]);
                if (hv[ch]) {
                    const int cc = ch * 64 + lane;
                    const int x = ns[0] + cc % w - rr, y = ns[1] + (cc / w) % w - rr, z = ns[2] + cc / (w * w) - rr;
                    double ph[3];
                    GeoPoint(a.gs, x, y, z, ph);
                    const int pos = count + __popcll(hits & below);
                    snode[first + pos] = qv[ch];
                    sw[first + pos] = GbInverse(a.tinyL, GbDist2(cs, ph));
                }
                count += __popcll(hits);
            }
        }
        if (lost) {
            if (lane == 0) atomicOr(overflow, 1);
            continue;
        }
        if (lane == src) {
            myEntry = e;
            myCount = count;
        }
    }
    // each ghost's record from its own lane
    if (myEntry >= 0) {
        GhostDev gh;
        memset(&gh, 0, sizeof gh);
        gh.node = (int)(((long)k * g.ny + j) * g.nx + i);
        gh.first = myEntry * kGhostSlot;
        gh.count = myCount;
        if (image) {
            gh.kind = 0;
            gh.wO = GbInverse(a.tinyL, GbDist2(pI, pO));
            const double arm[3] = {pO[0] - body.O[0], pO[1] - body.O[1], pO[2] - body.O[2]};
            GcCross(body.W, arm, gh.Vs);
            for (int s = 0; s < 3; ++s) gh.Vs[s] = body.V[s] + gh.Vs[s];
            for (int s = 0; s < 3; ++s) gh.N[s] = N[s];
            GcOrthogonalSpace(N, gh.Ta, gh.Tb);
            gh.noslip = (0.0 < body.cf);
            gh.isothermal = !(0.0 > body.T);
            gh.wallT = body.T;
        } else {
            gh.kind = 1;
        }
        ghosts[myEntry] = gh;
    }
}

// The surface-force table of one body on the device (the host's
// artra_build_force_table over the device's flags): one workgroup scans the
// body's node box in k-j-i order and appends its layer-2 ghosts in that
// order (a block-wide prefix of ballots), with the normal, arm and distance
// of each; it counts the body's interfacial layer-2 nodes.  counters[0] is
// the table's running end, counters[1 + 2 slot] and [2 + 2 slot] the
// body's ghost and interfacial counts.
constexpr int kForceBuildThreads = 1024;

__global__ __launch_bounds__(kForceBuildThreads) void force_build(Grid g, GeoSetup gs, GeoBox b, GhostBody body,
                                                                  int slot, int viscous, const int *did,
                                                                  const int *fid, const int *lid, const int *gst,
                                                                  ArtraForceNode *nodes, ArtraForceBody *bodies,
                                                                  int cap, int *counters)
{
    __shared__ int waveHits[kForceBuildThreads / 64];
    __shared__ int lidCount;
    const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int first = counters[0];
    if (tid == 0) lidCount = 0;
    __syncthreads();
    int end = first;
    for (long start = 0; start < b.count; start += kForceBuildThreads) {
        const long t = start + tid;
        bool hit = false;
        long q = 0;
        if (t < b.count) {
            int i, j, k;
            q = GeoNode(g, b, t, i, j, k);
            const int d = did[q];
            if (body.body + 1 == d) {
                hit = (2 == gst[q]);
                if (2 == lid[q]) atomicAdd(&lidCount, 1);
            }
        }
        const unsigned long long m = __ballot(hit);
        __syncthreads();                            // the last chunk's prefix is read
        if (lane == 0) waveHits[wave] = __popcll(m);
        __syncthreads();
        int before = end, total = 0;
        for (int w = 0; w < kForceBuildThreads / 64; ++w) {
            if (w < wave) before += waveHits[w];
            total += waveHits[w];
        }
        if (hit) {
            const int pos = before + __popcll(m & below);
            if (pos < cap) nodes[pos].node = (int)q;    // the geometry once every entry is placed
        }
        end += total;
    }
    __syncthreads();
    // each entry's normal, arm and distance, all entries at once
    for (int e = first + tid; e < end && e < cap; e += kForceBuildThreads) {
        const long q = nodes[e].node;
        const int i = (int)(q % g.nx), j = (int)((q / g.nx) % g.ny), k = (int)(q / ((long)g.nx * g.ny));
        double pG[3], pO[3], pI[3], N[3];
        GeoPoint(gs, i, j, k, pG);
        GcGeometricData(pG, fid[q], &body.pv, body.O, body.r, pO, pI, N);
        ArtraForceNode rec;
        rec.node = (int)q;
        rec.body = slot;
        for (int s = 0; s < 3; ++s) {
            rec.N[s] = N[s];
            rec.r[s] = pO[s] - body.O[s];
        }
        rec.dist = sqrt(GbDist2(pG, pO));
        nodes[e] = rec;
    }
    if (tid == 0) {
        counters[0] = end;
        counters[1 + 2 * slot] = end - first;
        counters[2 + 2 * slot] = lidCount;
        ArtraForceBody fb;
        fb.first = first;
        fb.count = end - first;
        fb.viscous = viscous;
        fb.pad = 0;
        for (int s = 0; s < 3; ++s) {
            fb.V[s] = body.V[s];
            fb.W[s] = body.W[s];
        }
        bodies[slot] = fb;
    }
}

// A table of filled nodes on the device with its stencil entries; the
// buffers grow as needed and are reused when the table is rebuilt.
struct IbTable {
    GhostDev *ghosts = nullptr;
    int *snode = nullptr;
    double *sw = nullptr;
    int count = 0;
    int ghostCap = 0, stencilCap = 0;
};

// The ghost node's state from the weighted sums of its stencil's primitive
// states (Uo) and weights (wsum): for a first-layer ghost the boundary
// point's state joins the stencil, then the state is reflected across the
// wall (immersed_boundary.c:345-392).
__device__ __forceinline__ void IbFinish(double *U, const GhostDev &gh, double Uo[6], double wsum, double gamma,
                                         double gasR)
{
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
    StoreNode(U, gh.node, out);
}

// One 64-lane block per filled node: the lanes load and convert 64 stencil
// entries at a time into LDS (the loop's memory latency overlapped), then
// lane c < 6 sums component c of the weighted primitive states and lane 6
// the weights, each over the entries in stencil order: the same operations
// as a loop over the stencil (immersed_boundary.c:393-440), so the node's
// state is bit for bit the reference's.  Grid: one block per entry of
// [begin, end).
__global__ __launch_bounds__(64) void ib_fill(double *U, const GhostDev *ghosts, int begin, const int *snode,
                                              const double *sw, double gamma, double gasR)
{
    __shared__ double stage[64 * 7];
    __shared__ double part[7];
    const int lane = threadIdx.x;
    const GhostDev gh = ghosts[begin + blockIdx.x];
    double acc = 0.0;                               // lane c < 6: Uo[c]; lane 6: the weight sum
    const int last = gh.first + gh.count;
    for (int base = gh.first; base < last; base += 64) {
        const int q = base + lane;
        if (q < last) {
            double Uh[5], Uoh[6];
            LoadNode(U, snode[q], Uh);
            Primitive(gamma, gasR, Uh, Uoh);
#pragma unroll
            for (int c = 0; c < 6; ++c) stage[lane * 7 + c] = Uoh[c];
            stage[lane * 7 + 6] = sw[q];
        }
        __syncthreads();
        const int m = min(64, last - base);
        if (lane < 6) {
            for (int e = 0; e < m; ++e) acc = acc + stage[e * 7 + lane] * stage[e * 7 + 6];
        } else if (lane == 6) {
            for (int e = 0; e < m; ++e) acc = acc + stage[e * 7 + 6];
        }
        __syncthreads();
    }
    if (lane < 7) part[lane] = acc;
    __syncthreads();
    if (lane == 0) {
        double Uo[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) Uo[c] = part[c];
        IbFinish(U, gh, Uo, part[6], gamma, gasR);
    }
}

// ----------------------------------------------------------- surface force
// Per layer-2 ghost node of a body: the pressure force p N, the wall shear
// mu (V_rel - (V_rel . N) N) / |pG - pO| of a no-slip body in a viscous
// flow, and their torque r x (Fp + Fv) (solid_dynamics.c:110-141), in the
// reference's operation order; terms[t] = Fp[3], Fv[3], Tt[3], p.
__global__ __launch_bounds__(256) void force_terms(const double *U, const ArtraForceNode *nodes, int count,
                                                   const ArtraForceBody *bodies, double gamma, double gasR,
                                                   double refMu, double refT, double *terms)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const ArtraForceNode e = nodes[t];
    const ArtraForceBody &b = bodies[e.body];
    double Un[5], Uo[6], Fp[3], Fv[3], Fs[3], Tt[3];
    LoadNode(U, e.node, Un);
    Primitive(gamma, gasR, Un, Uo);
#pragma unroll
    for (int s = 0; s < 3; ++s) Fp[s] = Uo[4] * e.N[s];
    if (b.viscous) {
        const double Ts = Uo[5] * refT;
        const double mu = refMu * (1.458e-6 * pow(Ts, 1.5) / (Ts + 110.4));
        double V[3];
        V[0] = b.W[1] * e.r[2] - b.W[2] * e.r[1];
        V[1] = b.W[2] * e.r[0] - b.W[0] * e.r[2];
        V[2] = b.W[0] * e.r[1] - b.W[1] * e.r[0];
#pragma unroll
        for (int s = 0; s < 3; ++s) V[s] = Uo[1 + s] - (b.V[s] + V[s]);
        const double Vn = V[0] * e.N[0] + V[1] * e.N[1] + V[2] * e.N[2];
#pragma unroll
        for (int s = 0; s < 3; ++s) Fv[s] = mu * (V[s] - Vn * e.N[s]) / e.dist;
    } else {
        Fv[0] = Fv[1] = Fv[2] = 0.0;
    }
#pragma unroll
    for (int s = 0; s < 3; ++s) Fs[s] = Fp[s] + Fv[s];
    Tt[0] = e.r[1] * Fs[2] - e.r[2] * Fs[1];
    Tt[1] = e.r[2] * Fs[0] - e.r[0] * Fs[2];
    Tt[2] = e.r[0] * Fs[1] - e.r[1] * Fs[0];
    double *o = terms + (size_t)t * 10;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        o[s] = Fp[s];
        o[3 + s] = Fv[s];
        o[6 + s] = Tt[s];
    }
    o[9] = Uo[4];
}

// Per body, the sums of its nodes' terms in table order (the reference's
// loop order, so the sums are those of its serial loop) and the pressure
// statistics: fvar = (p of the first node, sum (p - p0), sum (p - p0)^2).
// One 64-lane block per body: 64 nodes' terms at a time are staged in LDS,
// then lane q < 9 sums term q and lanes 9 and 10 the two statistics, each
// over the nodes in table order (the serial loop's additions, without its
// memory latency per node).
__global__ __launch_bounds__(64) void force_sum(const ArtraForceBody *bodies, int nbody, const double *terms,
                                                double *sums)
{
    __shared__ double stage[64 * 10];
    const int b = blockIdx.x;
    if (b >= nbody) return;                         // uniform over the block
    const int lane = threadIdx.x;
    const ArtraForceBody bd = bodies[b];
    const int last = bd.first + bd.count;
    const double p0 = (bd.count > 0) ? terms[(size_t)bd.first * 10 + 9] : 0.0;
    double acc = 0.0;
    for (int base = bd.first; base < last; base += 64) {
        const int t = base + lane;
        if (t < last) {
#pragma unroll
            for (int q = 0; q < 10; ++q) stage[lane * 10 + q] = terms[(size_t)t * 10 + q];
        }
        __syncthreads();
        const int m = min(64, last - base);
        if (lane < 9) {
            for (int e = 0; e < m; ++e) acc = acc + stage[e * 10 + lane];
        } else if (lane == 9) {
            for (int e = 0; e < m; ++e) acc = acc + stage[e * 10 + 9] - p0;
        } else if (lane == 10) {
            for (int e = 0; e < m; ++e) acc = acc + (stage[e * 10 + 9] - p0) * (stage[e * 10 + 9] - p0);
        }
        __syncthreads();
    }
    double *out = sums + (size_t)b * 12;
    if (lane < 9) out[lane] = acc;
    if (lane == 9) {
        out[9] = p0;
        out[10] = acc;
    }
    if (lane == 10) out[11] = acc;
}

// Primitive variables of level TO at a list of nodes (the probes).
__global__ __launch_bounds__(256) void gather_primitive(const double *U, const int *nodes, int count,
                                                        double gamma, double gasR, double *prim)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    double Un[5], Uo[6];
    LoadNode(U, nodes[t], Un);
    Primitive(gamma, gasR, Un, Uo);
#pragma unroll
    for (int c = 0; c < 6; ++c) prim[(size_t)t * 6 + c] = Uo[c];
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
        LoadNode(U, idx, Un);
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

// ------------------------------------------------ batched cases: helpers
// Per-case data of the batched time step.
struct BatchCase {
    const double *U;                // level TO
    const uint8_t *fluid;
    Grid g;
    double d[3];                    // cell sizes
    double cfl;
    double gamma, gasR;
};

struct CflChunk {
    int c;
    int pad;
    long start, count;              // interior nodes [start, start + count) of case c
};

// Per-axis maximum characteristic speed of every case (solve.c:108-153),
// one chunk of one case per block iteration.
__global__ __launch_bounds__(256) void cfl_batch(const CflChunk *chunks, int nchunks, const BatchCase *cases,
                                                 unsigned long long *vmax)
{
    __shared__ double red[3][256];
    for (long w = blockIdx.x; w < nchunks; w += gridDim.x) {
        const CflChunk ch = chunks[w];
        const BatchCase &bc = cases[ch.c];
        const Grid &g = bc.g;
        const int ex = g.hi[0] - g.lo[0], ey = g.hi[1] - g.lo[1];
        double m[3] = {0.0, 0.0, 0.0};
        for (long t = ch.start + threadIdx.x; t < ch.start + ch.count; t += blockDim.x) {
            const int i = g.lo[0] + (int)(t % ex);
            const int j = g.lo[1] + (int)((t / ex) % ey);
            const int k = g.lo[2] + (int)(t / ((long)ex * ey));
            const int idx = Index(g, i, j, k);
            if (!bc.fluid[idx]) continue;
            double Un[5], Uo[6];
            LoadNode(bc.U, idx, Un);
            Primitive(bc.gamma, bc.gasR, Un, Uo);
            const double c = sqrt(bc.gamma * bc.gasR * Uo[5]);
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const double V = fabs(Uo[s + 1]) + c;
                if (m[s] < V) m[s] = V;
            }
        }
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
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                atomicMax(vmax + 3 * ch.c + s, (unsigned long long)__double_as_longlong(red[s][0]));
            }
        }
        __syncthreads();
    }
}

constexpr int kClock = 6;           // doubles per batch case clock: now, end, rcInt, tmInt, stepC, stepN

// One step of the solver loop's clock (solve.c:59-78): dt = CFL min_s(d_s /
// vmax_s), rectified to the output instants and the end time.  ck = {now,
// end, rcInt, tmInt, stepC, stepN}.  A clock that has reached its end time
// or its step limit stays as it is and gives dt = 0; the step count only
// advances with the clock, as in the solver loop.
__device__ __forceinline__ double AdvanceClock(const double d[3], double cfl, const unsigned long long *vmax, double *ck)
{
    if (!(ck[0] < ck[1] && ck[4] < ck[5])) return 0.0;
    double v[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) v[s] = __longlong_as_double((long long)vmax[s]);
    const double tx = d[0] / v[0], ty = d[1] / v[1], tz = d[2] / v[2];
    const double tyz = (ty < tz) ? ty : tz;
    double h = cfl * ((tx < tyz) ? tx : tyz);
    double now = ck[0], rc = ck[2];
    const double end = ck[1], tm = ck[3];
    if (rc + h > tm) {
        h = tm - rc;
        rc = 0.0;
    } else {
        rc = rc + h;
    }
    now = now + h;
    if (now > end) {
        h = end - (now - h);
        now = end;
    }
    ck[0] = now;
    ck[2] = rc;
    ck[4] = ck[4] + 1.0;
    return h;
}

// Time step of every case of a batch; half[c] receives dt / 2, the step of
// each Strang-split directional operator.
__global__ void dt_batch(const BatchCase *cases, int C, const unsigned long long *vmax, double *clock, double *dt,
                         double *half)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const BatchCase &bc = cases[c];
    const double h = AdvanceClock(bc.d, bc.cfl, vmax + 3 * c, clock + kClock * c);
    dt[c] = h;
    half[c] = 0.5 * h;
}

// Cell sizes and CFL number of a single-case device march.
struct StepGeometry {
    double d[3];
    double cfl;
};

// Time step of a single case on the device: step[0] = dt, step[1] = dt / 2
// (the operators' step), hist[step[2]] = dt (the march's record of its
// steps, step[2] counting the steps of the current march call).
// The clock carries the output cadence of the solver loop (solve.c:57-62,
// 91-104): clock[7 + n] accumulates the steps of data kind n, clock[12 + n]
// is its interval; after a step at which some output is due (its
// accumulated time reaches the interval, or the loop ends) clock[6] is set
// and later steps leave the clock as it is, so that the host writes the
// output of exactly that state.
__global__ void dt_single(StepGeometry sg, const unsigned long long *vmax, double *clock, double *step, double *hist)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double h = 0.0;
    if (clock[6] == 0.0) {
        h = AdvanceClock(sg.d, sg.cfl, vmax, clock);
        if (h > 0.0) {
            bool due = false;
            for (int n = 0; n < 5; ++n) {
                clock[7 + n] = clock[7 + n] + h;
                if (clock[7 + n] >= clock[12 + n] || clock[0] == clock[1] || clock[4] == clock[5]) due = true;
            }
            if (due) clock[6] = 1.0;
        }
    }
    step[0] = h;
    step[1] = 0.5 * h;
    hist[(int)step[2]] = h;
    step[2] = step[2] + 1.0;
}

struct GatherCase {
    double *U;                      // level TO
    double gamma;
    double given[7][6];
};

// The resolved boundary fill of every case (bc_gather for many cases).
__global__ __launch_bounds__(256) void gather_batch(const GatherCase *cases, const int *__restrict__ cid,
                                                    const int *__restrict__ dst, const int *__restrict__ src,
                                                    long count)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const GatherCase &gc = cases[cid[t]];
    const int d = dst[t], q = src[t];
    double out[5];
    if (q >= 0) {
        LoadNode(gc.U, q, out);
    } else {
        Conservative(gc.gamma, gc.given[-3 - q], out);
    }
    StoreNode(gc.U, d, out);
}

__global__ void fill_i32(int *p, long n, int v)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// Plain streaming copy: the reference point for HBM counters and bandwidth.
// Field exchange between the component-major host layout and the
// node-major device layout (one thread per node).
__global__ __launch_bounds__(256) void to_node_major(const double *__restrict__ soa, double *__restrict__ aos, long N)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
#pragma unroll
    for (int c = 0; c < 5; ++c) aos[t * 5 + c] = soa[c * N + t];
}

__global__ __launch_bounds__(256) void to_component_major(const double *__restrict__ aos, double *__restrict__ soa,
                                                          long N)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
#pragma unroll
    for (int c = 0; c < 5; ++c) soa[c * N + t] = aos[t * 5 + c];
}

__global__ __launch_bounds__(256) void stream_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) dst[t] = src[t];
}

__global__ void fill_u64(unsigned long long *p, int n, unsigned long long v)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// The three per-axis speed maxima set to their floors (the body speeds).
__global__ void init_speeds(unsigned long long *vmax, unsigned long long x, unsigned long long y, unsigned long long z)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    vmax[0] = x;
    vmax[1] = y;
    vmax[2] = z;
}

}  // namespace

// ------------------------------------------------------------- device state
struct ArtraDevice {
    ArtraDevParams par;
    Grid g;
    double *U = nullptr;            // [3][N][5]: three levels, node-major
    double *exchange = nullptr;     // [5][N]: host-layout staging of one level
    uint8_t *fluid = nullptr;
    uint8_t *flagStage = nullptr;   // packed flags of a changed box (moving bodies)
    long flagStageCap = 0;
    unsigned long long *vmax = nullptr;
    double *marchBuf = nullptr;     // device march: speeds, clock, step, step record (kMarchChunk)
    // surface force and probes
    ArtraForceNode *forceNodes = nullptr;
    ArtraForceBody *forceBodies = nullptr;
    double *forceTerms = nullptr;   // [forceCount][10]
    double *forceSums = nullptr;    // [forceBodyCount][12]
    int forceCount = 0, forceBodyCount = 0;
    int forceNodeCap = 0, forceBodyCap = 0;     // allocated entries, kept across rebuilds
    int *forceCounters = nullptr;   // device-built force table: running end, per body counts
    int *forceCountsHost = nullptr; // their pinned host copy
    int forceCounterCap = 0;
    int *probeNodes = nullptr;      // gather buffers, grown as needed
    double *probePrim = nullptr;
    int probeCap = 0;
    const double *stepDev = nullptr; // when set, the operators read their step from here (device march)
    // the device march's step as a captured graph
    int useGraph = 1;
    int configGen = 0;              // bumped by every change to what a step launches
    int graphGen = -1;
    double graphKey[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    hipGraphExec_t stepGraph = nullptr;
    int ordinal = 0;                // the GPU this solver lives on
    hipStream_t stream = nullptr;
    std::vector<BcGroup> groups[4];          // mode 0: full fill; 1 + s: faces of axis s only
    BcArgs bc{};
    int gatherResolved = 0;
    int axisGather[3] = {0, 0, 0};  // the fill between the stages of a sweep along s is a gather (walls elsewhere)
    int useGather = 1;
    int *gDst[4] = {nullptr, nullptr, nullptr, nullptr};
    int *gSrc[4] = {nullptr, nullptr, nullptr, nullptr};
    long gCount[4] = {0, 0, 0, 0};
    IbTable ib;                     // ghost nodes, grouped by layer
    int ibImageOnly = 0;            // every ghost reconstructs from its image point (no layer reads another)
    IbTable joins;                  // nodes a moving body has just uncovered
    // the node classification of moving bodies (did, fid, lid, gst), kept
    // on the device and reclassified there after every move
    int *geoFlags = nullptr;        // [4][N]
    int *geoPacked = nullptr;       // [5][region nodes] for the host's copy
    int *geoPackedHost = nullptr;   // that copy (pinned)
    long geoPackedCap = 0;
    int *geoMoving = nullptr;       // [bodies]: 1 for a body that moves
    int *geoCounters = nullptr;     // ghost-table build: entries taken, overflow, each layer's end
    int *geoCountsHost = nullptr;   // their pinned host copy
    GeoSetup geo{};
    std::vector<GeoBodyDev> geoBodies;
    std::vector<int> layerStart;
    FluxModel mdl{};
    int segment = 32;
    int exact = 0;                  // 1: reference operation order (bit-faithful), 0: fast arithmetic
    int partialFill = 1;            // axis-only fills between the stages of inviscid split sweeps
    int hybrid = 0;                 // learned-indicator routing of WENO5 faces
    float *mlpParams = nullptr;     // FP32 biases/output vector (device)
    uint16_t *mlpW1 = nullptr, *mlpW2 = nullptr;
    float mlpThreshold = 0.0f;
    float smoothVariation = 1.0e-3f; // relative stencil variation below which the network is skipped
    float mlpB3 = 0.0f;
    unsigned long long *hybCounts = nullptr;
    // lean pass of the hybrid (rk_lean, then rk_line_list over the deferred lines)
    int *deferList = nullptr;       // lines the lean pass deferred (capacity: lines of the longest axis list)
    int *deferCount = nullptr;      // [1] their number in the current launch
    unsigned long long *deferTotal = nullptr;   // [3] deferred lines per axis since the last policy update
    int leanFirst[3] = {1, 1, 1};   // per axis: run the lean pass first (1) or the hybrid kernel on every line (0)
    int leanRetry[3] = {0, 0, 0};   // policy updates until an axis running the hybrid kernel directly tries again
    long leanLines[3] = {0, 0, 0};  // lines sent through the lean pass since the last policy update
    float *mlpBias = nullptr;       // b1, b2, w3 contiguous (96 FP32)
    // fused directional sweeps (rk_line)
    int fused = 1;                  // allowed by the caller
    int allFluid = 0;               // every interior node is fluid
    struct LineRules {
        int ok = 0;
        int count = 0;
        int dst[kLineRules], src[kLineRules];
        double given[2][6];
    } lines[3];
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
        sa.step = dev->stepDev;
        sa.dds = dev->par.dd[s];
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

// Entries [b, e) of a table, one block each.
int LaunchIb(ArtraDevice *dev, double *U, const IbTable &tb, int b, int e)
{
    int slot;
    if (Begin(dev, ARTRA_T_IB, &slot)) return 1;
    hipLaunchKernelGGL(ib_fill, dim3((unsigned)(e - b)), dim3(64), 0, dev->stream, U, tb.ghosts, b, tb.snode, tb.sw,
                       dev->par.gamma, dev->par.gasR);
    HIP_TRY(hipGetLastError());
    return End(dev, slot);
}

// The immersed-boundary ghost fill of one level, layer by layer
// (immersed_boundary.c:270-340).  When every ghost reconstructs from its
// image point (its stencil holds fluid nodes only, which the fill never
// writes), no layer reads another and one launch fills them all.
int ImmersedFill(ArtraDevice *dev, int level)
{
    double *U = Level(dev, level);
    const int layers = (int)dev->layerStart.size() - 1;
    if (dev->ibImageOnly) {
        const int b = dev->layerStart[0], e = dev->layerStart[layers];
        return (e > b) ? LaunchIb(dev, U, dev->ib, b, e) : 0;
    }
    for (int r = 0; r < layers; ++r) {
        const int b = dev->layerStart[r], e = dev->layerStart[r + 1];
        if (e > b && LaunchIb(dev, U, dev->ib, b, e)) return 1;
    }
    return 0;
}

// Replace a table's contents with a host table.  Kernels already enqueued
// may still read the old contents, so the stream drains first.
int UploadIbTable(ArtraDevice *dev, IbTable &tb, const ArtraGhostTable *t)
{
    HIP_TRY(hipStreamSynchronize(dev->stream));
    tb.count = 0;
    if (t->nghost > tb.ghostCap) {
        const int cap = t->nghost + t->nghost / 4 + 64;
        hipFree(tb.ghosts);
        tb.ghosts = nullptr;
        tb.ghostCap = 0;
        HIP_TRY(hipMalloc(&tb.ghosts, (size_t)cap * sizeof(GhostDev)));
        tb.ghostCap = cap;
    }
    if (t->nstencil > tb.stencilCap) {
        const int cap = t->nstencil + t->nstencil / 4 + 256;
        hipFree(tb.snode);
        hipFree(tb.sw);
        tb.snode = nullptr;
        tb.sw = nullptr;
        tb.stencilCap = 0;
        HIP_TRY(hipMalloc(&tb.snode, (size_t)cap * sizeof(int)));
        HIP_TRY(hipMalloc(&tb.sw, (size_t)cap * sizeof(double)));
        tb.stencilCap = cap;
    }
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
    HIP_TRY(hipMemcpy(tb.ghosts, host.data(), host.size() * sizeof(GhostDev), hipMemcpyHostToDevice));
    if (t->nstencil > 0) {
        HIP_TRY(hipMemcpy(tb.snode, t->stencil_node, (size_t)t->nstencil * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(tb.sw, t->stencil_w, (size_t)t->nstencil * sizeof(double), hipMemcpyHostToDevice));
    }
    tb.count = t->nghost;
    return 0;
}

void FreeIbTable(IbTable &tb)
{
    hipFree(tb.ghosts);
    hipFree(tb.snode);
    hipFree(tb.sw);
    tb = IbTable();
}

// mode 0: the whole external boundary; 1 + s: only the ghosts a sweep along
// axis s reads (valid between the stages of an inviscid split sweep).

int TreatBoundary(ArtraDevice *dev, int level, int mode = 0)
{
    double *U = Level(dev, level);
    // immersed boundary first
    if (ImmersedFill(dev, level)) return 1;
    if ((dev->gatherResolved || (mode > 0 && dev->axisGather[mode - 1])) && dev->useGather) {
        const long count = dev->gCount[mode];
        if (count <= 0) return 0;
        GatherArgs ga;
        ga.U = U;
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

// Launch shape of rk_line for axis p: threads, interior nodes per thread, LDS bytes.
struct LineShape {
    int T = 0, K = 0;
    size_t lds = 0;
    long lines = 0;
};

LineShape LineShapeOf(const ArtraDevice *dev, int p, bool hyb)
{
    const Grid &g = dev->g;
    const int ext[3] = {g.nx, g.ny, g.nz};
    const int t1 = (p == 0) ? 1 : 0, t2 = (p == 2) ? 1 : 2;
    LineShape sh;
    const int interior = g.hi[p] - g.lo[p];
    if (interior <= 0) return sh;
    sh.K = (interior <= 256) ? 1 : 2;
    const int per = (interior + sh.K - 1) / sh.K;
    sh.T = ((per + 63) / 64) * 64;
    if (sh.T > 256) return LineShape{};
    const int F = interior + 1;
    sh.lds = LineLdsBytes(ext[p], F, hyb);
    sh.lines = (long)(g.hi[t1] - g.lo[t1]) * (g.hi[t2] - g.lo[t2]);
    return sh;
}

// Whether one directional operator along p can run as one fused rk_line launch.
bool LineEligible(const ArtraDevice *dev, int p, bool hyb)
{
    if (!dev->fused || !dev->lines[p].ok || !dev->allFluid) return false;
    if (dev->layerStart.back() != 0) return false;               // immersed-boundary ghosts present
    if (dev->mdl.refMu > 0.0 || dev->par.multidim != 0) return false;
    if (!dev->gatherResolved || !dev->useGather || !dev->partialFill) return false;
    const LineShape sh = LineShapeOf(dev, p, hyb);
    return sh.T > 0 && sh.lines > 0 && sh.lds <= 64 * 1024;
}

template <int S, int NS, bool FAST, bool HYB>
void LaunchLineK(ArtraDevice *dev, const LineShape &sh, const LineArgs &la, const LineHybrid &hy)
{
    if (sh.K == 1) {
        hipLaunchKernelGGL((rk_line<S, NS, FAST, HYB, 1>), dim3((unsigned)sh.lines), dim3(sh.T), sh.lds, dev->stream,
                           la, dev->mdl, hy);
    } else {
        hipLaunchKernelGGL((rk_line<S, NS, FAST, HYB, 2>), dim3((unsigned)sh.lines), dim3(sh.T), sh.lds, dev->stream,
                           la, dev->mdl, hy);
    }
}

template <int S>
void LaunchLineS(ArtraDevice *dev, const LineShape &sh, const LineArgs &la, const LineHybrid &hy)
{
    const bool fast = !dev->exact;
    if (dev->par.sScheme == 1) {
        if (dev->hybrid) LaunchLineK<S, 6, true, true>(dev, sh, la, hy);
        else if (fast) LaunchLineK<S, 6, true, false>(dev, sh, la, hy);
        else LaunchLineK<S, 6, false, false>(dev, sh, la, hy);
    } else {
        if (fast) LaunchLineK<S, 4, true, false>(dev, sh, la, hy);
        else LaunchLineK<S, 4, false, false>(dev, sh, la, hy);
    }
}

// The hybrid's lean pass over every line of axis S, then the full hybrid
// body over the lines it deferred (the list's length stays on the device).
template <int S>
void LaunchLean(ArtraDevice *dev, const LineShape &sh, const LineArgs &la, const LineHybrid &hy, size_t leanLds,
                bool mean)
{
    const unsigned listGrid = (unsigned)std::min<long>(sh.lines, kListBlocks);
    if (mean) {
        hipLaunchKernelGGL((rk_lean<S, true>), dim3((unsigned)sh.lines), dim3(sh.T), leanLds, dev->stream, la, dev->mdl,
                           hy, dev->deferList, dev->deferCount, dev->deferTotal + S);
    } else {
        hipLaunchKernelGGL((rk_lean<S, false>), dim3((unsigned)sh.lines), dim3(sh.T), leanLds, dev->stream, la,
                           dev->mdl, hy, dev->deferList, dev->deferCount, dev->deferTotal + S);
    }
    hipLaunchKernelGGL(rk_line_list<S>, dim3(listGrid), dim3(sh.T), sh.lds, dev->stream, la, dev->mdl, hy,
                       dev->deferList, dev->deferCount);
}

// All RK stages of the directional operator along p in one launch; the
// result lands in TO, whose full boundary fill follows.
int LineSweep(ArtraDevice *dev, double dt, int p)
{
    const Grid &g = dev->g;
    const LineShape sh = LineShapeOf(dev, p, dev->hybrid != 0);
    const int ext[3] = {g.nx, g.ny, g.nz};
    const int stride[3] = {1, g.nx, g.nx * g.ny};
    const int t1 = (p == 0) ? 1 : 0, t2 = (p == 2) ? 1 : 2;
    LineArgs la{};
    la.U = Level(dev, 0);
    la.n = ext[p];
    la.lo = g.lo[p];
    la.hi = g.hi[p];
    la.n1 = g.hi[t1] - g.lo[t1];
    la.lo1 = g.lo[t1];
    la.lo2 = g.lo[t2];
    la.stride = stride[p];
    la.stride1 = stride[t1];
    la.stride2 = stride[t2];
    if (dev->par.tScheme == 0) {
        la.stages = 2;
        la.a[0] = 0.0; la.b[0] = 1.0;
        la.a[1] = 1.0 / 2.0; la.b[1] = 1.0 / 2.0;
    } else {
        la.stages = 3;
        la.a[0] = 0.0; la.b[0] = 1.0;
        la.a[1] = 3.0 / 4.0; la.b[1] = 1.0 / 4.0;
        la.a[2] = 1.0 / 3.0; la.b[2] = 2.0 / 3.0;
    }
    la.r = dt * dev->par.dd[p];
    la.dt = dev->stepDev;
    la.dd = dev->par.dd[p];
    la.gamma = dev->par.gamma;
    const ArtraDevice::LineRules &R = dev->lines[p];
    la.nrules = R.count;
    for (int e = 0; e < R.count; ++e) {
        la.rdst[e] = R.dst[e];
        la.rsrc[e] = R.src[e];
    }
    memcpy(la.given, R.given, sizeof la.given);
    LineHybrid hy{};
    if (dev->hybrid) {
        hy.W1 = dev->mlpW1;
        hy.W2 = dev->mlpW2;
        hy.bias = dev->mlpBias;
        hy.b3 = dev->mlpB3;
        hy.threshold = dev->mlpThreshold;
        hy.counts = dev->hybCounts;
        hy.smoothVariation = dev->smoothVariation;
    }
    int slot;
    if (Begin(dev, ARTRA_T_SWEEP, &slot)) return 1;
    const size_t leanLds = LeanLdsBytes(la.n, la.hi - la.lo + 1);
    if (dev->hybrid && sh.K == 1 && dev->leanFirst[p] && leanLds <= 64 * 1024) {
        // lean pass over every line, then the full hybrid kernel over the
        // lines it deferred (the list's length stays on the device)
        hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, dev->stream, dev->deferCount, 1L, 0);
        const bool mean = !dev->mdl.roe && !dev->mdl.sw;
        if (p == 0) LaunchLean<0>(dev, sh, la, hy, leanLds, mean);
        else if (p == 1) LaunchLean<1>(dev, sh, la, hy, leanLds, mean);
        else LaunchLean<2>(dev, sh, la, hy, leanLds, mean);
        dev->leanLines[p] += sh.lines;
    } else if (p == 0) {
        LaunchLineS<0>(dev, sh, la, hy);
    } else if (p == 1) {
        LaunchLineS<1>(dev, sh, la, hy);
    } else {
        LaunchLineS<2>(dev, sh, la, hy);
    }
    HIP_TRY(hipGetLastError());
    return End(dev, slot);
}

// Per-line copy rules of the between-stage fill along axis p, read off the
// resolved gather of the faces normal to p: every interior line must see the
// same rules (destination position along the line <- source position on the
// same line, or a constant state), covering every ghost the stencil reads.
void DeriveLineRules(ArtraDevice *dev, const ArtraBcGather *gt)
{
    const Grid &g = dev->g;
    const int ext[3] = {g.nx, g.ny, g.nz};
    const int half = (dev->par.sScheme == 1) ? 3 : 2;
    for (int p = 0; p < 3; ++p) {
        ArtraDevice::LineRules &R = dev->lines[p];
        R = ArtraDevice::LineRules{};
        if (!gt->resolved || g.hi[p] - g.lo[p] <= 0) continue;
        const int t1 = (p == 0) ? 1 : 0, t2 = (p == 2) ? 1 : 2;
        const long n1 = g.hi[t1] - g.lo[t1], n2 = g.hi[t2] - g.lo[t2];
        if (n1 <= 0 || n2 <= 0) continue;
        std::vector<std::vector<std::pair<int, int>>> per((size_t)(n1 * n2));
        int faces[2] = {-1, -1}, nfaces = 0;
        bool ok = true;
        for (long e = 0; e < gt->axis_count[p] && ok; ++e) {
            const int d = gt->axis_dst[p][e], q = gt->axis_src[p][e];
            const int c[3] = {d % g.nx, (d / g.nx) % g.ny, d / (g.nx * g.ny)};
            if (c[t1] < g.lo[t1] || c[t1] >= g.hi[t1] || c[t2] < g.lo[t2] || c[t2] >= g.hi[t2]) continue;
            int src;
            if (q >= 0) {
                const int cq[3] = {q % g.nx, (q / g.nx) % g.ny, q / (g.nx * g.ny)};
                if (cq[t1] != c[t1] || cq[t2] != c[t2] || cq[p] < g.lo[p] || cq[p] >= g.hi[p]) ok = false;
                src = cq[p];
            } else {
                const int face = -3 - q;
                int slot = -1;
                for (int k = 0; k < nfaces; ++k)
                    if (faces[k] == face) slot = k;
                if (slot < 0) {
                    if (nfaces == 2) {
                        ok = false;
                        break;
                    }
                    slot = nfaces;
                    faces[nfaces++] = face;
                }
                src = -1 - slot;
            }
            per[(size_t)((c[t1] - g.lo[t1]) + n1 * (c[t2] - g.lo[t2]))].emplace_back(c[p], src);
        }
        if (!ok) continue;
        for (auto &v : per) std::sort(v.begin(), v.end());
        const auto &ref = per[0];
        for (const auto &v : per) {
            if (v != ref) {
                ok = false;
                break;
            }
        }
        if (!ok || (int)ref.size() > kLineRules) continue;
        // copy rules read interior nodes, each feeding at most kRefillPerNode
        // ghosts (the owner of a node refills them when it updates it)
        for (const auto &r : ref) {
            if (r.second < 0) continue;
            int feeds = 0;
            for (const auto &o : ref) feeds += (o.second == r.second);
            ok = ok && r.second >= g.lo[p] && r.second < g.hi[p] && feeds <= kRefillPerNode;
        }
        if (!ok) continue;
        for (int pos = g.lo[p] - half; pos < g.hi[p] + half; ++pos) {
            if (pos >= g.lo[p] && pos < g.hi[p]) continue;
            bool hit = pos >= 0 && pos < ext[p];
            if (hit) {
                hit = false;
                for (const auto &r : ref) hit = hit || r.first == pos;
            }
            ok = ok && hit;
        }
        if (!ok) continue;
        R.count = (int)ref.size();
        for (int e = 0; e < R.count; ++e) {
            R.dst[e] = ref[e].first;
            R.src[e] = ref[e].second;
        }
        for (int k = 0; k < nfaces; ++k)
            for (int v = 0; v < 6; ++v) R.given[k][v] = dev->par.varBC[faces[k]][v];
        R.ok = 1;
    }
}

// Runge-Kutta stages of one directional operator (fluid_dynamics.c:121-148).
// Between the stages of an inviscid sweep along one axis only that axis's
// ghosts are read, so only they are filled.  After the last stage the fill
// covers the ghosts the next operator of the step reads (next = 1 + its
// axis), or everything at the end of the step (next = 0), exactly as
// TreatBoundary after every stage leaves them: each fill's values derive
// from interior nodes and from nodes written earlier in the same fill.
int Stages(ArtraDevice *dev, double dt, int p, int next)
{
    const int lazy = (p < 3 && !(dev->mdl.refMu > 0.0) && dev->partialFill);
    const int mid = lazy ? 1 + p : 0;
    const int last = lazy ? next : 0;
    if (p < 3 && LineEligible(dev, p, dev->hybrid != 0)) return LineSweep(dev, dt, p) || TreatBoundary(dev, 0, last);
    if (dev->hybrid) return Fail("the learned indicator is on but axis %d cannot run fused", p);
    if (dev->par.tScheme == 0) {
        if (Sweep(dev, dt, 0.0, 1.0, 0, 0, 1, p) || TreatBoundary(dev, 1, mid)) return 1;
        if (Sweep(dev, dt, 1.0 / 2.0, 1.0 / 2.0, 0, 1, 0, p) || TreatBoundary(dev, 0, last)) return 1;
        return 0;
    }
    if (Sweep(dev, dt, 0.0, 1.0, 0, 0, 1, p) || TreatBoundary(dev, 1, mid)) return 1;
    if (Sweep(dev, dt, 3.0 / 4.0, 1.0 / 4.0, 0, 1, 2, p) || TreatBoundary(dev, 2, mid)) return 1;
    if (Sweep(dev, dt, 1.0 / 3.0, 2.0 / 3.0, 0, 2, 0, p) || TreatBoundary(dev, 0, last)) return 1;
    return 0;
}

// One stage of the gravity source operator on levels (to, tn) -> tm.
int SourceStage(ArtraDevice *dev, double dt, double a, double b, int to, int tn, int tm)
{
    const Grid &g = dev->g;
    const long total = (long)(g.hi[0] - g.lo[0]) * (g.hi[1] - g.lo[1]) * (g.hi[2] - g.lo[2]);
    if (total <= 0) return 0;
    StageArgs sa{};
    sa.Utn = Level(dev, tn);
    sa.Uto = Level(dev, to);
    sa.Utm = Level(dev, tm);
    sa.fluid = dev->fluid;
    sa.a = a;
    sa.b = b;
    sa.r = dt;
    sa.toIsTn = (to == tn);
    sa.step = dev->stepDev;
    sa.dds = 1.0;
    int slot;
    if (Begin(dev, ARTRA_T_SWEEP, &slot)) return 1;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (dev->exact) {
        hipLaunchKernelGGL(source_stage<false>, grid, dim3(256), 0, dev->stream, g, sa, dev->par.g[0], dev->par.g[1],
                           dev->par.g[2]);
    } else {
        hipLaunchKernelGGL(source_stage<true>, grid, dim3(256), 0, dev->stream, g, sa, dev->par.g[0], dev->par.g[1],
                           dev->par.g[2]);
    }
    HIP_TRY(hipGetLastError());
    return End(dev, slot);
}

// Runge-Kutta stages of the gravity source operator (fluid_dynamics.c:121-148
// with s = PHI), each followed by the full boundary fill as in the reference.
int SourceStages(ArtraDevice *dev, double dt)
{
    if (dev->par.tScheme == 0) {
        if (SourceStage(dev, dt, 0.0, 1.0, 0, 0, 1) || TreatBoundary(dev, 1)) return 1;
        return SourceStage(dev, dt, 1.0 / 2.0, 1.0 / 2.0, 0, 1, 0) || TreatBoundary(dev, 0);
    }
    if (SourceStage(dev, dt, 0.0, 1.0, 0, 0, 1) || TreatBoundary(dev, 1)) return 1;
    if (SourceStage(dev, dt, 3.0 / 4.0, 1.0 / 4.0, 0, 1, 2) || TreatBoundary(dev, 2)) return 1;
    return SourceStage(dev, dt, 1.0 / 3.0, 2.0 / 3.0, 0, 2, 0) || TreatBoundary(dev, 0);
}

// The lean pass's per-axis policy, from the lines it deferred since the
// last update (the stream must be idle): an axis where most lines have
// troubled faces (a shock across every line) runs the hybrid kernel on
// every line, and tries the lean pass again a few updates later.
int UpdateLeanPolicy(ArtraDevice *dev)
{
    if (!dev->deferTotal) return 0;
    unsigned long long deferred[4];
    HIP_TRY(hipMemcpy(deferred, dev->deferTotal, sizeof deferred, hipMemcpyDeviceToHost));
    for (int q = 0; q < 3; ++q) {
        if (dev->leanLines[q] > 0) {
            dev->leanFirst[q] = 2 * (long)deferred[q] < dev->leanLines[q];
            if (!dev->leanFirst[q]) dev->leanRetry[q] = 8;
        } else if (!dev->leanFirst[q] && --dev->leanRetry[q] <= 0) {
            dev->leanFirst[q] = 1;
        }
        dev->leanLines[q] = 0;
    }
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, dev->stream, dev->deferTotal, 4, 0ull);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Per-axis speed maxima of level TO folded with the floors (body speeds),
// into vmax on the device (solve.c:108-153).
int EnqueueSpeeds(ArtraDevice *dev, const double floor[3], unsigned long long *vmax)
{
    unsigned long long init[3];
    for (int s = 0; s < 3; ++s) {
        const double f = (floor[s] > 0.0) ? floor[s] : 0.0;
        memcpy(init + s, &f, sizeof f);
    }
    int slot;
    if (Begin(dev, ARTRA_T_CFL, &slot)) return 1;
    hipLaunchKernelGGL(init_speeds, dim3(1), dim3(1), 0, dev->stream, vmax, init[0], init[1], init[2]);
    const Grid &g = dev->g;
    const long total = (long)(g.hi[0] - g.lo[0]) * (g.hi[1] - g.lo[1]) * (g.hi[2] - g.lo[2]);
    long blocks = (total + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(cfl_max, dim3((unsigned)blocks), dim3(256), 0, dev->stream, Level(dev, 0), dev->fluid, g,
                       dev->par.gamma, dev->par.gasR, vmax);
    HIP_TRY(hipGetLastError());
    return End(dev, slot);
}

// Device memory of a device-stepped march: the speed maxima (3 x u64, one
// spare), the clock (kMarchClock), the step (dt, dt / 2) and the record of
// up to cap steps, all 8-byte words.
struct MarchScratch {
    unsigned long long *vmax;
    double *clock;
    double *step;
    double *hist;
    int cap;
};

MarchScratch MarchScratchAt(void *base, int cap)
{
    MarchScratch m;
    m.vmax = (unsigned long long *)base;
    m.clock = (double *)base + 4;
    m.step = m.clock + kMarchClock;
    m.hist = m.step + 3;
    m.cap = cap;
    return m;
}

// One time step with its dt computed on the device: the CFL maxima, the
// clock's step (dt_single) and EvolveFluidDynamics reading its step from
// device memory.
int EnqueueStep(ArtraDevice *dev, const MarchScratch &ms, const double floor[3], double cfl)
{
    StepGeometry sg;
    for (int s = 0; s < 3; ++s) sg.d[s] = dev->par.d[s];
    sg.cfl = cfl;
    if (EnqueueSpeeds(dev, floor, ms.vmax)) return 1;
    hipLaunchKernelGGL(dt_single, dim3(1), dim3(1), 0, dev->stream, sg, ms.vmax, ms.clock, ms.step, ms.hist);
    if (hipGetLastError() != hipSuccess) return Fail("dt_single launch failed");
    return artra_dev_evolve(dev, 0.0);
}

// The step as a captured graph, replayed once per step: one launch instead
// of the step's 20 to 150 kernel launches.  Rebuilt when anything the
// step's launches depend on changed (configGen) or the body-speed floors or
// the CFL number differ; not used with kernel timing (its events) or the
// learned indicator (its lean-pass policy changes the launches between
// march calls).
int StepGraph(ArtraDevice *dev, const MarchScratch &ms, const double floor[3], double cfl, hipGraphExec_t *out)
{
    *out = nullptr;
    if (!dev->useGraph || dev->timing || dev->hybrid) return 0;
    const double key[5] = {floor[0], floor[1], floor[2], cfl, (double)(uintptr_t)ms.clock};
    if (dev->stepGraph && dev->graphGen == dev->configGen && 0 == memcmp(key, dev->graphKey, sizeof key)) {
        *out = dev->stepGraph;
        return 0;
    }
    if (dev->stepGraph) {
        HIP_TRY(hipGraphExecDestroy(dev->stepGraph));
        dev->stepGraph = nullptr;
    }
    hipGraph_t graph = nullptr;
    dev->stepDev = ms.step + 1;
    HIP_TRY(hipStreamBeginCapture(dev->stream, hipStreamCaptureModeThreadLocal));
    const int rc = EnqueueStep(dev, ms, floor, cfl);
    const hipError_t e = hipStreamEndCapture(dev->stream, &graph);
    dev->stepDev = nullptr;
    if (rc || e != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        return rc ? 1 : Fail("step capture failed: %s", hipGetErrorString(e));
    }
    const hipError_t ie = hipGraphInstantiate(&dev->stepGraph, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        dev->stepGraph = nullptr;
        return Fail("step graph instantiation failed: %s", hipGetErrorString(ie));
    }
    dev->graphGen = dev->configGen;
    memcpy(dev->graphKey, key, sizeof key);
    *out = dev->stepGraph;
    return 0;
}

// steps time steps with the dt computed on the device, with no host round
// trip in between.  clock (in/out) is the solver loop's clock, dts[i] the
// steps taken (0 once the clock has reached its end), *advanced the steps
// that moved the clock.  allowGraph: the step may be replayed from a
// captured graph (never for the allocation-free march: instantiation
// allocates).
int MarchSteps(ArtraDevice *dev, const MarchScratch &ms, double clock[kMarchClock], const double floor[3], double cfl,
               int steps, double *dts, int *advanced, bool allowGraph)
{
    *advanced = 0;
    if (steps <= 0) return 0;
    if (steps > ms.cap) return Fail("a device march takes at most %d steps per call (%d asked)", ms.cap, steps);
    const double step0 = clock[4];
    // the clock and the step words (dt, dt / 2, the record index from 0)
    // are contiguous
    double head[kMarchClock + 3];
    memcpy(head, clock, kMarchClock * sizeof(double));
    head[kMarchClock] = head[kMarchClock + 1] = head[kMarchClock + 2] = 0.0;
    HIP_TRY(hipMemcpyAsync(ms.clock, head, sizeof head, hipMemcpyHostToDevice, dev->stream));
    hipGraphExec_t graph = nullptr;
    if (allowGraph && StepGraph(dev, ms, floor, cfl, &graph)) return 1;
    int rc = 0;
    if (graph) {
        for (int i = 0; i < steps && !rc; ++i) {
            if (hipGraphLaunch(graph, dev->stream) != hipSuccess) rc = Fail("step graph launch failed");
        }
    } else {
        dev->stepDev = ms.step + 1;
        for (int i = 0; i < steps && !rc; ++i) rc = EnqueueStep(dev, ms, floor, cfl);
        dev->stepDev = nullptr;
    }
    if (rc) return 1;
    HIP_TRY(hipMemcpyAsync(clock, ms.clock, kMarchClock * sizeof(double), hipMemcpyDeviceToHost, dev->stream));
    HIP_TRY(hipMemcpyAsync(dts, ms.hist, (size_t)steps * sizeof(double), hipMemcpyDeviceToHost, dev->stream));
    HIP_TRY(hipStreamSynchronize(dev->stream));
    *advanced = (int)(clock[4] - step0);
    if (dev->hybrid && UpdateLeanPolicy(dev)) return 1;
    return Harvest(dev);
}

}  // namespace

extern "C" {

const char *artra_dev_error(void)
{
    return gError;
}

// Every interior node fluid (the fused sweeps' condition).
static int AllFluid(const Grid &g, const uint8_t *fluid)
{
    for (int k = g.lo[2]; k < g.hi[2]; ++k)
        for (int j = g.lo[1]; j < g.hi[1]; ++j)
            for (int i = g.lo[0]; i < g.hi[0]; ++i) {
                if (!fluid[((long)k * g.ny + j) * g.nx + i]) return 0;
            }
    return 1;
}

int artra_dev_create(const ArtraDevParams *par, const uint8_t *fluid, ArtraDevice **out)
{
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return Fail("no HIP device is visible");
    int ordinal = 0;
    HIP_TRY(hipGetDevice(&ordinal));
    if (ordinal < 0 || ordinal >= kMaxDevices) return Fail("device ordinal %d beyond the %d supported", ordinal, kMaxDevices);
    {
        // the address of this GPU's copy of the allocation-free march's
        // scratch, looked up once per device (the current one)
        std::lock_guard<std::mutex> hold(gScratchLock);
        if (!gScratch[ordinal]) {
            void *p = nullptr;
            HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            gScratch[ordinal] = p;
        }
    }
    ArtraDevice *dev = new ArtraDevice();
    dev->ordinal = ordinal;
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
    if (e == hipSuccess) e = hipMalloc(&dev->marchBuf, MarchScratchBytes(kMarchChunk));
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
    dev->allFluid = AllFluid(g, fluid);
    *out = dev;
    return 0;
}

void artra_dev_destroy(ArtraDevice *dev)
{
    if (!dev) return;
    if (dev->stream) hipStreamSynchronize(dev->stream);
    if (dev->stepGraph) hipGraphExecDestroy(dev->stepGraph);
    for (hipEvent_t ev : dev->evPool) hipEventDestroy(ev);
    hipFree(dev->U);
    hipFree(dev->exchange);
    hipFree(dev->fluid);
    hipFree(dev->flagStage);
    hipFree(dev->geoFlags);
    hipFree(dev->geoPacked);
    hipHostFree(dev->geoPackedHost);
    hipFree(dev->geoMoving);
    hipFree(dev->geoCounters);
    hipHostFree(dev->geoCountsHost);
    for (GeoBodyDev &b : dev->geoBodies) {
        hipFree(b.f); hipFree(b.e); hipFree(b.block); hipHostFree(b.stage); hipFree(b.sph); hipFree(b.bsph);
    }
    hipFree(dev->vmax);
    hipFree(dev->marchBuf);
    hipFree(dev->forceNodes);
    hipFree(dev->forceBodies);
    hipFree(dev->forceTerms);
    hipFree(dev->forceSums);
    hipFree(dev->forceCounters);
    hipHostFree(dev->forceCountsHost);
    hipFree(dev->probeNodes);
    hipFree(dev->probePrim);
    FreeIbTable(dev->ib);
    FreeIbTable(dev->joins);
    for (int m = 0; m < 4; ++m) {
        hipFree(dev->gDst[m]);
        hipFree(dev->gSrc[m]);
    }
    hipFree(dev->mlpParams);
    hipFree(dev->mlpW1);
    hipFree(dev->mlpW2);
    hipFree(dev->hybCounts);
    hipFree(dev->deferList);
    hipFree(dev->deferCount);
    hipFree(dev->deferTotal);
    hipFree(dev->mlpBias);
    if (dev->stream) hipStreamDestroy(dev->stream);
    delete dev;
}

int artra_dev_set_boundary(ArtraDevice *dev, int mode, const ArtraBcBox *boxes, int nbox, const int *group_start,
                           int ngroup)
{
    ++dev->configGen;
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
    ++dev->configGen;
    for (int m = 0; m < 4; ++m) {
        hipFree(dev->gDst[m]);
        hipFree(dev->gSrc[m]);
        dev->gDst[m] = dev->gSrc[m] = nullptr;
        dev->gCount[m] = 0;
    }
    dev->gatherResolved = 0;
    for (int s = 0; s < 3; ++s) dev->axisGather[s] = 0;
    DeriveLineRules(dev, gt);
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

int artra_dev_set_axis_gather(ArtraDevice *dev, int s, const ArtraBcGather *gt)
{
    ++dev->configGen;
    if (s < 0 || s > 2) return Fail("bad axis %d", s);
    if (dev->gatherResolved) return 0;              // the full gather's axis tables serve already
    const int m = 1 + s;
    hipFree(dev->gDst[m]);
    hipFree(dev->gSrc[m]);
    dev->gDst[m] = dev->gSrc[m] = nullptr;
    dev->gCount[m] = 0;
    dev->axisGather[s] = 0;
    if (!gt->resolved) return 0;
    const long n = gt->count;
    if (n > 0) {
        HIP_TRY(hipMalloc(&dev->gDst[m], n * sizeof(int)));
        HIP_TRY(hipMalloc(&dev->gSrc[m], n * sizeof(int)));
        HIP_TRY(hipMemcpy(dev->gDst[m], gt->dst, n * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(dev->gSrc[m], gt->src, n * sizeof(int), hipMemcpyHostToDevice));
    }
    dev->gCount[m] = n;
    dev->axisGather[s] = 1;
    return 0;
}

int artra_dev_set_fill_options(ArtraDevice *dev, int use_gather, int partial)
{
    ++dev->configGen;
    dev->useGather = (use_gather != 0);
    dev->partialFill = (partial != 0);
    return 0;
}

int artra_dev_set_ghosts(ArtraDevice *dev, const ArtraGhostTable *t)
{
    ++dev->configGen;
    dev->layerStart.assign(1, 0);
    dev->ibImageOnly = 0;
    if (UploadIbTable(dev, dev->ib, t)) return 1;
    if (t->nghost > 0) dev->layerStart.assign(t->layer_start, t->layer_start + t->layers + 1);
    int image = 1;
    for (int q = 0; q < t->nghost; ++q) image = image && (0 == t->ghosts[q].kind);
    dev->ibImageOnly = image;
    return 0;
}

int artra_dev_reconstruct(ArtraDevice *dev, int level, const ArtraGhostTable *joins)
{
    if (level < 0 || level > 2) return Fail("bad time level %d", level);
    if (joins->nghost == 0) return 0;
    if (UploadIbTable(dev, dev->joins, joins)) return 1;
    return LaunchIb(dev, Level(dev, level), dev->joins, 0, joins->nghost);
}

int artra_dev_set_fluid(ArtraDevice *dev, const uint8_t *fluid, const int (*box)[2])
{
    ++dev->configGen;
    const Grid &g = dev->g;
    HIP_TRY(hipStreamSynchronize(dev->stream));
    if (!box) {
        HIP_TRY(hipMemcpy(dev->fluid, fluid, (size_t)g.N, hipMemcpyHostToDevice));
    } else {
        const int ext[3] = {g.nx, g.ny, g.nz};
        int lo[3], hi[3];
        long count = 1;
        for (int s = 0; s < 3; ++s) {
            lo[s] = box[s][0] > 0 ? box[s][0] : 0;
            hi[s] = box[s][1] < ext[s] ? box[s][1] : ext[s];
            count *= (hi[s] > lo[s]) ? (hi[s] - lo[s]) : 0;
        }
        if (count > 0) {
            // the box packed on the host, one copy, scattered on the device
            std::vector<uint8_t> packed((size_t)count);
            long n = 0;
            for (int k = lo[2]; k < hi[2]; ++k)
                for (int j = lo[1]; j < hi[1]; ++j) {
                    memcpy(packed.data() + n, fluid + ((long)k * g.ny + j) * g.nx + lo[0], (size_t)(hi[0] - lo[0]));
                    n += hi[0] - lo[0];
                }
            if (count > dev->flagStageCap) {
                hipFree(dev->flagStage);
                dev->flagStage = nullptr;
                dev->flagStageCap = 0;
                HIP_TRY(hipMalloc(&dev->flagStage, (size_t)count));
                dev->flagStageCap = count;
            }
            HIP_TRY(hipMemcpy(dev->flagStage, packed.data(), (size_t)count, hipMemcpyHostToDevice));
            FlagBox fbx;
            for (int s = 0; s < 3; ++s) {
                fbx.lo[s] = lo[s];
                fbx.n[s] = hi[s] - lo[s];
            }
            hipLaunchKernelGGL(scatter_flags, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, dev->stream,
                               dev->fluid, dev->flagStage, g, fbx, count);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(dev->stream));
        }
    }
    dev->allFluid = AllFluid(g, fluid);
    return 0;
}

// The host hands fields over component-major ([5][N], the layout of the C
// interface); the device keeps them node-major.  The exchange goes through
// a device buffer of one level, allocated at the first transfer.
static int ExchangeBuffer(ArtraDevice *dev)
{
    if (!dev->exchange) HIP_TRY(hipMalloc(&dev->exchange, (size_t)5 * dev->g.N * sizeof(double)));
    return 0;
}

int artra_dev_upload(ArtraDevice *dev, int level, const double *soa)
{
    if (ExchangeBuffer(dev)) return 1;
    const long N = dev->g.N;
    HIP_TRY(hipMemcpyAsync(dev->exchange, soa, (size_t)5 * N * sizeof(double), hipMemcpyHostToDevice, dev->stream));
    hipLaunchKernelGGL(to_node_major, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, dev->stream, dev->exchange,
                       Level(dev, level), N);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(dev->stream));
    return 0;
}

int artra_dev_download(ArtraDevice *dev, int level, double *soa)
{
    if (ExchangeBuffer(dev)) return 1;
    const long N = dev->g.N;
    hipLaunchKernelGGL(to_component_major, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, dev->stream,
                       Level(dev, level), dev->exchange, N);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(soa, dev->exchange, (size_t)5 * N * sizeof(double), hipMemcpyDeviceToHost, dev->stream));
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

int artra_dev_treat_immersed(ArtraDevice *dev, int level)
{
    if (level < 0 || level > 2) return Fail("bad time level %d", level);
    return ImmersedFill(dev, level);
}

int artra_dev_max_speed(ArtraDevice *dev, const double floor[3], double vmax[3])
{
    if (EnqueueSpeeds(dev, floor, dev->vmax)) return 1;
    unsigned long long host[3];
    HIP_TRY(hipMemcpyAsync(host, dev->vmax, sizeof host, hipMemcpyDeviceToHost, dev->stream));
    HIP_TRY(hipStreamSynchronize(dev->stream));
    for (int s = 0; s < 3; ++s) memcpy(vmax + s, host + s, sizeof(double));
    return Harvest(dev);
}

int artra_dev_set_force_table(ArtraDevice *dev, const ArtraForceNode *nodes, int count, const ArtraForceBody *bodies,
                              int nbody)
{
    // the tables keep their allocations while they fit (moving bodies
    // rebuild them every half-step); the copies below wait for the stream
    HIP_TRY(hipStreamSynchronize(dev->stream));
    dev->forceCount = dev->forceBodyCount = 0;
    if (nbody <= 0) return 0;
    if (nbody > dev->forceBodyCap) {
        hipFree(dev->forceBodies);
        hipFree(dev->forceSums);
        dev->forceBodies = nullptr;
        dev->forceSums = nullptr;
        dev->forceBodyCap = 0;
        HIP_TRY(hipMalloc(&dev->forceBodies, nbody * sizeof(ArtraForceBody)));
        HIP_TRY(hipMalloc(&dev->forceSums, (size_t)nbody * 12 * sizeof(double)));
        dev->forceBodyCap = nbody;
    }
    HIP_TRY(hipMemcpy(dev->forceBodies, bodies, nbody * sizeof(ArtraForceBody), hipMemcpyHostToDevice));
    if (count > dev->forceNodeCap) {
        hipFree(dev->forceNodes);
        hipFree(dev->forceTerms);
        dev->forceNodes = nullptr;
        dev->forceTerms = nullptr;
        dev->forceNodeCap = 0;
        const int cap = count + count / 2;
        HIP_TRY(hipMalloc(&dev->forceNodes, (size_t)cap * sizeof(ArtraForceNode)));
        HIP_TRY(hipMalloc(&dev->forceTerms, (size_t)cap * 10 * sizeof(double)));
        dev->forceNodeCap = cap;
    }
    if (count > 0) HIP_TRY(hipMemcpy(dev->forceNodes, nodes, count * sizeof(ArtraForceNode), hipMemcpyHostToDevice));
    dev->forceCount = count;
    dev->forceBodyCount = nbody;
    return 0;
}

int artra_dev_surface_force(ArtraDevice *dev, double *sums)
{
    if (dev->forceBodyCount <= 0) return 0;
    if (dev->forceCount > 0) {
        hipLaunchKernelGGL(force_terms, dim3((dev->forceCount + 255) / 256), dim3(256), 0, dev->stream,
                           Level(dev, 0), dev->forceNodes, dev->forceCount, dev->forceBodies,
                           dev->par.gamma, dev->par.gasR, dev->par.refMu, dev->par.refT, dev->forceTerms);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(force_sum, dim3(dev->forceBodyCount), dim3(64), 0, dev->stream, dev->forceBodies,
                       dev->forceBodyCount, dev->forceTerms, dev->forceSums);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(sums, dev->forceSums, (size_t)dev->forceBodyCount * 12 * sizeof(double),
                           hipMemcpyDeviceToHost, dev->stream));
    HIP_TRY(hipStreamSynchronize(dev->stream));
    return 0;
}

int artra_dev_gather_primitive(ArtraDevice *dev, const int *nodes, int count, double *prim)
{
    if (count <= 0) return 0;
    if (count > dev->probeCap) {
        hipFree(dev->probeNodes);
        hipFree(dev->probePrim);
        dev->probeNodes = nullptr;
        dev->probePrim = nullptr;
        dev->probeCap = 0;
        HIP_TRY(hipMalloc(&dev->probeNodes, (size_t)count * sizeof(int)));
        HIP_TRY(hipMalloc(&dev->probePrim, (size_t)count * 6 * sizeof(double)));
        dev->probeCap = count;
    }
    for (int t = 0; t < count; ++t) {
        if (nodes[t] < 0 || nodes[t] >= dev->g.N) return Fail("probe node %d outside the grid", nodes[t]);
    }
    HIP_TRY(hipMemcpyAsync(dev->probeNodes, nodes, (size_t)count * sizeof(int), hipMemcpyHostToDevice, dev->stream));
    hipLaunchKernelGGL(gather_primitive, dim3((count + 255) / 256), dim3(256), 0, dev->stream, Level(dev, 0),
                       dev->probeNodes, count, dev->par.gamma, dev->par.gasR, dev->probePrim);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(prim, dev->probePrim, (size_t)count * 6 * sizeof(double), hipMemcpyDeviceToHost,
                           dev->stream));
    HIP_TRY(hipStreamSynchronize(dev->stream));
    return 0;
}

int artra_dev_march_steps(ArtraDevice *dev, double clock[ARTRA_DEV_MARCH_CLOCK], const double floor[3], double cfl, int steps,
                          double *dts, int *advanced)
{
    return MarchSteps(dev, MarchScratchAt(dev->marchBuf, kMarchChunk), clock, floor, cfl, steps, dts, advanced, true);
}

int artra_dev_march_noalloc(ArtraDevice *dev, double clock[ARTRA_DEV_MARCH_CLOCK], const double floor[3], double cfl, int steps,
                            double *dts, int *advanced)
{
    *advanced = 0;
    if (steps < 0 || steps > kNoAllocSteps) {
        return Fail("the allocation-free march takes 0 to %d steps per call (%d asked)", kNoAllocSteps, steps);
    }
    if (dev->timing) return Fail("kernel timing is on: its events are allocated, so the allocation-free march refuses it");
    void *base = gScratch[dev->ordinal];
    if (!base) return Fail("no scratch address for device %d", dev->ordinal);
    return MarchSteps(dev, MarchScratchAt(base, kNoAllocSteps), clock, floor, cfl, steps, dts, advanced, false);
}

int artra_dev_evolve(ArtraDevice *dev, double dt)
{
    const double h = 0.5 * dt;
    // gravity source half-steps around the fluid operators (fluid_dynamics.c:57-59, 112-114)
    if (dev->par.sState && SourceStages(dev, h)) return 1;
    if (dev->par.multidim == 1) {
        if (Stages(dev, h, 3, 0) || Stages(dev, h, 3, 0)) return 1;
    } else {
        static const int order[9][6] = {{2, 1, 0, 0, 1, 2},    {2, 1, 1, 2, -1, -1},   {2, 0, 0, 2, -1, -1},
                                        {1, 0, 0, 1, -1, -1},  {-1, -1, -1, -1, -1, -1}, {2, 2, -1, -1, -1, -1},
                                        {-1, -1, -1, -1, -1, -1}, {1, 1, -1, -1, -1, -1}, {0, 0, -1, -1, -1, -1}};
        const int c = dev->par.collapse;
        if (c < 0 || c > 8 || order[c][0] < 0) return Fail("unsupported collapse flag %d", c);
        for (int q = 0; q < 6 && order[c][q] >= 0; ++q) {
            const int more = (q + 1 < 6 && order[c][q + 1] >= 0);
            if (Stages(dev, h, order[c][q], more ? 1 + order[c][q + 1] : 0)) return 1;
        }
    }
    if (dev->par.sState && SourceStages(dev, h)) return 1;
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
    ++dev->configGen;
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

int artra_dev_set_indicator(ArtraDevice *dev, const float *params, float threshold)
{
    ++dev->configGen;
    if (params == nullptr) {
        dev->hybrid = 0;
        return 0;
    }
    if (dev->par.sScheme != 1) return Fail("the learned indicator routes WENO5 faces; this case uses WENO3");
    if (dev->exact) return Fail("the learned indicator runs with the fast arithmetic only");
    static const int order[9][6] = {{2, 1, 0, 0, 1, 2},    {2, 1, 1, 2, -1, -1},   {2, 0, 0, 2, -1, -1},
                                    {1, 0, 0, 1, -1, -1},  {-1, -1, -1, -1, -1, -1}, {2, 2, -1, -1, -1, -1},
                                    {-1, -1, -1, -1, -1, -1}, {1, 1, -1, -1, -1, -1}, {0, 0, -1, -1, -1, -1}};
    const int c = dev->par.collapse;
    for (int q = 0; c >= 0 && c <= 8 && q < 6 && order[c][q] >= 0; ++q) {
        if (!LineEligible(dev, order[c][q], true)) {
            return Fail("the learned indicator runs inside the fused line sweeps: it needs an inviscid split case "
                        "without immersed bodies whose line fill copies along lines and whose lines fit in LDS "
                        "(axis %d does not qualify)", order[c][q]);
        }
    }
    if (!dev->mlpParams) {
        HIP_TRY(hipMalloc(&dev->mlpParams, ARTRA_MLP_PARAMS * sizeof(float)));
        HIP_TRY(hipMalloc(&dev->mlpW1, 512 * sizeof(uint16_t)));
        HIP_TRY(hipMalloc(&dev->mlpW2, 1024 * sizeof(uint16_t)));
        HIP_TRY(hipMalloc(&dev->mlpBias, 96 * sizeof(float)));
        HIP_TRY(hipMalloc(&dev->hybCounts, kCountBytes));
        long most = 1;
        for (int q = 0; q < 3; ++q) most = std::max(most, LineShapeOf(dev, q, true).lines);
        HIP_TRY(hipMalloc(&dev->deferList, most * sizeof(int)));
        HIP_TRY(hipMalloc(&dev->deferCount, 4 * sizeof(int)));
        HIP_TRY(hipMalloc(&dev->deferTotal, 4 * sizeof(unsigned long long)));
    }
    HIP_TRY(hipMemset(dev->deferTotal, 0, 4 * sizeof(unsigned long long)));
    for (int q = 0; q < 3; ++q) {
        dev->leanFirst[q] = 1;
        dev->leanRetry[q] = 0;
        dev->leanLines[q] = 0;
    }
    std::vector<uint16_t> w1(512), w2(1024);
    for (int q = 0; q < 512; ++q) w1[q] = artra::f32_to_bf16(params[ARTRA_MLP_W1 + q]);
    for (int q = 0; q < 1024; ++q) w2[q] = artra::f32_to_bf16(params[ARTRA_MLP_W2 + q]);
    std::vector<float> bias(96);
    for (int q = 0; q < 32; ++q) {
        bias[q] = params[ARTRA_MLP_B1 + q];
        bias[32 + q] = params[ARTRA_MLP_B2 + q];
        bias[64 + q] = params[ARTRA_MLP_W3 + q];
    }
    HIP_TRY(hipMemcpy(dev->mlpParams, params, ARTRA_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dev->mlpW1, w1.data(), 512 * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dev->mlpW2, w2.data(), 1024 * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dev->mlpBias, bias.data(), 96 * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(dev->hybCounts, 0, kCountBytes));
    dev->mlpB3 = params[ARTRA_MLP_B3];
    dev->mlpThreshold = threshold;
    dev->hybrid = 1;
    return 0;
}

int artra_dev_set_indicator_options(ArtraDevice *dev, int every_stage, float smooth_variation)
{
    ++dev->configGen;
    if (every_stage != 0)
        return Fail("the indicator classifies once per directional operator; per-stage classification is not provided");
    if (!(smooth_variation >= 0.0f)) return Fail("the smooth-variation gate must be non-negative");
    dev->smoothVariation = smooth_variation;
    return 0;
}

int artra_dev_set_fused(ArtraDevice *dev, int fused)
{
    ++dev->configGen;
    if (!fused && dev->hybrid) return Fail("the learned indicator needs the fused line sweeps");
    dev->fused = (fused != 0);
    return 0;
}

int artra_dev_fused_axes(ArtraDevice *dev)
{
    int mask = 0;
    for (int p = 0; p < 3; ++p)
        if (LineEligible(dev, p, dev->hybrid != 0)) mask |= 1 << p;
    return mask;
}

// Sum of the tally slots (and their reset).
static int SumCounts(unsigned long long *slots, unsigned long long counts[kTallies], int reset)
{
    std::vector<unsigned long long> h((size_t)kCountSlots * kCountStride);
    HIP_TRY(hipMemcpy(h.data(), slots, kCountBytes, hipMemcpyDeviceToHost));
    for (int v = 0; v < kTallies; ++v) counts[v] = 0;
    for (int q = 0; q < kCountSlots; ++q)
        for (int v = 0; v < kTallies; ++v) counts[v] += h[(size_t)q * kCountStride + v];
    if (reset) HIP_TRY(hipMemset(slots, 0, kCountBytes));
    return 0;
}

int artra_dev_hybrid_counts(ArtraDevice *dev, unsigned long long counts[kTallies], int reset)
{
    if (!dev->hybCounts) {
        for (int v = 0; v < kTallies; ++v) counts[v] = 0;
        return 0;
    }
    HIP_TRY(hipStreamSynchronize(dev->stream));
    if (UpdateLeanPolicy(dev)) return 1;
    return SumCounts(dev->hybCounts, counts, reset);
}

long artra_dev_face_count(ArtraDevice *dev, int axis)
{
    const Grid &g = dev->g;
    long n = g.hi[axis] - g.lo[axis] + 1;
    for (int t = 0; t < 3; ++t) {
        if (t != axis) n *= g.hi[t] - g.lo[t];
    }
    return n;
}

int artra_dev_sample_faces(ArtraDevice *dev, int axis, float *feat, float *label)
{
    const long n = artra_dev_face_count(dev, axis);
    if (n <= 0) return 0;
    // the six-node stencil of the faces lo-1/2 .. hi-1/2 spans nodes lo-3 .. hi+2
    const int extent = (axis == 0) ? dev->g.nx : (axis == 1) ? dev->g.ny : dev->g.nz;
    if (dev->g.lo[axis] - 3 < 0 || dev->g.hi[axis] + 2 > extent - 1) {
        return Fail("axis %d has too few ghost layers for the six-node face stencil", axis);
    }
    float *df = nullptr, *dl = nullptr;
    HIP_TRY(hipMalloc(&df, n * artra::kFeat * sizeof(float)));
    HIP_TRY(hipMalloc(&dl, n * sizeof(float)));
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (axis == 0) hipLaunchKernelGGL(sample_faces<0>, dim3(blocks), dim3(256), 0, dev->stream, dev->g, Level(dev, 0), dev->mdl, df, dl);
    else if (axis == 1) hipLaunchKernelGGL(sample_faces<1>, dim3(blocks), dim3(256), 0, dev->stream, dev->g, Level(dev, 0), dev->mdl, df, dl);
    else hipLaunchKernelGGL(sample_faces<2>, dim3(blocks), dim3(256), 0, dev->stream, dev->g, Level(dev, 0), dev->mdl, df, dl);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    if (e == hipSuccess) e = hipMemcpy(feat, df, n * artra::kFeat * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(label, dl, n * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(df);
    hipFree(dl);
    if (e != hipSuccess) return Fail("face sampling failed: %s", hipGetErrorString(e));
    return 0;
}

int artra_dev_set_graph(ArtraDevice *dev, int on)
{
    ++dev->configGen;
    dev->useGraph = on ? 1 : 0;
    return 0;
}

int artra_dev_set_exact(ArtraDevice *dev, int exact)
{
    ++dev->configGen;
    dev->exact = (exact != 0);
    return 0;
}

int artra_dev_set_tuning(ArtraDevice *dev, int march_segment, int xblock)
{
    ++dev->configGen;
    (void)xblock;
    if (march_segment > 0) dev->segment = march_segment;
    return 0;
}


}  // extern "C"

// ------------------------------------------------------------ batched cases
struct ArtraBatch {
    std::vector<ArtraDevice *> cases;
    int C = 0;
    hipStream_t stream = nullptr;
    std::vector<int> threads;               // threads per line of each case (tuned)
    int grid = 0;                           // persistent blocks of the line launches
    LineArgs *dArgs[3] = {nullptr, nullptr, nullptr};
    BatchTile *dTiles[3] = {nullptr, nullptr, nullptr};
    int ntiles[3] = {0, 0, 0};
    size_t lds[3] = {0, 0, 0};
    BatchCase *dCases = nullptr;
    CflChunk *dChunks = nullptr;
    int nchunks = 0;
    GatherCase *dGather = nullptr;
    int *gCase = nullptr, *gDst = nullptr, *gSrc = nullptr;
    long gCount = 0;
    unsigned long long *vmax = nullptr;
    double *clock = nullptr, *dt = nullptr, *half = nullptr;
    int hybrid = 0;
    LineHybrid hy{};
    float *mlpBias = nullptr;
    uint16_t *mlpW1 = nullptr, *mlpW2 = nullptr;
    unsigned long long *counts = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graphHybrid = -1;
    int cus = 256;
};

namespace {

const int kBatchOrder[6] = {2, 1, 0, 0, 1, 2};

void BatchDropGraph(ArtraBatch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Tiles of axis p for the current threads-per-line choice, in case order.
int BatchBuildTiles(ArtraBatch *b, int p)
{
    std::vector<BatchTile> tiles;
    size_t lds = 0;
    for (int c = 0; c < b->C; ++c) {
        const ArtraDevice *dev = b->cases[c];
        const Grid &g = dev->g;
        const int t1 = (p == 0) ? 1 : 0, t2 = (p == 2) ? 1 : 2;
        const long lines = (long)(g.hi[t1] - g.lo[t1]) * (g.hi[t2] - g.lo[t2]);
        const int T = b->threads[c], P = 256 / T;
        const int ext[3] = {g.nx, g.ny, g.nz};
        lds = std::max(lds, (size_t)P * LineLdsBytes(ext[p], g.hi[p] - g.lo[p] + 1, true));
        for (long first = 0; first < lines; first += P) {
            BatchTile bt{};
            bt.c = c;
            bt.T = T;
            bt.count = (int)std::min<long>(P, lines - first);
            bt.first = first;
            tiles.push_back(bt);
        }
    }
    if (lds > 64 * 1024) return Fail("batched lines need %zu bytes of LDS per block (64 KB at most)", lds);
    hipFree(b->dTiles[p]);
    b->dTiles[p] = nullptr;
    HIP_TRY(hipMalloc(&b->dTiles[p], tiles.size() * sizeof(BatchTile)));
    HIP_TRY(hipMemcpy(b->dTiles[p], tiles.data(), tiles.size() * sizeof(BatchTile), hipMemcpyHostToDevice));
    b->ntiles[p] = (int)tiles.size();
    b->lds[p] = lds;
    BatchDropGraph(b);
    return 0;
}

void BatchLaunchLines(ArtraBatch *b, int p, hipStream_t st)
{
    const unsigned grid = (unsigned)std::max(1, std::min(b->grid, b->ntiles[p]));
    const bool hyb = b->hybrid != 0;
    const FluxModel mdl = b->cases[0]->mdl;
#define ARTRA_BATCH_LAUNCH(S)                                                                                   \
    do {                                                                                                        \
        if (hyb)                                                                                                \
            hipLaunchKernelGGL((rk_batch<S, 6, true>), dim3(grid), dim3(256), b->lds[p], st, b->dTiles[p],      \
                               b->ntiles[p], b->dArgs[p], mdl, b->hy);                                          \
        else                                                                                                    \
            hipLaunchKernelGGL((rk_batch<S, 6, false>), dim3(grid), dim3(256), b->lds[p], st, b->dTiles[p],     \
                               b->ntiles[p], b->dArgs[p], mdl, b->hy);                                          \
    } while (0)
    if (p == 0) ARTRA_BATCH_LAUNCH(0);
    else if (p == 1) ARTRA_BATCH_LAUNCH(1);
    else ARTRA_BATCH_LAUNCH(2);
#undef ARTRA_BATCH_LAUNCH
}

// One time step of every case: speeds, dt, then Z Y X X Y Z, each operator a
// persistent launch over all cases' lines followed by every case's fill.
int BatchStep(ArtraBatch *b, hipStream_t st)
{
    hipLaunchKernelGGL(fill_u64, dim3((3 * b->C + 255) / 256), dim3(256), 0, st, b->vmax, 3 * b->C, 0ull);
    hipLaunchKernelGGL(cfl_batch, dim3((unsigned)std::min(b->nchunks, 4 * b->cus)), dim3(256), 0, st, b->dChunks,
                       b->nchunks, b->dCases, b->vmax);
    hipLaunchKernelGGL(dt_batch, dim3((b->C + 63) / 64), dim3(64), 0, st, b->dCases, b->C, b->vmax, b->clock, b->dt,
                       b->half);
    for (int q = 0; q < 6; ++q) {
        BatchLaunchLines(b, kBatchOrder[q], st);
        if (b->gCount > 0) {
            hipLaunchKernelGGL(gather_batch, dim3((unsigned)((b->gCount + 255) / 256)), dim3(256), 0, st, b->dGather,
                               b->gCase, b->gDst, b->gSrc, b->gCount);
        }
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int artra_dev_batch_create(ArtraDevice *const *devs, int count, const double *clock, const double *cfl,
                           ArtraBatch **out)
{
    *out = nullptr;
    if (count < 1) return Fail("a batch needs at least one case");
    const ArtraDevice *d0 = devs[0];
    for (int c = 0; c < count; ++c) {
        const ArtraDevice *dev = devs[c];
        if (dev->par.collapse != 0) return Fail("batched case %d: only three-dimensional cases are batched", c);
        if (dev->par.sState) return Fail("batched case %d: the gravity source operator is not batched", c);
        for (int p = 0; p < 3; ++p) {
            if (!LineEligible(dev, p, true)) {
                return Fail("batched case %d: axis %d cannot run as fused line sweeps (inviscid, split, no "
                            "immersed body, resolvable fills and lines that fit in LDS are required)", c, p);
            }
        }
        if (dev->par.sScheme != 1 || dev->exact) return Fail("batched case %d: WENO5 with the fast arithmetic only", c);
        if (dev->par.tScheme != d0->par.tScheme || dev->mdl.gamma != d0->mdl.gamma || dev->mdl.roe != d0->mdl.roe ||
            dev->mdl.sw != d0->mdl.sw) {
            return Fail("batched case %d: every case must share the time scheme and flux model", c);
        }
        if (dev->g.hi[0] - dev->g.lo[0] > 256 || dev->g.hi[1] - dev->g.lo[1] > 256 || dev->g.hi[2] - dev->g.lo[2] > 256)
            return Fail("batched case %d: lines longer than 256 interior nodes", c);
    }
    ArtraBatch *b = new ArtraBatch();
    b->cases.assign(devs, devs + count);
    b->C = count;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        b->cus = prop.multiProcessorCount;
    }
    b->grid = 4 * b->cus;
    auto fail = [&](int rc) {
        artra_dev_batch_destroy(b);
        return rc;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(Fail("stream creation failed"));
    // threads per line: the smallest wavefront multiple that covers the line with two nodes per lane
    b->threads.resize(count);
    for (int c = 0; c < count; ++c) {
        const Grid &g = devs[c]->g;
        int longest = 0;
        for (int p = 0; p < 3; ++p) longest = std::max(longest, g.hi[p] - g.lo[p]);
        int T = 64;
        while (T < longest) T *= 2;         // one interior node per thread
        // fewer, wider lines per block where the LDS of 256 / T lines would not fit
        const int ext[3] = {g.nx, g.ny, g.nz};
        for (int p = 0; p < 3; ++p) {
            while (T < 256 && (size_t)(256 / T) * LineLdsBytes(ext[p], g.hi[p] - g.lo[p] + 1, true) > 64 * 1024) T *= 2;
        }
        b->threads[c] = std::min(T, 256);
    }
    // line arguments of every case and axis
    std::vector<LineArgs> hostArgs[3];
    for (int p = 0; p < 3; ++p) {
        std::vector<LineArgs> args(count);
        for (int c = 0; c < count; ++c) {
            ArtraDevice *dv = devs[c];
            const Grid &g = dv->g;
            const int ext[3] = {g.nx, g.ny, g.nz};
            const int stride[3] = {1, g.nx, g.nx * g.ny};
            const int t1 = (p == 0) ? 1 : 0, t2 = (p == 2) ? 1 : 2;
            LineArgs la{};
            la.U = Level(dv, 0);
            la.n = ext[p];
            la.lo = g.lo[p];
            la.hi = g.hi[p];
            la.n1 = g.hi[t1] - g.lo[t1];
            la.lo1 = g.lo[t1];
            la.lo2 = g.lo[t2];
            la.stride = stride[p];
            la.stride1 = stride[t1];
            la.stride2 = stride[t2];
            if (dv->par.tScheme == 0) {
                la.stages = 2;
                la.a[0] = 0.0; la.b[0] = 1.0;
                la.a[1] = 1.0 / 2.0; la.b[1] = 1.0 / 2.0;
            } else {
                la.stages = 3;
                la.a[0] = 0.0; la.b[0] = 1.0;
                la.a[1] = 3.0 / 4.0; la.b[1] = 1.0 / 4.0;
                la.a[2] = 1.0 / 3.0; la.b[2] = 2.0 / 3.0;
            }
            la.r = 0.0;
            la.dd = dv->par.dd[p];
            la.gamma = dv->par.gamma;
            const ArtraDevice::LineRules &R = dv->lines[p];
            la.nrules = R.count;
            for (int e = 0; e < R.count; ++e) {
                la.rdst[e] = R.dst[e];
                la.rsrc[e] = R.src[e];
            }
            memcpy(la.given, R.given, sizeof la.given);
            args[c] = la;
        }
        if (hipMalloc(&b->dArgs[p], count * sizeof(LineArgs)) != hipSuccess) return fail(Fail("out of device memory"));
        hostArgs[p] = args;
    }
    if (hipMalloc(&b->vmax, 3 * count * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&b->clock, kClock * count * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->dt, count * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->half, count * sizeof(double)) != hipSuccess) {
        return fail(Fail("out of device memory"));
    }
    for (int p = 0; p < 3; ++p) {
        // every operator of a case advances by half of the case's step, read on the device
        for (int c = 0; c < count; ++c) hostArgs[p][c].dt = b->half + c;
        if (hipMemcpy(b->dArgs[p], hostArgs[p].data(), count * sizeof(LineArgs), hipMemcpyHostToDevice) != hipSuccess)
            return fail(Fail("upload of the line arguments failed"));
    }
    if (hipMemcpy(b->clock, clock, kClock * count * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return fail(Fail("upload of the case clocks failed"));
    for (int p = 0; p < 3; ++p)
        if (BatchBuildTiles(b, p)) return fail(1);
    // CFL chunks and per-case data
    std::vector<BatchCase> bcs(count);
    std::vector<CflChunk> chunks;
    std::vector<GatherCase> gcs(count);
    for (int c = 0; c < count; ++c) {
        ArtraDevice *dv = devs[c];
        BatchCase &x = bcs[c];
        x.U = Level(dv, 0);
        x.fluid = dv->fluid;
        x.g = dv->g;
        for (int s = 0; s < 3; ++s) x.d[s] = dv->par.d[s];
        x.cfl = cfl[c];
        x.gamma = dv->par.gamma;
        x.gasR = dv->par.gasR;
        const Grid &g = dv->g;
        const long total = (long)(g.hi[0] - g.lo[0]) * (g.hi[1] - g.lo[1]) * (g.hi[2] - g.lo[2]);
        const long per = 256L * 16;
        for (long st = 0; st < total; st += per) chunks.push_back(CflChunk{c, 0, st, std::min(per, total - st)});
        GatherCase &gc = gcs[c];
        gc.U = Level(dv, 0);
        gc.gamma = dv->par.gamma;
        for (int q = 0; q < 7; ++q)
            for (int v = 0; v < 6; ++v) gc.given[q][v] = dv->par.varBC[q][v];
        b->gCount += dv->gCount[0];
    }
    b->nchunks = (int)chunks.size();
    if (hipMalloc(&b->dCases, count * sizeof(BatchCase)) != hipSuccess ||
        hipMalloc(&b->dChunks, chunks.size() * sizeof(CflChunk)) != hipSuccess ||
        hipMalloc(&b->dGather, count * sizeof(GatherCase)) != hipSuccess) {
        return fail(Fail("out of device memory"));
    }
    if (hipMemcpy(b->dCases, bcs.data(), count * sizeof(BatchCase), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->dChunks, chunks.data(), chunks.size() * sizeof(CflChunk), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->dGather, gcs.data(), count * sizeof(GatherCase), hipMemcpyHostToDevice) != hipSuccess) {
        return fail(Fail("upload of the batch tables failed"));
    }
    if (b->gCount > 0) {
        if (hipMalloc(&b->gCase, b->gCount * sizeof(int)) != hipSuccess ||
            hipMalloc(&b->gDst, b->gCount * sizeof(int)) != hipSuccess ||
            hipMalloc(&b->gSrc, b->gCount * sizeof(int)) != hipSuccess) {
            return fail(Fail("out of device memory"));
        }
        long off = 0;
        for (int c = 0; c < count; ++c) {
            const long n = devs[c]->gCount[0];
            if (n == 0) continue;
            if (hipMemcpy(b->gDst + off, devs[c]->gDst[0], n * sizeof(int), hipMemcpyDeviceToDevice) != hipSuccess ||
                hipMemcpy(b->gSrc + off, devs[c]->gSrc[0], n * sizeof(int), hipMemcpyDeviceToDevice) != hipSuccess) {
                return fail(Fail("copy of the boundary tables failed"));
            }
            hipLaunchKernelGGL(fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, b->gCase + off, n, c);
            off += n;
        }
        if (hipStreamSynchronize(b->stream) != hipSuccess) return fail(Fail("boundary table set-up failed"));
    }
    *out = b;
    return 0;
}

void artra_dev_batch_destroy(ArtraBatch *b)
{
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    BatchDropGraph(b);
    for (int p = 0; p < 3; ++p) {
        hipFree(b->dArgs[p]);
        hipFree(b->dTiles[p]);
    }
    hipFree(b->dCases);
    hipFree(b->dChunks);
    hipFree(b->dGather);
    hipFree(b->gCase);
    hipFree(b->gDst);
    hipFree(b->gSrc);
    hipFree(b->vmax);
    hipFree(b->clock);
    hipFree(b->dt);
    hipFree(b->half);
    hipFree(b->mlpBias);
    hipFree(b->mlpW1);
    hipFree(b->mlpW2);
    hipFree(b->counts);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

int artra_dev_batch_set_indicator(ArtraBatch *b, const float *params, float threshold)
{
    BatchDropGraph(b);
    if (!params) {
        b->hybrid = 0;
        return 0;
    }
    if (!b->mlpBias) {
        HIP_TRY(hipMalloc(&b->mlpBias, 96 * sizeof(float)));
        HIP_TRY(hipMalloc(&b->mlpW1, 512 * sizeof(uint16_t)));
        HIP_TRY(hipMalloc(&b->mlpW2, 1024 * sizeof(uint16_t)));
        HIP_TRY(hipMalloc(&b->counts, kCountBytes));
    }
    std::vector<uint16_t> w1(512), w2(1024);
    for (int q = 0; q < 512; ++q) w1[q] = artra::f32_to_bf16(params[ARTRA_MLP_W1 + q]);
    for (int q = 0; q < 1024; ++q) w2[q] = artra::f32_to_bf16(params[ARTRA_MLP_W2 + q]);
    std::vector<float> bias(96);
    for (int q = 0; q < 32; ++q) {
        bias[q] = params[ARTRA_MLP_B1 + q];
        bias[32 + q] = params[ARTRA_MLP_B2 + q];
        bias[64 + q] = params[ARTRA_MLP_W3 + q];
    }
    HIP_TRY(hipMemcpy(b->mlpW1, w1.data(), 512 * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->mlpW2, w2.data(), 1024 * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->mlpBias, bias.data(), 96 * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(b->counts, 0, kCountBytes));
    b->hy.W1 = b->mlpW1;
    b->hy.W2 = b->mlpW2;
    b->hy.bias = b->mlpBias;
    b->hy.b3 = params[ARTRA_MLP_B3];
    b->hy.threshold = threshold;
    b->hy.counts = b->counts;
    b->hy.smoothVariation = 1.0e-3f;
    b->hybrid = 1;
    return 0;
}

int artra_dev_batch_counts(ArtraBatch *b, unsigned long long counts[kTallies], int reset)
{
    for (int v = 0; v < kTallies; ++v) counts[v] = 0;
    if (!b->counts) return 0;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return SumCounts(b->counts, counts, reset);
}

// steps time steps of every case; with use_graph the step is captured once
// in a HIP graph and replayed.  ms (optional) receives the wall time of the
// steps measured with events on the batch stream.
int artra_dev_batch_march(ArtraBatch *b, int steps, int use_graph, double *ms)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = 0;
    if (use_graph && (!b->exec || b->graphHybrid != b->hybrid)) {
        BatchDropGraph(b);
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int r1 = BatchStep(b, b->stream);
        const hipError_t ce = hipStreamEndCapture(b->stream, &b->graph);
        if (r1) return 1;
        HIP_TRY(ce);
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graphHybrid = b->hybrid;
    }
    HIP_TRY(hipEventRecord(e0, b->stream));
    for (int q = 0; q < steps && !rc; ++q) {
        if (use_graph) {
            const hipError_t e = hipGraphLaunch(b->exec, b->stream);
            if (e != hipSuccess) rc = Fail("graph launch failed: %s", hipGetErrorString(e));
        } else {
            rc = BatchStep(b, b->stream);
        }
    }
    HIP_TRY(hipEventRecord(e1, b->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float el = 0.0f;
    HIP_TRY(hipEventElapsedTime(&el, e0, e1));
    if (ms) *ms = el;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

int artra_dev_batch_clock(ArtraBatch *b, double *clock)
{
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(clock, b->clock, kClock * b->C * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// Tile shape of every case: threads[c] threads per line (64, 128 or 256, at
// least the case's longest interior line, with 256 / T lines in LDS).
int artra_dev_batch_set_threads(ArtraBatch *b, const int *threads)
{
    for (int c = 0; c < b->C; ++c) {
        const Grid &g = b->cases[c]->g;
        int longest = 0;
        for (int p = 0; p < 3; ++p) longest = std::max(longest, g.hi[p] - g.lo[p]);
        const int T = threads[c];
        if ((T != 64 && T != 128 && T != 256) || T < longest)
            return Fail("batched case %d: %d threads per line is not a tile shape for lines of %d nodes", c, T, longest);
    }
    const std::vector<int> saved = b->threads;
    b->threads.assign(threads, threads + b->C);
    for (int p = 0; p < 3; ++p) {
        if (BatchBuildTiles(b, p)) {
            b->threads = saved;
            for (int q = 0; q < 3; ++q) BatchBuildTiles(b, q);
            return 1;
        }
    }
    return 0;
}

int artra_dev_batch_set_clock(ArtraBatch *b, const double *clock)
{
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(b->clock, clock, kClock * b->C * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

// Autotuner: for every grid size in the batch (cases of equal line length)
// time the three directional launches with each admissible threads-per-line
// choice (the tile shape: 256 / T lines of T threads per block), keep the
// fastest; then time whole steps for persistent grids of 2, 4 and 8 blocks
// per compute unit.  The field is restored afterwards (the tuning steps run
// on a scratch copy of level TO in level TN of every case).
int artra_dev_batch_tune(ArtraBatch *b, int reps, int *threads_out, int *grid_out)
{
    if (reps < 1) reps = 1;
    // save TO of every case into TN, restore at the end
    for (ArtraDevice *dv : b->cases) {
        HIP_TRY(hipMemcpyAsync(Level(dv, 1), Level(dv, 0), 5 * dv->g.N * sizeof(double), hipMemcpyDeviceToDevice,
                               b->stream));
    }
    std::vector<double> clockSave(kClock * b->C);
    HIP_TRY(hipMemcpyAsync(clockSave.data(), b->clock, clockSave.size() * sizeof(double), hipMemcpyDeviceToHost,
                           b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // one dt so the line launches have a valid step
    if (BatchStep(b, b->stream)) return 1;
    // distinct longest-line lengths
    std::vector<int> lengths;
    for (ArtraDevice *dv : b->cases) {
        int L = 0;
        for (int p = 0; p < 3; ++p) L = std::max(L, dv->g.hi[p] - dv->g.lo[p]);
        if (std::find(lengths.begin(), lengths.end(), L) == lengths.end()) lengths.push_back(L);
    }
    const std::vector<int> saved = b->threads;
    for (int L : lengths) {
        double best = 1e300;
        int bestT = 0;
        for (int T = 64; T <= 256; T *= 2) {
            if (T < L) continue;                // one interior node per thread
            for (int c = 0; c < b->C; ++c) {
                int Lc = 0;
                for (int p = 0; p < 3; ++p) Lc = std::max(Lc, b->cases[c]->g.hi[p] - b->cases[c]->g.lo[p]);
                b->threads[c] = (Lc == L) ? T : saved[c];
            }
            bool ok = true;
            for (int p = 0; p < 3 && ok; ++p) ok = BatchBuildTiles(b, p) == 0;
            if (!ok) continue;
            HIP_TRY(hipEventRecord(e0, b->stream));
            for (int r = 0; r < reps; ++r)
                for (int p = 0; p < 3; ++p) BatchLaunchLines(b, p, b->stream);
            HIP_TRY(hipEventRecord(e1, b->stream));
            HIP_TRY(hipEventSynchronize(e1));
            HIP_TRY(hipGetLastError());
            float el = 0.0f;
            HIP_TRY(hipEventElapsedTime(&el, e0, e1));
            if (el < best) {
                best = el;
                bestT = T;
            }
        }
        if (bestT == 0) return Fail("no admissible tile shape for lines of %d nodes", L);
        for (int c = 0; c < b->C; ++c) {
            int Lc = 0;
            for (int p = 0; p < 3; ++p) Lc = std::max(Lc, b->cases[c]->g.hi[p] - b->cases[c]->g.lo[p]);
            if (Lc == L) b->threads[c] = bestT;
        }
    }
    for (int p = 0; p < 3; ++p)
        if (BatchBuildTiles(b, p)) return 1;
    double bestG = 1e300;
    int grid = b->grid;
    for (int m = 2; m <= 8; m *= 2) {
        b->grid = m * b->cus;
        HIP_TRY(hipEventRecord(e0, b->stream));
        for (int r = 0; r < reps; ++r)
            if (BatchStep(b, b->stream)) return 1;
        HIP_TRY(hipEventRecord(e1, b->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float el = 0.0f;
        HIP_TRY(hipEventElapsedTime(&el, e0, e1));
        if (el < bestG) {
            bestG = el;
            grid = b->grid;
        }
    }
    b->grid = grid;
    BatchDropGraph(b);
    // restore the fields and clocks
    for (ArtraDevice *dv : b->cases) {
        HIP_TRY(hipMemcpyAsync(Level(dv, 0), Level(dv, 1), 5 * dv->g.N * sizeof(double), hipMemcpyDeviceToDevice,
                               b->stream));
    }
    HIP_TRY(hipMemcpyAsync(b->clock, clockSave.data(), clockSave.size() * sizeof(double), hipMemcpyHostToDevice,
                           b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->counts) HIP_TRY(hipMemset(b->counts, 0, kCountBytes));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    for (int c = 0; c < b->C; ++c)
        if (threads_out) threads_out[c] = b->threads[c];
    if (grid_out) *grid_out = b->grid;
    return 0;
}

}  // extern "C"

// ------------------------------------------ moving-body classification API
extern "C" {

int artra_dev_geo_init(ArtraDevice *dev, const ArtraGeoSetup *setup, const int *flags, int nbody)
{
    const long N = dev->g.N;
    HIP_TRY(hipStreamSynchronize(dev->stream));
    if (!dev->geoFlags) HIP_TRY(hipMalloc(&dev->geoFlags, (size_t)4 * N * sizeof(int)));
    HIP_TRY(hipMemcpy(dev->geoFlags, flags, (size_t)4 * N * sizeof(int), hipMemcpyHostToDevice));
    if ((int)dev->geoBodies.size() != nbody) {
        for (GeoBodyDev &b : dev->geoBodies) {
            hipFree(b.f); hipFree(b.e); hipFree(b.block); hipHostFree(b.stage); hipFree(b.sph); hipFree(b.bsph);
        }
        dev->geoBodies.assign(nbody, GeoBodyDev());
        hipFree(dev->geoMoving);
        dev->geoMoving = nullptr;
        HIP_TRY(hipMalloc(&dev->geoMoving, (size_t)(nbody > 0 ? nbody : 1) * sizeof(int)));
    }
    dev->geo = *setup;
    return 0;
}

int artra_dev_geo_classify(ArtraDevice *dev, const ArtraGeoBody *bodies, int nbody, const int region[3][2],
                           const int **packed, long *count)
{
    *count = 0;
    *packed = nullptr;
    if (!dev->geoFlags || nbody != (int)dev->geoBodies.size()) return Fail("moving-body classification not set up");
    ++dev->configGen;
    dev->allFluid = 0;                              // bodies present: the stage-by-stage sweeps
    const Grid &g = dev->g;
    const long N = g.N;
    int *did = dev->geoFlags, *fid = did + N, *lid = fid + N, *gst = lid + N;
    HIP_TRY(hipStreamSynchronize(dev->stream));
    std::vector<int> moving(nbody > 0 ? nbody : 1, 0);
    for (int n = 0; n < nbody; ++n) {
        const ArtraGeoBody &bd = bodies[n];
        moving[n] = bd.moving;
        if (bd.pv.faceN <= 0) continue;
        GeoBodyDev &gb = dev->geoBodies[n];
        const size_t fb = (size_t)bd.pv.faceN * 3 * sizeof(double), eb = (size_t)bd.pv.edgeN * 3 * sizeof(double),
                     vb = (size_t)bd.pv.vertN * 3 * sizeof(double);
        if (!gb.uploaded) {
            gb.faceN = bd.pv.faceN;
            gb.edgeN = bd.pv.edgeN;
            gb.vertN = bd.pv.vertN;
            HIP_TRY(hipMalloc(&gb.f, (size_t)gb.faceN * 3 * sizeof(int)));
            HIP_TRY(hipMalloc(&gb.e, (size_t)gb.edgeN * 4 * sizeof(int)));
            HIP_TRY(hipMalloc(&gb.block, 2 * vb + fb + eb));
            HIP_TRY(hipMalloc(&gb.sph, (size_t)gb.faceN * sizeof(double4)));
            HIP_TRY(hipMalloc(&gb.bsph, (size_t)((gb.faceN + 31) / 32) * sizeof(double4)));
            HIP_TRY(hipHostMalloc(&gb.stage, 2 * vb + fb + eb));
            gb.v = gb.block;
            gb.Nf = gb.v + vb / sizeof(double);
            gb.Ne = gb.Nf + fb / sizeof(double);
            gb.Nv = gb.Ne + eb / sizeof(double);
            HIP_TRY(hipMemcpy(gb.f, bd.pv.f, (size_t)gb.faceN * 3 * sizeof(int), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(gb.e, bd.pv.e, (size_t)gb.edgeN * 4 * sizeof(int), hipMemcpyHostToDevice));
        } else if (gb.faceN != bd.pv.faceN || gb.edgeN != bd.pv.edgeN || gb.vertN != bd.pv.vertN) {
            return Fail("body %d changed its facet count", n);
        }
        if (!gb.uploaded || bd.moving) {
            // the stream is idle here (synchronized above): the staging copy
            // of the last upload is free
            char *st = (char *)gb.stage;
            memcpy(st, bd.pv.v, vb);
            memcpy(st + vb, bd.pv.Nf, fb);
            memcpy(st + vb + fb, bd.pv.Ne, eb);
            memcpy(st + vb + fb + eb, bd.pv.Nv, vb);
            HIP_TRY(hipMemcpyAsync(gb.block, gb.stage, 2 * vb + fb + eb, hipMemcpyHostToDevice, dev->stream));
            ArtraPolyView dv = bd.pv;
            dv.f = (const int (*)[3])gb.f;
            dv.v = (const double (*)[3])gb.v;
            hipLaunchKernelGGL(geo_spheres, dim3((unsigned)((gb.faceN + 255) / 256)), dim3(256), 0, dev->stream, dv,
                               gb.sph, gb.bsph);
            HIP_TRY(hipGetLastError());
        }
        gb.uploaded = true;
    }
    HIP_TRY(hipMemcpy(dev->geoMoving, moving.data(), moving.size() * sizeof(int), hipMemcpyHostToDevice));
    // the region within the interior range
    GeoBox rb;
    rb.count = 1;
    for (int s = 0; s < 3; ++s) {
        const int lo = region[s][0] > g.lo[s] ? region[s][0] : g.lo[s];
        const int hi = region[s][1] < g.hi[s] ? region[s][1] : g.hi[s];
        rb.lo[s] = lo;
        rb.n[s] = hi > lo ? hi - lo : 0;
        rb.count *= rb.n[s];
    }
    if (rb.count == 0) return 0;
    if (5 * rb.count > dev->geoPackedCap) {
        hipFree(dev->geoPacked);
        hipHostFree(dev->geoPackedHost);
        dev->geoPacked = nullptr;
        dev->geoPackedHost = nullptr;
        dev->geoPackedCap = 0;
        HIP_TRY(hipMalloc(&dev->geoPacked, (size_t)5 * rb.count * sizeof(int)));
        HIP_TRY(hipHostMalloc(&dev->geoPackedHost, (size_t)5 * rb.count * sizeof(int)));
        dev->geoPackedCap = 5 * rb.count;
    }
    const GeoSetup gs = dev->geo;
    hipLaunchKernelGGL(geo_reset, dim3((unsigned)((rb.count + 255) / 256)), dim3(256), 0, dev->stream, g, rb, did, fid,
                       lid, gst, dev->geoMoving, nbody);
    HIP_TRY(hipGetLastError());
    for (int n = 0; n < nbody; ++n) {
        const ArtraGeoBody &bd = bodies[n];
        if (!bd.moving) continue;
        GeoBox bb;
        bb.count = 1;
        for (int s = 0; s < 3; ++s) {
            bb.lo[s] = bd.box[s][0];
            bb.n[s] = bd.box[s][1] > bd.box[s][0] ? bd.box[s][1] - bd.box[s][0] : 0;
            bb.count *= bb.n[s];
        }
        if (bb.count == 0) continue;
        if (bd.pv.faceN <= 0) {
            hipLaunchKernelGGL(geo_sphere, dim3((unsigned)((bb.count + 255) / 256)), dim3(256), 0, dev->stream, g, gs,
                               bb, did, fid, bd.O[0], bd.O[1], bd.O[2], bd.r, n);
        } else {
            const GeoBodyDev &gb = dev->geoBodies[n];
            ArtraPolyView pv;
            pv.faceN = gb.faceN;
            pv.edgeN = gb.edgeN;
            pv.vertN = gb.vertN;
            pv.f = (const int (*)[3])gb.f;
            pv.e = (const int (*)[4])gb.e;
            pv.v = (const double (*)[3])gb.v;
            pv.Nf = (const double (*)[3])gb.Nf;
            pv.Ne = (const double (*)[3])gb.Ne;
            pv.Nv = (const double (*)[3])gb.Nv;
            hipLaunchKernelGGL(geo_mesh, dim3((unsigned)((bb.count + 3) / 4)), dim3(256), 0, dev->stream, g, gs, bb,
                               did, fid, pv, gb.sph, gb.bsph, n);
        }
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(geo_interface, dim3((unsigned)((rb.count + 255) / 256)), dim3(256), 0, dev->stream, g, gs, rb,
                       did, fid, lid, gst, dev->fluid, dev->geoPacked);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(dev->geoPackedHost, dev->geoPacked, (size_t)5 * rb.count * sizeof(int),
                           hipMemcpyDeviceToHost, dev->stream));
    HIP_TRY(hipStreamSynchronize(dev->stream));
    *packed = dev->geoPackedHost;
    *count = rb.count;
    return 0;
}

}  // extern "C"

extern "C" {

int artra_dev_build_ghosts(ArtraDevice *dev, const ArtraGeoBody *bodies, int nbody, double dd[3], double tinyL,
                           int ibmLayer, int *fallback)
{
    *fallback = 0;
    if (!dev->geoFlags) return Fail("moving-body classification not set up");
    ++dev->configGen;
    const Grid &g = dev->g;
    const long N = g.N;
    const int *did = dev->geoFlags, *fid = did + N, *gst = fid + 2 * N;
    IbTable &tb = dev->ib;
    const int layers = dev->geo.gl;
    if (layers + 3 > kGeoCounters) return Fail("%d ghost layers exceed the table's %d", layers, kGeoCounters - 3);
    // room for the last table's ghosts and half again
    const int want = std::max(4096, tb.count + tb.count / 2);
    if (want > tb.ghostCap || (long)want * kGhostSlot > tb.stencilCap) {
        HIP_TRY(hipStreamSynchronize(dev->stream));
        hipFree(tb.ghosts);
        hipFree(tb.snode);
        hipFree(tb.sw);
        tb = IbTable();
        HIP_TRY(hipMalloc(&tb.ghosts, (size_t)want * sizeof(GhostDev)));
        HIP_TRY(hipMalloc(&tb.snode, (size_t)want * kGhostSlot * sizeof(int)));
        HIP_TRY(hipMalloc(&tb.sw, (size_t)want * kGhostSlot * sizeof(double)));
        tb.ghostCap = want;
        tb.stencilCap = want * kGhostSlot;
    }
    if (!dev->geoCounters) HIP_TRY(hipMalloc(&dev->geoCounters, kGeoCounters * sizeof(int)));
    if (!dev->geoCountsHost) HIP_TRY(hipHostMalloc(&dev->geoCountsHost, kGeoCounters * sizeof(int)));
    HIP_TRY(hipMemsetAsync(dev->geoCounters, 0, 2 * sizeof(int), dev->stream));
    GhostBuildArgs a;
    a.g = g;
    a.gs = dev->geo;
    for (int s = 0; s < 3; ++s) a.dd[s] = dd[s];
    a.tinyL = tinyL;
    a.ibmLayer = ibmLayer;
    a.cap = tb.ghostCap;
    for (int r = 1; r <= layers; ++r) {
        for (int n = 0; n < nbody; ++n) {
            const ArtraGeoBody &bd = bodies[n];
            GeoBox bb;
            bb.count = 1;
            for (int s = 0; s < 3; ++s) {
                bb.lo[s] = bd.box[s][0];
                bb.n[s] = bd.box[s][1] > bd.box[s][0] ? bd.box[s][1] - bd.box[s][0] : 0;
                bb.count *= bb.n[s];
            }
            if (bb.count == 0) continue;
            GhostBody gbody;
            gbody.pv = bd.pv;
            if (bd.pv.faceN > 0) {
                const GeoBodyDev &gb = dev->geoBodies[n];
                if (!gb.uploaded) return Fail("body %d has no device arrays", n);
                gbody.pv.f = (const int (*)[3])gb.f;
                gbody.pv.e = (const int (*)[4])gb.e;
                gbody.pv.v = (const double (*)[3])gb.v;
                gbody.pv.Nf = (const double (*)[3])gb.Nf;
                gbody.pv.Ne = (const double (*)[3])gb.Ne;
                gbody.pv.Nv = (const double (*)[3])gb.Nv;
            }
            for (int s = 0; s < 3; ++s) {
                gbody.O[s] = bd.O[s];
                gbody.V[s] = bd.V[s];
                gbody.W[s] = bd.W[s];
            }
            gbody.r = bd.r;
            gbody.cf = bd.cf;
            gbody.T = bd.T;
            gbody.body = n;
            hipLaunchKernelGGL(ib_build, dim3((unsigned)((bb.count + 255) / 256)), dim3(256), 0, dev->stream, a, bb, r,
                               gbody, did, fid, gst, tb.ghosts, tb.snode, tb.sw, dev->geoCounters,
                               dev->geoCounters + 1);
            HIP_TRY(hipGetLastError());
        }
        // the table's end after layer r, kept on the device until all are built
        HIP_TRY(hipMemcpyAsync(dev->geoCounters + 2 + r, dev->geoCounters, sizeof(int), hipMemcpyDeviceToDevice,
                               dev->stream));
    }
    int *c = dev->geoCountsHost;
    HIP_TRY(hipMemcpyAsync(c, dev->geoCounters, (size_t)(layers + 3) * sizeof(int), hipMemcpyDeviceToHost,
                           dev->stream));
    HIP_TRY(hipStreamSynchronize(dev->stream));
    if (c[1] || c[0] > tb.ghostCap) {
        tb.count = std::min(c[0], tb.ghostCap);  // the next build allocates room for them
        *fallback = 1;
        return 0;
    }
    std::vector<int> layerEnd(layers + 1, 0);
    for (int r = 1; r <= layers; ++r) layerEnd[r] = c[2 + r];
    tb.count = layerEnd[layers];
    dev->layerStart.assign(layerEnd.begin(), layerEnd.end());
    dev->ibImageOnly = (ibmLayer >= layers);
    return 0;
}


int artra_dev_build_force_table(ArtraDevice *dev, const ArtraGeoBody *bodies, int nbody, const int *slotBody,
                                const int (*box)[3][2], const int *viscous, int nb, int *gstN, int *lidN)
{
    if (!dev->geoFlags) return Fail("moving-body classification not set up");
    dev->forceCount = dev->forceBodyCount = 0;
    if (nb <= 0) return 0;
    const long N = dev->g.N;
    const int *did = dev->geoFlags, *fid = did + N, *lid = fid + N, *gst = lid + N;
    if (nb > dev->forceBodyCap || 1 + 2 * nb > dev->forceCounterCap) {
        HIP_TRY(hipStreamSynchronize(dev->stream));
        hipFree(dev->forceBodies);
        hipFree(dev->forceSums);
        hipFree(dev->forceCounters);
        hipHostFree(dev->forceCountsHost);
        dev->forceBodies = nullptr;
        dev->forceSums = nullptr;
        dev->forceCounters = nullptr;
        dev->forceCountsHost = nullptr;
        dev->forceBodyCap = dev->forceCounterCap = 0;
        HIP_TRY(hipMalloc(&dev->forceBodies, nb * sizeof(ArtraForceBody)));
        HIP_TRY(hipMalloc(&dev->forceSums, (size_t)nb * 12 * sizeof(double)));
        HIP_TRY(hipMalloc(&dev->forceCounters, (size_t)(1 + 2 * nb) * sizeof(int)));
        HIP_TRY(hipHostMalloc(&dev->forceCountsHost, (size_t)(1 + 2 * nb) * sizeof(int)));
        dev->forceBodyCap = nb;
        dev->forceCounterCap = 1 + 2 * nb;
    }
    // a second pass when the table outgrew its allocation
    for (int pass = 0; pass < 2; ++pass) {
        HIP_TRY(hipMemsetAsync(dev->forceCounters, 0, sizeof(int), dev->stream));
        for (int b = 0; b < nb; ++b) {
            const int n = slotBody[b];
            if (n < 0 || n >= nbody) return Fail("force body %d out of range", n);
            const ArtraGeoBody &bd = bodies[n];
            GeoBox bb;
            bb.count = 1;
            for (int s = 0; s < 3; ++s) {
                bb.lo[s] = box[b][s][0];
                bb.n[s] = box[b][s][1] > box[b][s][0] ? box[b][s][1] - box[b][s][0] : 0;
                bb.count *= bb.n[s];
            }
            GhostBody gbody;
            gbody.pv = bd.pv;
            if (bd.pv.faceN > 0) {
                const GeoBodyDev &gb = dev->geoBodies[n];
                if (!gb.uploaded) return Fail("body %d has no device arrays", n);
                gbody.pv.f = (const int (*)[3])gb.f;
                gbody.pv.e = (const int (*)[4])gb.e;
                gbody.pv.v = (const double (*)[3])gb.v;
                gbody.pv.Nf = (const double (*)[3])gb.Nf;
                gbody.pv.Ne = (const double (*)[3])gb.Ne;
                gbody.pv.Nv = (const double (*)[3])gb.Nv;
            }
            for (int s = 0; s < 3; ++s) {
                gbody.O[s] = bd.O[s];
                gbody.V[s] = bd.V[s];
                gbody.W[s] = bd.W[s];
            }
            gbody.r = bd.r;
            gbody.cf = bd.cf;
            gbody.T = bd.T;
            gbody.body = n;
            hipLaunchKernelGGL(force_build, dim3(1), dim3(kForceBuildThreads), 0, dev->stream, dev->g, dev->geo, bb,
                               gbody, b, viscous[b], did, fid, lid, gst, dev->forceNodes, dev->forceBodies,
                               dev->forceNodeCap, dev->forceCounters);
            HIP_TRY(hipGetLastError());
        }
        int *c = dev->forceCountsHost;
        HIP_TRY(hipMemcpyAsync(c, dev->forceCounters, (size_t)(1 + 2 * nb) * sizeof(int), hipMemcpyDeviceToHost,
                               dev->stream));
        HIP_TRY(hipStreamSynchronize(dev->stream));
        if (c[0] <= dev->forceNodeCap) {
            for (int b = 0; b < nb; ++b) {
                gstN[b] = c[1 + 2 * b];
                lidN[b] = c[2 + 2 * b];
            }
            dev->forceCount = c[0];
            dev->forceBodyCount = nb;
            return 0;
        }
        hipFree(dev->forceNodes);
        hipFree(dev->forceTerms);
        dev->forceNodes = nullptr;
        dev->forceTerms = nullptr;
        dev->forceNodeCap = 0;
        const int cap = c[0] + c[0] / 2;
        HIP_TRY(hipMalloc(&dev->forceNodes, (size_t)cap * sizeof(ArtraForceNode)));
        HIP_TRY(hipMalloc(&dev->forceTerms, (size_t)cap * 10 * sizeof(double)));
        dev->forceNodeCap = cap;
    }
    return Fail("surface-force table did not fit after growing");
}

}  // extern "C"
