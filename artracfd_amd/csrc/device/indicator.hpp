// indicator.hpp - learned troubled-face indicator: features, labels, the
// 12-32-32-1 MLP on BF16 MFMA with FP32 accumulation, and the linear
// fifth-order face flux used where the indicator says the data are smooth.
//
// Features of a face (12 FP32 values): the six-node stencils (i-2 .. i+3)
// of density and pressure, each shifted by its face average
// c = (q_i + q_{i+1}) / 2 and divided by  max_m |q_m - c| + 0.01 |c|.
// The 0.01 |c| floor keeps variations small against the local level small
// (smooth), while large relative variations are seen by their shape only.
//
// Label of a face (training only): the largest deviation of the WENO5
// nonlinear weights from the linear ones, max over the five characteristic
// fields and both split sides of sum_k |omega_k - C_k|, computed on the
// solver's own data with the flux's own arithmetic.
//
// MLP layout: W1 is 16 x 32 (rows 12..15 zero), W2 32 x 32, row-major [in][out],
// BF16; biases and the output vector FP32; hidden layers use ReLU; the
// output is a logit.  A face is "troubled" (full WENO5) when its logit
// exceeds a threshold; otherwise the linear upwind flux is used.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "flux.hpp"

namespace artra {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kFeat = 12;
constexpr int kFeatPad = 16;
constexpr int kHidden = 32;

__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
    // round to nearest even (finite inputs)
    union { float f; uint32_t u; } v;
    v.f = f;
    const uint32_t rounding = 0x7fffu + ((v.u >> 16) & 1u);
    return (uint16_t)((v.u + rounding) >> 16);
}

__host__ __device__ __forceinline__ float bf16_to_f32(uint16_t h)
{
    union { float f; uint32_t u; } v;
    v.u = (uint32_t)h << 16;
    return v.f;
}

// FP32 density and pressure of one node (conservative state in FP64).
__device__ __forceinline__ void node_rho_p(double gamma, const double U[5], float &rho, float &p)
{
    const float g1 = (float)(gamma - 1.0);
    const float r = (float)U[0];
    const float mx = (float)U[1], my = (float)U[2], mz = (float)U[3];
    const float inv = 1.0f / r;
    rho = r;
    p = g1 * ((float)U[4] - 0.5f * (mx * mx + my * my + mz * mz) * inv);
}

// Features of a face from the six nodes' density and pressure (the
// normalising reciprocal is the hardware one, within 1 ulp).
__device__ __forceinline__ void features_from(const float rho[6], const float p[6], float x[kFeat])
{
    const float *src[2] = {rho, p};
#pragma unroll
    for (int v = 0; v < 2; ++v) {
        const float *q = src[v];
        const float c = 0.5f * (q[2] + q[3]);
        float dev = 0.0f;
#pragma unroll
        for (int m = 0; m < 6; ++m) dev = fmaxf(dev, fabsf(q[m] - c));
        const float inv = __builtin_amdgcn_rcpf(dev + 0.01f * fabsf(c) + 1.0e-30f);
#pragma unroll
        for (int m = 0; m < 6; ++m) x[6 * v + m] = (q[m] - c) * inv;
    }
}

// Whether the six-node density and pressure stencils of a face vary by at
// most `variation` of their face averages; such faces can be routed to the
// linear flux without consulting the network (at 1e-3 their features are
// all below 0.1 in magnitude, where WENO5 and the linear flux differ by far
// less than the routing tolerance).
__device__ __forceinline__ bool near_uniform(const float rho[6], const float p[6], float variation)
{
    const float cr = 0.5f * (rho[2] + rho[3]), cp = 0.5f * (p[2] + p[3]);
    float dr = 0.0f, dp = 0.0f;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        dr = fmaxf(dr, fabsf(rho[m] - cr));
        dp = fmaxf(dp, fabsf(p[m] - cp));
    }
    return dr <= variation * fabsf(cr) && dp <= variation * fabsf(cp);
}

// Face features from the FP64 stencil states (conservative).
template <int NS>
__device__ __forceinline__ void face_features(double gamma, const double st[NS][5], float x[kFeat])
{
    float rho[6], p[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const int q = (NS == 6) ? m : (m == 0 ? 0 : (m == 5 ? NS - 1 : m - 1));
        node_rho_p(gamma, st[q], rho[m], p[m]);
    }
    features_from(rho, p, x);
}

// sum_k |omega_k - C_k| of one WENO5 evaluation (Jiang-Shu weights, C = 0.1, 0.6, 0.3).
__host__ __device__ __forceinline__ double weno5_weight_deviation(double f0, double f1, double f2, double f3, double f4)
{
    const double eps = 1.0e-6;
    const double a0 = fma(-2.0, f1, f0 + f2), b0 = fma(3.0, f2, fma(-4.0, f1, f0));
    const double a1 = fma(-2.0, f2, f1 + f3), b1 = f1 - f3;
    const double a2 = fma(-2.0, f3, f2 + f4), b2 = fma(3.0, f2, fma(-4.0, f3, f4));
    const double e0 = eps + fma(13.0 / 12.0 * a0, a0, 0.25 * b0 * b0);
    const double e1 = eps + fma(13.0 / 12.0 * a1, a1, 0.25 * b1 * b1);
    const double e2 = eps + fma(13.0 / 12.0 * a2, a2, 0.25 * b2 * b2);
    const double d0 = e0 * e0, d1 = e1 * e1, d2 = e2 * e2;
    const double t0 = 0.1 * (d1 * d2), t1 = 0.6 * (d0 * d2), t2 = 0.3 * (d0 * d1);
    const double inv = 1.0 / (t0 + t1 + t2);
    return fabs(t0 * inv - 0.1) + fabs(t1 * inv - 0.6) + fabs(t2 * inv - 0.3);
}

// Face state shared by the WENO and linear fluxes (fast arithmetic).
struct FaceState {
    double u, v, w, hT, c, q, b, d, un, g1;
    double lp[5], ln[5];
};

// b = (gamma - 1) / (2 c^2) and d = 1 / (2 c) of the eigenvectors; only the
// characteristic projections need them.
__host__ __device__ __forceinline__ void face_scales(FaceState &fs)
{
    const double rc = 1.0 / fs.c;
    fs.d = 0.5 * rc;
    fs.b = 0.5 * fs.g1 * rc * rc;
}

template <int S>
__host__ __device__ __forceinline__ void face_state(const FluxModel &mdl, const double UL[5], const double UR[5], FaceState &fs)
{
    const double gamma = mdl.gamma, g1 = gamma - 1.0;
    double rL, rR;
    fast::divide_pair(1.0, UL[0], 1.0, UR[0], rL, rR);
    const double uL = UL[1] * rL, vL = UL[2] * rL, wL = UL[3] * rL;
    const double uR = UR[1] * rR, vR = UR[2] * rR, wR = UR[3] * rR;
    const double hTL = fma(gamma, UL[4] * rL, -0.5 * g1 * fma(uL, uL, fma(vL, vL, wL * wL)));
    const double hTR = fma(gamma, UR[4] * rR, -0.5 * g1 * fma(uR, uR, fma(vR, vR, wR * wR)));
    if (mdl.roe) {
        const double D = sqrt(UR[0] * rL);
        const double inv = 1.0 / (1.0 + D);
        fs.u = fma(D, uR, uL) * inv;
        fs.v = fma(D, vR, vL) * inv;
        fs.w = fma(D, wR, wL) * inv;
        fs.hT = fma(D, hTR, hTL) * inv;
    } else {
        fs.u = 0.5 * (uL + uR);
        fs.v = 0.5 * (vL + vR);
        fs.w = 0.5 * (wL + wR);
        fs.hT = 0.5 * (hTL + hTR);
    }
    fs.q = 0.5 * fma(fs.u, fs.u, fma(fs.v, fs.v, fs.w * fs.w));
    fs.c = sqrt(g1 * (fs.hT - fs.q));
    fs.g1 = g1;
    fs.un = (S == 0) ? fs.u : (S == 1) ? fs.v : fs.w;
    const double lam[5] = {fs.un - fs.c, fs.un, fs.un, fs.un, fs.un + fs.c};
    if (mdl.sw) {
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const double root = sqrt(fma(lam[r], lam[r], 1.0e-6));
            fs.lp[r] = 0.5 * (lam[r] + root);
            fs.ln[r] = 0.5 * (lam[r] - root);
        }
    } else {
        const double star = fabs(fs.un) + fs.c;
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            fs.lp[r] = 0.5 * (lam[r] + star);
            fs.ln[r] = 0.5 * (lam[r] - star);
        }
    }
}

// Characteristic variables W = L U of one state (structured form).
template <int S>
__host__ __device__ __forceinline__ void characteristic(const FaceState &fs, const double U[5], double W[5])
{
    const double vel[3] = {fs.u, fs.v, fs.w};
    const double K = fma(fs.q, U[0], U[4]) - fma(fs.u, U[1], fma(fs.v, U[2], fs.w * U[3]));
    const double M = fma(fs.un, U[0], -U[1 + S]);
    const double bK = fs.b * K, dM = fs.d * M;
    const double m2b = -2.0 * fs.b;
    W[0] = bK + dM;
    W[4] = bK - dM;
#pragma unroll
    for (int t = 0; t < 3; ++t) W[1 + t] = (t == S) ? fma(m2b, K, U[0]) : fma(m2b * vel[t], K, U[1 + t]);
}

// F = R h (right eigenvectors along S).
template <int S>
__host__ __device__ __forceinline__ void project_back(const FaceState &fs, const double h[5], double F[5])
{
    const double vel[3] = {fs.u, fs.v, fs.w};
    const double ha = h[0] + h[4], hd = h[4] - h[0];
    F[0] = ha + h[1 + S];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        double acc = fma(vel[m], ha, (m == S) ? fs.c * hd : 0.0);
        acc = (m == S) ? fma(vel[m], h[1 + S], acc) : acc + h[1 + m];
        F[1 + m] = acc;
    }
    double e = fma(fs.hT, ha, fs.un * fs.c * hd);
#pragma unroll
    for (int t = 0; t < 3; ++t) e = fma((t == S) ? fma(fs.un, fs.un, -fs.q) : vel[t], h[1 + t], e);
    F[4] = e;
}

// Linear fifth-order upwind flux: the WENO5 flux with every nonlinear weight
// replaced by its linear value; linearity moves the stencil combination in
// front of the projection: h = lp * L (sum c+ U) + ln * L (sum c- U).
// linear_averages forms the two combinations (sum c+ U, sum c- U) from the
// six-node stencil, linear_flux_from projects them.
__host__ __device__ __forceinline__ void linear_averages(const double st[6][5], double Up[5], double Un[5])
{
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        Up[c] = fma(2.0, st[0][c], fma(-13.0, st[1][c], fma(47.0, st[2][c], fma(27.0, st[3][c], -3.0 * st[4][c])))) *
                (1.0 / 60.0);
        Un[c] = fma(2.0, st[5][c], fma(-13.0, st[4][c], fma(47.0, st[3][c], fma(27.0, st[2][c], -3.0 * st[1][c])))) *
                (1.0 / 60.0);
    }
}

template <int S>
__host__ __device__ __forceinline__ void linear_flux_from(FaceState fs, const double Up[5], const double Un[5], double F[5])
{
    face_scales(fs);
    double Wp[5], Wn[5], h[5];
    characteristic<S>(fs, Up, Wp);
    characteristic<S>(fs, Un, Wn);
#pragma unroll
    for (int r = 0; r < 5; ++r) h[r] = fma(fs.lp[r], Wp[r], fs.ln[r] * Wn[r]);
    project_back<S>(fs, h, F);
}

// Face state of the arithmetic mean from per-node velocities and total
// enthalpies (precomputed once per node); the sound speed uses the hardware
// square root, which is ample for the linear flux's dissipation coefficient.
template <int S>
__device__ __forceinline__ void face_state_mean(double gamma, const double nL[4], const double nR[4], FaceState &fs)
{
    fs.u = 0.5 * (nL[0] + nR[0]);
    fs.v = 0.5 * (nL[1] + nR[1]);
    fs.w = 0.5 * (nL[2] + nR[2]);
    fs.hT = 0.5 * (nL[3] + nR[3]);
    fs.q = 0.5 * fma(fs.u, fs.u, fma(fs.v, fs.v, fs.w * fs.w));
    fs.g1 = gamma - 1.0;
    fs.c = __builtin_amdgcn_sqrt(fs.g1 * (fs.hT - fs.q));
    fs.un = (S == 0) ? fs.u : (S == 1) ? fs.v : fs.w;
}

// Velocities and total enthalpy of one node (as face_state computes them per side).
__device__ __forceinline__ void node_velocity_enthalpy(double gamma, const double U[5], double nv[4])
{
    const double r = 1.0 / U[0];
    nv[0] = U[1] * r;
    nv[1] = U[2] * r;
    nv[2] = U[3] * r;
    nv[3] = fma(gamma, U[4] * r, -0.5 * (gamma - 1.0) * fma(nv[0], nv[0], fma(nv[1], nv[1], nv[2] * nv[2])));
}

// Sum and difference of the two linear combinations, sum c+ U + sum c- U and
// sum c+ U - sum c- U, of one component's six stencil values in symmetric
// form.
__host__ __device__ __forceinline__ void linear_sum_diff_1(const double w[6], double &Sg, double &Dl)
{
    const double a05 = w[0] + w[5], a14 = w[1] + w[4], a23 = w[2] + w[3];
    const double d05 = w[0] - w[5], d14 = w[1] - w[4], d23 = w[2] - w[3];
    Sg = fma(2.0, a05, fma(-16.0, a14, 74.0 * a23)) * (1.0 / 60.0);
    Dl = fma(2.0, d05, fma(-10.0, d14, 20.0 * d23)) * (1.0 / 60.0);
}

// The same for every component of the six-node stencil st[node][component].
__host__ __device__ __forceinline__ void linear_sum_diff(const double st[6][5], double Sg[5], double Dl[5])
{
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const double w[6] = {st[0][c], st[1][c], st[2][c], st[3][c], st[4][c], st[5][c]};
        linear_sum_diff_1(w, Sg[c], Dl[c]);
    }
}

// The same linear flux for the local Lax-Friedrichs split, where
// lp/ln = (lam +/- lam*) / 2 with one lam* = |u_s| + c for every field:
// R diag(lp) L Up + R diag(ln) L Um = A (Up + Um) / 2 + lam* (Up - Um) / 2,
// A = R diag(lam) L the Euler flux Jacobian at the face state, applied in
// closed form (P is the linearised pressure of the vector it acts on):
//   (A V)_0 = V_{1+s},  (A V)_{1+t} = u_t V_{1+s} + u_s (V_{1+t} - u_t V_0) + [t = s] P,
//   (A V)_4 = H (V_{1+s} - u_s V_0) + u_s (P + V_4),
//   P = (gamma - 1) (q V_0 - u V_1 - v V_2 - w V_3 + V_4).
template <int S>
__host__ __device__ __forceinline__ void linear_flux_llf_sd(double gamma, const FaceState &fs, const double Sg[5],
                                                   const double Dl[5], double F[5])
{
    const double vel[3] = {fs.u, fs.v, fs.w};
    const double un = fs.un, m = Sg[1 + S];
    const double P = (gamma - 1.0) * (fma(fs.q, Sg[0], Sg[4]) - fma(fs.u, Sg[1], fma(fs.v, Sg[2], fs.w * Sg[3])));
    const double star = 0.5 * (fabs(un) + fs.c);
    double A[5];
    A[0] = m;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const double a = fma(vel[t], m, un * fma(-vel[t], Sg[0], Sg[1 + t]));
        A[1 + t] = (t == S) ? a + P : a;
    }
    A[4] = fma(fs.hT, fma(-un, Sg[0], m), un * (P + Sg[4]));
#pragma unroll
    for (int c = 0; c < 5; ++c) F[c] = fma(0.5, A[c], star * Dl[c]);
}

template <int S>
__host__ __device__ __forceinline__ void linear_flux_llf(double gamma, const FaceState &fs, const double Up[5], const double Um[5],
                                                double F[5])
{
    double Sg[5], Dl[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        Sg[c] = Up[c] + Um[c];
        Dl[c] = Up[c] - Um[c];
    }
    linear_flux_llf_sd<S>(gamma, fs, Sg, Dl, F);
}

template <int S>
__host__ __device__ __forceinline__ void linear_flux(const FaceState &fs, const double st[6][5], double F[5])
{
    double Up[5], Un[5];
    linear_averages(st, Up, Un);
    linear_flux_from<S>(fs, Up, Un, F);
}

// The WENO5 sum of one characteristic field of the face flux, h_r =
// WENO5(lp W_r) over the nodes f-2 .. f+2 plus WENO5(ln W_r) over the
// mirrored nodes f+3 .. f-1 (for a line held node-major, [n][5]), as a
// numerator and denominator (fs needs face_scales applied).
template <int S>
__device__ __forceinline__ void weno_field_parts(const FaceState &fs, const double *st, int f, int r, bool sw,
                                                 double &num, double &den)
{
    const double vel[3] = {fs.u, fs.v, fs.w};
    const double m2b = -2.0 * fs.b;
    const bool edge = (r == 0) || (r == 4);
    const double sgn = (r == 0) ? 1.0 : -1.0;
    const int t = edge ? 0 : r - 1;
    const double cK = (t == S) ? m2b : m2b * ((t == 0) ? vel[0] : (t == 1) ? vel[1] : vel[2]);
    const int idx = (t == S) ? 0 : 1 + t;
    const double lam = (r == 0) ? fs.un - fs.c : (r == 4) ? fs.un + fs.c : fs.un;
    double lp, ln;
    if (sw) {
        const double root = sqrt(fma(lam, lam, 1.0e-6));
        lp = 0.5 * (lam + root);
        ln = 0.5 * (lam - root);
    } else {
        const double star = fabs(fs.un) + fs.c;
        lp = 0.5 * (lam + star);
        ln = 0.5 * (lam - star);
    }
    double w[6];                                    // the field's value at nodes f-2 .. f+3
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const int node = f - 2 + m;
        const double *Un = st + node * 5;
        const double U0 = Un[0], U1 = Un[1], U2 = Un[2], U3 = Un[3], U4 = Un[4];
        const double K = fma(fs.q, U0, U4) - fma(fs.u, U1, fma(fs.v, U2, fs.w * U3));
        const double M = fma(fs.un, U0, -Un[1 + S]);
        const double bK = fs.b * K, dM = fs.d * M;
        const double Ui = Un[idx];
        w[m] = edge ? bK + sgn * dM : fma(cK, K, Ui);
    }
    fast::weno5_field_parts(w, lp, ln, num, den);
}

// The field sums of a face from their numerators and denominators
// num/den[r] = parts[2 r], parts[2 r + 1], with the divisions
// fast::convective_flux makes (fields 0 and 4, and 1 and 2, share one), so
// that h is bit-identical to it.
__host__ __device__ __forceinline__ void fields_from_parts(const double *parts, double h[5])
{
    fast::divide_pair(parts[0], parts[1], parts[8], parts[9], h[0], h[4]);
    fast::divide_pair(parts[2], parts[3], parts[4], parts[5], h[1], h[2]);
    h[3] = parts[6] / parts[7];
}

// The fast WENO5 flux of the face between nodes f and f + 1 of a line held
// node-major, [n][5] (in LDS), bit-identical to fast::convective_flux on the same
// six nodes.  The fields are formed one at a time from the per-node
// projections K and M (acoustic fields 0 and 4 first, then the entropy and
// shear fields, whose node components are read again), so that only one
// field's stencil values are live: the sweep then fits in 128 registers.
template <int S>
__device__ __forceinline__ void weno_flux_line(const FluxModel &mdl, const double *st, int f, double F[5])
{
    FaceState fs;
    {
        double UL[5], UR[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            UL[c] = st[f * 5 + c];
            UR[c] = st[(f + 1) * 5 + c];
        }
        face_state<S>(mdl, UL, UR, fs);
    }
    face_scales(fs);
    double K[6], M[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const int node = f - 2 + m;
        const double *Un = st + node * 5;
        const double U0 = Un[0], U1 = Un[1], U2 = Un[2], U3 = Un[3], U4 = Un[4];
        K[m] = fma(fs.q, U0, U4) - fma(fs.u, U1, fma(fs.v, U2, fs.w * U3));
        M[m] = fma(fs.un, U0, -Un[1 + S]);
    }
    double parts[10];
    {
        double wa[6], wb[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const double bK = fs.b * K[m], dM = fs.d * M[m];
            wa[m] = bK + dM;
            wb[m] = bK - dM;
        }
        fast::weno5_field_parts(wa, fs.lp[0], fs.ln[0], parts[0], parts[1]);
        fast::weno5_field_parts(wb, fs.lp[4], fs.ln[4], parts[8], parts[9]);
    }
    const double vel[3] = {fs.u, fs.v, fs.w};
    const double m2b = -2.0 * fs.b;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const double cK = (t == S) ? m2b : m2b * vel[t];
        const int idx = (t == S) ? 0 : 1 + t;
        double w[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) w[m] = fma(cK, K[m], st[(f - 2 + m) * 5 + idx]);
        fast::weno5_field_parts(w, fs.lp[1 + t], fs.ln[1 + t], parts[2 + 2 * t], parts[3 + 2 * t]);
    }
    double h[5];
    fields_from_parts(parts, h);
    project_back<S>(fs, h, F);
}

// Full WENO5 flux from the same face state (matches fast::convective_flux).
// With label != nullptr also returns the weight-deviation label.
template <int S>
__host__ __device__ __forceinline__ void weno_flux(FaceState fs, const double st[6][5], double F[5], double *label)
{
    face_scales(fs);
    double W[5][6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        double t[5];
        characteristic<S>(fs, st[m], t);
#pragma unroll
        for (int r = 0; r < 5; ++r) W[r][m] = t[r];
    }
    double h[5], num[5], den[5];
    double dev = 0.0;
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        const double lp = fs.lp[r], ln = fs.ln[r];
        fast::weno5_field_parts(W[r], lp, ln, num[r], den[r]);
        if (label) {
            dev = fmax(dev, weno5_weight_deviation(lp * W[r][0], lp * W[r][1], lp * W[r][2], lp * W[r][3],
                                                   lp * W[r][4]));
            dev = fmax(dev, weno5_weight_deviation(ln * W[r][5], ln * W[r][4], ln * W[r][3], ln * W[r][2],
                                                   ln * W[r][1]));
        }
    }
    if (label) *label = dev;
    double parts[10];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        parts[2 * r] = num[r];
        parts[2 * r + 1] = den[r];
    }
    fields_from_parts(parts, h);
    project_back<S>(fs, h, F);
}

// ------------------------------------------------------------ the MLP
struct MlpDev {
    const uint16_t *W1;   // [16][32] bf16
    const uint16_t *W2;   // [32][32] bf16
    const float *b1;      // [32]
    const float *b2;      // [32]
    const float *w3;      // [32]
    float b3;
    float threshold;      // logit above which a face is troubled
};

// Per-lane B fragments and biases (column j = lane % 32).
struct MlpFrags {
    bf16x4 w1[2];
    bf16x4 w2[4];
    float b1, b2, w3;
};

__device__ __forceinline__ void mlp_fragments(const MlpDev &net, MlpFrags &f)
{
    const int l = threadIdx.x & 63;
    const int j = l & 31, kh = 4 * (l >> 5);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int t = 0; t < 4; ++t) f.w1[ks][t] = (short)net.W1[(8 * ks + kh + t) * kHidden + j];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int t = 0; t < 4; ++t) f.w2[ks][t] = (short)net.W2[(8 * ks + kh + t) * kHidden + j];
    f.b1 = net.b1[j];
    f.b2 = net.b2[j];
    f.w3 = net.w3[j];
}

// Inference on one wavefront (64 faces, lane l holds face l), computed
// transposed, H^T = W^T X^T, with v_mfma_f32_32x32x16_bf16 (K = 16 per
// instruction): lane (j, half) of a 32x32 result holds the hidden units
// 8 g + 4 half + r (g, r = 0..3) of sample j.  Layer 2 consumes them in
// place: its K index is permuted so that K-slice ks of lane (j, half) is
// exactly the accumulator registers 8 ks .. 8 ks + 7, i.e. hidden units
// 16 ks + 8 (t / 4) + 4 half + t % 4 (t = 0..7); the W2 fragments use the
// same permutation.  The two 32-sample column blocks are faces 0..31 and
// 32..63; 6 MFMAs per 64 faces, activations never leave registers.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

struct MlpWave {
    bf16x8 w1t;        // A[h][8 half + t] = W1[8 half + t][h]
    bf16x8 w2t[2];     // A[h][K(ks, half, t)] = W2[16 ks + 8 (t / 4) + 4 half + t % 4][h]
};

__device__ __forceinline__ __bf16 bf16_of(uint16_t bits)
{
    return __builtin_bit_cast(__bf16, bits);
}

__device__ __forceinline__ void mlp_wave_fragments(const uint16_t *W1, const uint16_t *W2, MlpWave &f)
{
    const int l = __lane_id();
    const int h = l & 31, half = l >> 5;
#pragma unroll
    for (int t = 0; t < 8; ++t) f.w1t[t] = bf16_of(W1[(8 * half + t) * kHidden + h]);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int unit = 16 * ks + 8 * (t >> 2) + 4 * half + (t & 3);
            f.w2t[ks][t] = bf16_of(W2[unit * kHidden + h]);
        }
}

// max(x, 0) as a signed-integer maximum of the bit pattern: non-negative
// floats order like their patterns, negative ones have negative patterns.
// One instruction, where the float maximum of an MFMA result costs two
// (the compiler first quiets a possible signalling NaN).
__device__ __forceinline__ float relu(float x)
{
    return __builtin_bit_cast(float, max(__builtin_bit_cast(int, x), 0));
}

// Two FP32 values to packed BF16 (round to nearest even, v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t pack_bf16(float a, float b)
{
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

__device__ __forceinline__ bf16x8 frag8(uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = {a, b, c, d};
    return __builtin_bit_cast(bf16x8, v);
}

// Logit of this lane's face.  Every lane of the wavefront must call it with
// the whole wavefront active.  bias points to 96 FP32 values (b1[32], b2[32],
// w3[32]) that every lane can read (LDS or global).
__device__ __forceinline__ float mlp_logit_wave(const float x[kFeat], const MlpWave &f, const float *bias, float b3)
{
    const int l = __lane_id();
    const bool upper = l >= 32;
    const uint32_t p0 = pack_bf16(x[0], x[1]), p1 = pack_bf16(x[2], x[3]);
    const uint32_t p2 = pack_bf16(x[4], x[5]), p3 = pack_bf16(x[6], x[7]);
    const uint32_t p4 = pack_bf16(x[8], x[9]), p5 = pack_bf16(x[10], x[11]);
    // lower lanes need features 0..7 of their upper partner, upper lanes features 8..15 of their lower partner
    const uint32_t s0 = upper ? p0 : p4, s1 = upper ? p1 : p5, s2 = upper ? p2 : 0u, s3 = upper ? p3 : 0u;
    const uint32_t r0 = (uint32_t)__shfl_xor((int)s0, 32), r1 = (uint32_t)__shfl_xor((int)s1, 32);
    const uint32_t r2 = (uint32_t)__shfl_xor((int)s2, 32), r3 = (uint32_t)__shfl_xor((int)s3, 32);
    bf16x8 xb[2];
    if (!upper) {
        xb[0] = frag8(p0, p1, p2, p3);
        xb[1] = frag8(r0, r1, r2, r3);
    } else {
        xb[0] = frag8(r0, r1, 0u, 0u);
        xb[1] = frag8(p4, p5, 0u, 0u);
    }
    const int kh = upper ? 4 : 0;
    float part[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        floatx16 acc;
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[4 * g + t] = bias[8 * g + kh + t];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.w1t, xb[c], acc, 0, 0, 0);
        bf16x8 hb[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
            hb[ks] = frag8(pack_bf16(relu(acc[8 * ks + 0]), relu(acc[8 * ks + 1])),
                           pack_bf16(relu(acc[8 * ks + 2]), relu(acc[8 * ks + 3])),
                           pack_bf16(relu(acc[8 * ks + 4]), relu(acc[8 * ks + 5])),
                           pack_bf16(relu(acc[8 * ks + 6]), relu(acc[8 * ks + 7])));
        floatx16 acc2;
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc2[4 * g + t] = bias[32 + 8 * g + kh + t];
        acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.w2t[0], hb[0], acc2, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.w2t[1], hb[1], acc2, 0, 0, 0);
        float s = 0.0f;
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int t = 0; t < 4; ++t) s = fmaf(relu(acc2[4 * g + t]), bias[64 + 8 * g + kh + t], s);
        part[c] = s;
    }
    // sample j of block c: halves in lanes j and j + 32
    const float mine = upper ? part[1] : part[0];
    const float other = __shfl_xor(upper ? part[0] : part[1], 32);
    return b3 + (upper ? other + mine : mine + other);
}

}  // namespace artra
