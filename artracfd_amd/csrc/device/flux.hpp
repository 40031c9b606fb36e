// flux.hpp - per-face numerical fluxes evaluated on the device.
//
// The convective flux is the characteristic-wise, flux-vector-split WENO
// reconstruction of the solver (convective_flux.c:48-124 with the state
// algebra of cfd_commons.c:79-246 and weno5.c:35-62 / weno3.c:35-55); the
// viscous flux is its second-order central face flux (diffusive_flux.c:58-312).
// Every expression keeps the reference's operand order and the translation
// units are compiled with -ffp-contract=off, so a face flux is bit-identical
// to the CPU evaluation (up to the sign of an exact zero).
#pragma once

#include <hip/hip_runtime.h>

namespace artra {

constexpr int kComp = 5;

struct FluxModel {
    double gamma;
    double cv;
    double refMu;   // nondimensional viscosity scale (> 0 enables viscous flux)
    double refT;
    int roe;        // 1: Roe average, 0: arithmetic mean
    int sw;         // 1: Steger-Warming split, 0: local Lax-Friedrichs
};

__host__ __device__ __forceinline__ double sq(double x) { return x * x; }

// Jiang-Shu fifth-order WENO of one characteristic component from the
// five upwind-ordered values f0..f4 (value sought at the face after f2).
__host__ __device__ __forceinline__ double weno5(double f0, double f1, double f2, double f3, double f4)
{
    const double eps = 1.0e-6;
    const double IS0 = (13.0 / 12.0) * sq(f0 - 2.0 * f1 + f2) + (1.0 / 4.0) * sq(f0 - 4.0 * f1 + 3.0 * f2);
    const double IS1 = (13.0 / 12.0) * sq(f1 - 2.0 * f2 + f3) + (1.0 / 4.0) * sq(f1 - f3);
    const double IS2 = (13.0 / 12.0) * sq(f2 - 2.0 * f3 + f4) + (1.0 / 4.0) * sq(3.0 * f2 - 4.0 * f3 + f4);
    const double a0 = (1.0 / 10.0) / sq(eps + IS0);
    const double a1 = (6.0 / 10.0) / sq(eps + IS1);
    const double a2 = (3.0 / 10.0) / sq(eps + IS2);
    const double w0 = a0 / (a0 + a1 + a2);
    const double w1 = a1 / (a0 + a1 + a2);
    const double w2 = a2 / (a0 + a1 + a2);
    const double q0 = (1.0 / 6.0) * (2.0 * f0 - 7.0 * f1 + 11.0 * f2);
    const double q1 = (1.0 / 6.0) * (-f1 + 5.0 * f2 + 2.0 * f3);
    const double q2 = (1.0 / 6.0) * (2.0 * f2 + 5.0 * f3 - f4);
    return w0 * q0 + w1 * q1 + w2 * q2;
}

// Third-order WENO from f0..f2 (value sought at the face after f1).
__host__ __device__ __forceinline__ double weno3(double f0, double f1, double f2)
{
    const double eps = 1.0e-6;
    const double IS0 = sq(f1 - f0);
    const double IS1 = sq(f2 - f1);
    const double a0 = (1.0 / 3.0) / sq(eps + IS0);
    const double a1 = (2.0 / 3.0) / sq(eps + IS1);
    const double w0 = a0 / (a0 + a1);
    const double w1 = a1 / (a0 + a1);
    const double q0 = (1.0 / 2.0) * (-f0 + 3.0 * f1);
    const double q1 = (1.0 / 2.0) * (f1 + f2);
    return w0 * q0 + w1 * q1;
}

// Left eigenvector matrix of the Euler flux Jacobian along S at the face
// state (u, v, w, c); rows are characteristic fields (acoustic -, entropy /
// shear x3 in axis order, acoustic +).
template <int S>
__host__ __device__ __forceinline__ void left_vectors(double gamma, double u, double v, double w, double c, double L[5][5])
{
    const double q = 0.5 * (u * u + v * v + w * w);
    const double b = (gamma - 1.0) / (2.0 * c * c);
    const double d = 1.0 / (2.0 * c);
    if (S == 0) {
        L[0][0] = b * q + d * u;         L[0][1] = -b * u - d;            L[0][2] = -b * v;
        L[0][3] = -b * w;                L[0][4] = b;
        L[1][0] = -2.0 * b * q + 1.0;    L[1][1] = 2.0 * b * u;           L[1][2] = 2.0 * b * v;
        L[1][3] = 2.0 * b * w;           L[1][4] = -2.0 * b;
        L[2][0] = -2.0 * b * q * v;      L[2][1] = 2.0 * b * v * u;       L[2][2] = 2.0 * b * v * v + 1.0;
        L[2][3] = 2.0 * b * w * v;       L[2][4] = -2.0 * b * v;
        L[3][0] = -2.0 * b * q * w;      L[3][1] = 2.0 * b * w * u;       L[3][2] = 2.0 * b * w * v;
        L[3][3] = 2.0 * b * w * w + 1.0; L[3][4] = -2.0 * b * w;
        L[4][0] = b * q - d * u;         L[4][1] = -b * u + d;            L[4][2] = -b * v;
        L[4][3] = -b * w;                L[4][4] = b;
    } else if (S == 1) {
        L[0][0] = b * q + d * v;         L[0][1] = -b * u;                L[0][2] = -b * v - d;
        L[0][3] = -b * w;                L[0][4] = b;
        L[1][0] = -2.0 * b * q * u;      L[1][1] = 2.0 * b * u * u + 1.0; L[1][2] = 2.0 * b * v * u;
        L[1][3] = 2.0 * b * w * u;       L[1][4] = -2.0 * b * u;
        L[2][0] = -2.0 * b * q + 1.0;    L[2][1] = 2.0 * b * u;           L[2][2] = 2.0 * b * v;
        L[2][3] = 2.0 * b * w;           L[2][4] = -2.0 * b;
        L[3][0] = -2.0 * b * q * w;      L[3][1] = 2.0 * b * w * u;       L[3][2] = 2.0 * b * w * v;
        L[3][3] = 2.0 * b * w * w + 1.0; L[3][4] = -2.0 * b * w;
        L[4][0] = b * q - d * v;         L[4][1] = -b * u;                L[4][2] = -b * v + d;
        L[4][3] = -b * w;                L[4][4] = b;
    } else {
        L[0][0] = b * q + d * w;         L[0][1] = -b * u;                L[0][2] = -b * v;
        L[0][3] = -b * w - d;            L[0][4] = b;
        L[1][0] = -2.0 * b * q * u;      L[1][1] = 2.0 * b * u * u + 1.0; L[1][2] = 2.0 * b * v * u;
        L[1][3] = 2.0 * b * w * u;       L[1][4] = -2.0 * b * u;
        L[2][0] = -2.0 * b * q * v;      L[2][1] = 2.0 * b * v * u;       L[2][2] = 2.0 * b * v * v + 1.0;
        L[2][3] = 2.0 * b * w * v;       L[2][4] = -2.0 * b * v;
        L[3][0] = -2.0 * b * q + 1.0;    L[3][1] = 2.0 * b * u;           L[3][2] = 2.0 * b * v;
        L[3][3] = 2.0 * b * w;           L[3][4] = -2.0 * b;
        L[4][0] = b * q - d * w;         L[4][1] = -b * u;                L[4][2] = -b * v;
        L[4][3] = -b * w + d;            L[4][4] = b;
    }
}

// F = R (hp + hn) with R the right eigenvectors along S.  The reference
// accumulates all 25 products from 0.0; the structural zeros and ones of R
// contribute exact no-ops, so only the nonzero products are formed, in the
// same column order, starting from 0.0.
template <int S>
__host__ __device__ __forceinline__ void inverse_project(double u, double v, double w, double hT, double c, const double h[5],
                                                double F[5])
{
    const double q = 0.5 * (u * u + v * v + w * w);
    const double vel[3] = {u, v, w};
    const double un = vel[S];
    // row 0: columns 0, 1+S, 4 are 1
    F[0] = ((0.0 + h[0]) + h[1 + S]) + h[4];
    // rows 1..3 (momentum m)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const double r0 = (m == S) ? vel[m] - c : vel[m];
        const double r4 = (m == S) ? vel[m] + c : vel[m];
        double acc = 0.0 + r0 * h[0];
        // column 1+t: entropy column (t == S) carries vel[m] only when m == S; shear column t carries 1 at m == t
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            if (t == S) {
                if (m == S) acc = acc + vel[m] * h[1 + t];
            } else if (t == m) {
                acc = acc + h[1 + t];
            }
        }
        F[1 + m] = acc + r4 * h[4];
    }
    // row 4
    {
        double acc = 0.0 + (hT - un * c) * h[0];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const double r = (t == S) ? vel[S] * vel[S] - q : vel[t];
            acc = acc + r * h[1 + t];
        }
        F[4] = acc + (hT + un * c) * h[4];
    }
}

// Convective numerical flux at the face between stencil nodes K0 and K0+1,
// from the conservative states st[0..NS-1] (nodes i-2..i+3 for WENO5,
// i-1..i+2 for WENO3; NS = 6 or 4).
template <int S, int NS>
__host__ __device__ __forceinline__ void convective_flux(const FluxModel &mdl, const double st[NS][5], double F[5])
{
    constexpr int K0 = (NS == 6) ? 2 : 1;
    const double gamma = mdl.gamma;
    // face average (cfd_commons.c:79-112)
    const double *UL = st[K0];
    const double *UR = st[K0 + 1];
    const double rhoL = UL[0];
    const double uL = UL[1] / UL[0], vL = UL[2] / UL[0], wL = UL[3] / UL[0];
    const double hTL = (UL[4] / UL[0]) * gamma - 0.5 * (uL * uL + vL * vL + wL * wL) * (gamma - 1.0);
    const double rhoR = UR[0];
    const double uR = UR[1] / UR[0], vR = UR[2] / UR[0], wR = UR[3] / UR[0];
    const double hTR = (UR[4] / UR[0]) * gamma - 0.5 * (uR * uR + vR * vR + wR * wR) * (gamma - 1.0);
    double u, v, w, hT;
    if (mdl.roe) {
        const double D = sqrt(rhoR / rhoL);
        u = (uL + D * uR) / (1.0 + D);
        v = (vL + D * vR) / (1.0 + D);
        w = (wL + D * wR) / (1.0 + D);
        hT = (hTL + D * hTR) / (1.0 + D);
    } else {
        u = 0.5 * (uL + uR);
        v = 0.5 * (vL + vR);
        w = 0.5 * (wL + wR);
        hT = 0.5 * (hTL + hTR);
    }
    const double c = sqrt((gamma - 1.0) * (hT - 0.5 * (u * u + v * v + w * w)));
    const double vel[3] = {u, v, w};
    const double un = vel[S];
    // split eigenvalues (cfd_commons.c:113-148)
    const double lam[5] = {un - c, un, un, un, un + c};
    double lp[5], ln[5];
    if (mdl.sw) {
        const double eps = 1.0e-3;
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            lp[r] = 0.5 * (lam[r] + sqrt(lam[r] * lam[r] + eps * eps));
            ln[r] = 0.5 * (lam[r] - sqrt(lam[r] * lam[r] + eps * eps));
        }
    } else {
        const double star = fabs(lam[2]) + lam[4] - lam[2];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            lp[r] = 0.5 * (lam[r] + star);
            ln[r] = 0.5 * (lam[r] - star);
        }
    }
    double L[5][5];
    left_vectors<S>(gamma, u, v, w, c, L);
    // characteristic variables of the stencil, then WENO per field
    double hsum[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        double W[NS];
#pragma unroll
        for (int m = 0; m < NS; ++m) {
            double acc = 0.0;
#pragma unroll
            for (int cc = 0; cc < 5; ++cc) acc = acc + L[r][cc] * st[m][cc];
            W[m] = acc;
        }
        double hp, hn;
        if (NS == 6) {
            hp = weno5(lp[r] * W[0], lp[r] * W[1], lp[r] * W[2], lp[r] * W[3], lp[r] * W[4]);
            hn = weno5(ln[r] * W[5], ln[r] * W[4], ln[r] * W[3], ln[r] * W[2], ln[r] * W[1]);
        } else {
            hp = weno3(lp[r] * W[0], lp[r] * W[1], lp[r] * W[2]);
            hn = weno3(ln[r] * W[3], ln[r] * W[2], ln[r] * W[1]);
        }
        hsum[r] = hp + hn;
    }
    inverse_project<S>(u, v, w, hT, c, hsum, F);
}

// Viscous numerical flux at the face between node P and its +S neighbour E.
// g[a] are the states at P -/+ the first tangential axis, g[b] along the
// second, and the same for E: nb[0]=P-a, nb[1]=P+a, nb[2]=P-b, nb[3]=P+b,
// nb[4]=E-a, nb[5]=E+a, nb[6]=E-b, nb[7]=E+b.  Tangential axes: X face
// (y, z), Y face (x, z), Z face (x, y).
template <int S>
__host__ __device__ __forceinline__ void viscous_flux(const FluxModel &mdl, const double dd[3], const double P[5],
                                             const double E[5], const double nb[8][5], double F[5])
{
    constexpr int TA = (S == 0) ? 1 : 0;
    constexpr int TB = (S == 2) ? 1 : 2;
    const double cv = mdl.cv;
    const double velP[3] = {P[1] / P[0], P[2] / P[0], P[3] / P[0]};
    const double velE[3] = {E[1] / E[0], E[2] / E[0], E[3] / E[0]};
    const double TP = (P[4] - 0.5 * (P[1] * P[1] + P[2] * P[2] + P[3] * P[3]) / P[0]) / (P[0] * cv);
    const double TE = (E[4] - 0.5 * (E[1] * E[1] + E[2] * E[2] + E[3] * E[3]) / E[0]) / (E[0] * cv);
    // g[m][n] = d vel_m / d x_n
    double g[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        g[m][S] = (velE[m] - velP[m]) * dd[S];
        g[m][TA] = 0.25 * (nb[1][1 + m] / nb[1][0] + nb[5][1 + m] / nb[5][0] - nb[0][1 + m] / nb[0][0] -
                           nb[4][1 + m] / nb[4][0]) * dd[TA];
        g[m][TB] = 0.25 * (nb[3][1 + m] / nb[3][0] + nb[7][1 + m] / nb[7][0] - nb[2][1 + m] / nb[2][0] -
                           nb[6][1 + m] / nb[6][0]) * dd[TB];
    }
    const double dT = (TE - TP) * dd[S];
    const double uh = 0.5 * (velP[0] + velE[0]);
    const double vh = 0.5 * (velP[1] + velE[1]);
    const double wh = 0.5 * (velP[2] + velE[2]);
    const double Ts = 0.5 * (TP + TE) * mdl.refT;
    const double mu = mdl.refMu * (1.458e-6 * pow(Ts, 1.5) / (Ts + 110.4));
    const double heatK = mdl.gamma * mdl.cv * mu / 0.71;
    const double divV = g[0][0] + g[1][1] + g[2][2];
    F[0] = 0.0;
    if (S == 0) {
        F[1] = mu * (g[0][0] + g[0][0] - (2.0 / 3.0) * divV);
        F[2] = mu * (g[0][1] + g[1][0]);
        F[3] = mu * (g[0][2] + g[2][0]);
    } else if (S == 1) {
        F[1] = mu * (g[1][0] + g[0][1]);
        F[2] = mu * (g[1][1] + g[1][1] - (2.0 / 3.0) * divV);
        F[3] = mu * (g[1][2] + g[2][1]);
    } else {
        F[1] = mu * (g[2][0] + g[0][2]);
        F[2] = mu * (g[2][1] + g[1][2]);
        F[3] = mu * (g[2][2] + g[2][2] - (2.0 / 3.0) * divV);
    }
    F[4] = heatK * dT + F[1] * uh + F[2] * vh + F[3] * wh;
}

}  // namespace artra

// ---------------------------------------------------------------------------
// Fast variant.  Same numerical flux, evaluated with fewer FP64 instructions;
// results agree with the reference order to rounding level (checked against
// the 1e-10 relative-L2 contract on every golden case):
//  * one reciprocal per density instead of one division per velocity;
//  * the characteristic projection uses the structure of the left
//    eigenvectors: with K = q U0 - vel.mom + U4 and M = un U0 - U_{1+S},
//    W = (bK + dM, U0 - 2bK [entropy] / U_{1+t} - 2b vel_t K [shear], bK - dM),
//    12 instead of 50 operations per stencil node;
//  * WENO weights without their own divisions: with d_k = (eps + IS_k)^2,
//    omega_k = C_k prod_{j!=k} d_j / sum_l C_l prod_{j!=l} d_j, and the two
//    split sides of a field summed over one common denominator
//    (weno5_field_parts);
//  * fused multiply-adds throughout.
namespace artra {
namespace fast {

// One characteristic field of the split WENO5 flux, h = WENO5(lp W0..lp W4)
// + WENO5(ln W5..ln W1), from the field's six stencil values W0..W5 and its
// split eigenvalues, as num / den: Jiang-Shu WENO5 (weno5 above) of the
// scaled values lp W and ln W, arranged for fewer operations and one
// division per field (the two sides over a common denominator):
//  * the smoothness indicators are quadratic forms, IS(l f) = l^2 IS(f), so
//    they are formed once on W (their second differences shared between the
//    two sides) and scaled by lp^2 or ln^2;
//  * 12 IS = 13 a^2 + 3 b^2 and eps are scaled by 12 together, a common
//    factor that cancels in the weights t_k = C_k prod_{j != k} (eps + IS_j)^2
//    (C scaled by 10 likewise);
//  * the candidate values are the central node plus differences,
//    6 q_k = 6 W2 + R_k (6 W3 + R_k on the mirrored side), so
//    h = lp W2 + ln W3 + (lp N+ D- + ln N- D+) / (6 D+ D-), N = sum t_k R_k,
//    D = sum t_k, returned as num = base * den + (...).
// den >= 6 (1.2e-5)^8: no underflow; it overflows only for smoothness
// indicators far beyond any flow state (about 1e18).
__host__ __device__ __forceinline__ void weno5_field_parts(const double W[6], double lp, double ln, double &num,
                                                           double &den)
{
    const double eps12 = 12.0e-6;
    const double d0 = W[1] - W[0], d1 = W[2] - W[1], d2 = W[3] - W[2], d3 = W[4] - W[3], d4 = W[5] - W[4];
    const double D0 = d1 - d0, D1 = d2 - d1, D2 = d3 - d2, D3 = d4 - d3;   // second differences
    const double A0 = (13.0 * D0) * D0, A1 = (13.0 * D1) * D1, A2 = (13.0 * D2) * D2, A3 = (13.0 * D3) * D3;
    // first-derivative terms of the six sub-stencils: plus side (W0..W4), minus side (W5..W1)
    const double bp0 = fma(2.0, d1, D0), bp1 = d1 + d2, bp2 = fma(-2.0, d2, D2);
    const double bm0 = fma(-2.0, d3, D3), bm1 = d2 + d3, bm2 = fma(2.0, d2, D1);
    const double lp2 = lp * lp, ln2 = ln * ln;
    const double ep0 = fma(lp2, fma(3.0 * bp0, bp0, A0), eps12);
    const double ep1 = fma(lp2, fma(3.0 * bp1, bp1, A1), eps12);
    const double ep2 = fma(lp2, fma(3.0 * bp2, bp2, A2), eps12);
    const double em0 = fma(ln2, fma(3.0 * bm0, bm0, A3), eps12);
    const double em1 = fma(ln2, fma(3.0 * bm1, bm1, A2), eps12);
    const double em2 = fma(ln2, fma(3.0 * bm2, bm2, A1), eps12);
    const double gp0 = ep0 * ep0, gp1 = ep1 * ep1, gp2 = ep2 * ep2;
    const double gm0 = em0 * em0, gm1 = em1 * em1, gm2 = em2 * em2;
    const double tp0 = gp1 * gp2, tp1 = 6.0 * (gp0 * gp2), tp2 = 3.0 * (gp0 * gp1);
    const double tm0 = gm1 * gm2, tm1 = 6.0 * (gm0 * gm2), tm2 = 3.0 * (gm0 * gm1);
    // 6 q_k - 6 (central node) of each side's candidates
    const double Rp0 = fma(5.0, d1, -2.0 * d0), Rp1 = fma(2.0, d2, d1), Rp2 = fma(4.0, d2, -d3);
    const double Rm0 = fma(2.0, d4, -5.0 * d3), Rm1 = -fma(2.0, d2, d3), Rm2 = fma(-4.0, d2, d1);
    const double Np = fma(tp0, Rp0, fma(tp1, Rp1, tp2 * Rp2)), Dp = tp0 + tp1 + tp2;
    const double Nm = fma(tm0, Rm0, fma(tm1, Rm1, tm2 * Rm2)), Dm = tm0 + tm1 + tm2;
    den = 6.0 * (Dp * Dm);
    const double base = fma(lp, W[2], ln * W[3]);
    num = fma(base, den, fma(lp * Np, Dm, (ln * Nm) * Dp));
}

// Two quotients from one division: a/b and c/d as a (d q) and c (b q), q = 1/(b d).
__host__ __device__ __forceinline__ void divide_pair(double a, double b, double c, double d, double &x, double &y)
{
    const double q = 1.0 / (b * d);
    x = a * (d * q);
    y = c * (b * q);
}

__host__ __device__ __forceinline__ double weno3(double f0, double f1, double f2)
{
    const double eps = 1.0e-6;
    const double e0 = eps + (f1 - f0) * (f1 - f0), e1 = eps + (f2 - f1) * (f2 - f1);
    const double t0 = (1.0 / 3.0) * (e1 * e1), t1 = (2.0 / 3.0) * (e0 * e0);
    const double Q0 = fma(3.0, f1, -f0), Q1 = f1 + f2;
    return fma(t0, Q0, t1 * Q1) / (2.0 * (t0 + t1));
}

template <int S, int NS>
__host__ __device__ __forceinline__ void convective_flux(const FluxModel &mdl, const double st[NS][5], double F[5])
{
    constexpr int K0 = (NS == 6) ? 2 : 1;
    const double gamma = mdl.gamma;
    const double g1 = gamma - 1.0;
    const double *UL = st[K0];
    const double *UR = st[K0 + 1];
    double rL, rR;
    divide_pair(1.0, UL[0], 1.0, UR[0], rL, rR);
    const double uL = UL[1] * rL, vL = UL[2] * rL, wL = UL[3] * rL;
    const double uR = UR[1] * rR, vR = UR[2] * rR, wR = UR[3] * rR;
    const double hTL = fma(gamma, UL[4] * rL, -0.5 * g1 * fma(uL, uL, fma(vL, vL, wL * wL)));
    const double hTR = fma(gamma, UR[4] * rR, -0.5 * g1 * fma(uR, uR, fma(vR, vR, wR * wR)));
    double u, v, w, hT;
    if (mdl.roe) {
        const double D = sqrt(UR[0] * rL);
        const double inv = 1.0 / (1.0 + D);
        u = fma(D, uR, uL) * inv;
        v = fma(D, vR, vL) * inv;
        w = fma(D, wR, wL) * inv;
        hT = fma(D, hTR, hTL) * inv;
    } else {
        u = 0.5 * (uL + uR);
        v = 0.5 * (vL + vR);
        w = 0.5 * (wL + wR);
        hT = 0.5 * (hTL + hTR);
    }
    const double q = 0.5 * fma(u, u, fma(v, v, w * w));
    const double c = sqrt(g1 * (hT - q));
    const double rc = 1.0 / c;
    const double d = 0.5 * rc;
    const double b = 0.5 * g1 * rc * rc;
    const double vel[3] = {u, v, w};
    const double un = vel[S];
    const double lam[5] = {un - c, un, un, un, un + c};
    double lp[5], ln[5];
    if (mdl.sw) {
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const double root = sqrt(fma(lam[r], lam[r], 1.0e-6));
            lp[r] = 0.5 * (lam[r] + root);
            ln[r] = 0.5 * (lam[r] - root);
        }
    } else {
        const double star = fabs(un) + c;
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            lp[r] = 0.5 * (lam[r] + star);
            ln[r] = 0.5 * (lam[r] - star);
        }
    }
    // characteristic variables: per stencil node K = q U0 - vel.mom + U4 and
    // M = un U0 - U_{1+S}; the fields are formed one at a time from them
    // (acoustic fields 0 and 4, then the entropy and shear fields), so that
    // only one field's stencil values are live at once
    const double m2b = -2.0 * b;
    double sh[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) sh[t] = m2b * vel[t];
    double K[NS], M[NS];
#pragma unroll
    for (int m = 0; m < NS; ++m) {
        const double *U = st[m];
        K[m] = fma(q, U[0], U[4]) - fma(u, U[1], fma(v, U[2], w * U[3]));
        M[m] = fma(un, U[0], -U[1 + S]);
    }
    double h[5];
    if (NS == 6) {
        double num[5], den[5];
        {
            double wa[6], wb[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                const double bK = b * K[m], dM = d * M[m];
                wa[m] = bK + dM;
                wb[m] = bK - dM;
            }
            weno5_field_parts(wa, lp[0], ln[0], num[0], den[0]);
            weno5_field_parts(wb, lp[4], ln[4], num[4], den[4]);
        }
        divide_pair(num[0], den[0], num[4], den[4], h[0], h[4]);
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            double wf[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) wf[m] = (t == S) ? fma(m2b, K[m], st[m][0]) : fma(sh[t], K[m], st[m][1 + t]);
            weno5_field_parts(wf, lp[1 + t], ln[1 + t], num[1 + t], den[1 + t]);
        }
        divide_pair(num[1], den[1], num[2], den[2], h[1], h[2]);
        h[3] = num[3] / den[3];
    } else {
        double W[5][NS];
#pragma unroll
        for (int m = 0; m < NS; ++m) {
            const double bK = b * K[m], dM = d * M[m];
            W[0][m] = bK + dM;
            W[4][m] = bK - dM;
#pragma unroll
            for (int t = 0; t < 3; ++t) W[1 + t][m] = (t == S) ? fma(m2b, K[m], st[m][0]) : fma(sh[t], K[m], st[m][1 + t]);
        }
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const double hp = weno3(lp[r] * W[r][0], lp[r] * W[r][1], lp[r] * W[r][2]);
            const double hn = weno3(ln[r] * W[r][3], ln[r] * W[r][2], ln[r] * W[r][1]);
            h[r] = hp + hn;
        }
    }
    // F = R h with the right eigenvectors along S
    const double ha = h[0] + h[4], hd = h[4] - h[0];
    F[0] = ha + h[1 + S];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        double acc = fma(vel[m], ha, (m == S) ? c * hd : 0.0);
        acc = (m == S) ? fma(vel[m], h[1 + S], acc) : acc + h[1 + m];
        F[1 + m] = acc;
    }
    double e = fma(hT, ha, un * c * hd);
#pragma unroll
    for (int t = 0; t < 3; ++t) e = fma((t == S) ? fma(un, un, -q) : vel[t], h[1 + t], e);
    F[4] = e;
}

template <int S>
__host__ __device__ __forceinline__ void viscous_flux(const FluxModel &mdl, const double dd[3], const double P[5],
                                             const double E[5], const double nb[8][5], double F[5])
{
    constexpr int TA = (S == 0) ? 1 : 0;
    constexpr int TB = (S == 2) ? 1 : 2;
    const double rP = 1.0 / P[0], rE = 1.0 / E[0];
    const double velP[3] = {P[1] * rP, P[2] * rP, P[3] * rP};
    const double velE[3] = {E[1] * rE, E[2] * rE, E[3] * rE};
    const double rcv = 1.0 / mdl.cv;
    const double TP = (P[4] * rP - 0.5 * fma(velP[0], velP[0], fma(velP[1], velP[1], velP[2] * velP[2]))) * rcv;
    const double TE = (E[4] * rE - 0.5 * fma(velE[0], velE[0], fma(velE[1], velE[1], velE[2] * velE[2]))) * rcv;
    double rn[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) rn[q] = 1.0 / nb[q][0];
    double g[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        g[m][S] = (velE[m] - velP[m]) * dd[S];
        g[m][TA] = 0.25 * dd[TA] *
                   (fma(nb[1][1 + m], rn[1], nb[5][1 + m] * rn[5]) - fma(nb[0][1 + m], rn[0], nb[4][1 + m] * rn[4]));
        g[m][TB] = 0.25 * dd[TB] *
                   (fma(nb[3][1 + m], rn[3], nb[7][1 + m] * rn[7]) - fma(nb[2][1 + m], rn[2], nb[6][1 + m] * rn[6]));
    }
    const double Ts = 0.5 * (TP + TE) * mdl.refT;
    const double mu = mdl.refMu * (1.458e-6 * Ts * sqrt(Ts) / (Ts + 110.4));
    const double heatK = mdl.gamma * mdl.cv * (1.0 / 0.71) * mu;
    const double divV = g[0][0] + g[1][1] + g[2][2];
    F[0] = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        F[1 + m] = (m == S) ? mu * fma(2.0, g[S][S], (-2.0 / 3.0) * divV) : mu * (g[m][S] + g[S][m]);
    }
    const double uh = 0.5 * (velP[0] + velE[0]);
    const double vh = 0.5 * (velP[1] + velE[1]);
    const double wh = 0.5 * (velP[2] + velE[2]);
    F[4] = fma(heatK, (TE - TP) * dd[S], fma(F[1], uh, fma(F[2], vh, F[3] * wh)));
}

}  // namespace fast
}  // namespace artra
