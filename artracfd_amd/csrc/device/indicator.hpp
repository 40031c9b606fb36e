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

// Face features from the FP64 stencil states (conservative).
template <int NS>
__host__ __device__ __forceinline__ void face_features(double gamma, const double st[NS][5], float x[kFeat])
{
    float rho[6], p[6];
    const float g1 = (float)(gamma - 1.0);
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const int q = (NS == 6) ? m : (m == 0 ? 0 : (m == 5 ? NS - 1 : m - 1));
        const float r = (float)st[q][0];
        const float mx = (float)st[q][1], my = (float)st[q][2], mz = (float)st[q][3];
        const float inv = 1.0f / r;
        rho[m] = r;
        p[m] = g1 * ((float)st[q][4] - 0.5f * (mx * mx + my * my + mz * mz) * inv);
    }
    const float *src[2] = {rho, p};
#pragma unroll
    for (int v = 0; v < 2; ++v) {
        const float *q = src[v];
        const float c = 0.5f * (q[2] + q[3]);
        float dev = 0.0f;
#pragma unroll
        for (int m = 0; m < 6; ++m) dev = fmaxf(dev, fabsf(q[m] - c));
        const float inv = 1.0f / (dev + 0.01f * fabsf(c) + 1.0e-30f);
#pragma unroll
        for (int m = 0; m < 6; ++m) x[6 * v + m] = (q[m] - c) * inv;
    }
}

// sum_k |omega_k - C_k| of one WENO5 evaluation (same weights as fast::weno5).
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
    double u, v, w, hT, c, q, b, d, un;
    double lp[5], ln[5];
};

template <int S>
__host__ __device__ __forceinline__ void face_state(const FluxModel &mdl, const double UL[5], const double UR[5], FaceState &fs)
{
    const double gamma = mdl.gamma, g1 = gamma - 1.0;
    const double rL = 1.0 / UL[0], rR = 1.0 / UR[0];
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
    const double rc = 1.0 / fs.c;
    fs.d = 0.5 * rc;
    fs.b = 0.5 * g1 * rc * rc;
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
__host__ __device__ __forceinline__ void linear_flux_from(const FaceState &fs, const double Up[5], const double Un[5], double F[5])
{
    double Wp[5], Wn[5], h[5];
    characteristic<S>(fs, Up, Wp);
    characteristic<S>(fs, Un, Wn);
#pragma unroll
    for (int r = 0; r < 5; ++r) h[r] = fma(fs.lp[r], Wp[r], fs.ln[r] * Wn[r]);
    project_back<S>(fs, h, F);
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
__host__ __device__ __forceinline__ void linear_flux_llf(double gamma, const FaceState &fs, const double Up[5], const double Um[5],
                                                double F[5])
{
    double Sg[5], Dl[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        Sg[c] = Up[c] + Um[c];
        Dl[c] = Up[c] - Um[c];
    }
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
__host__ __device__ __forceinline__ void linear_flux(const FaceState &fs, const double st[6][5], double F[5])
{
    double Up[5], Un[5];
    linear_averages(st, Up, Un);
    linear_flux_from<S>(fs, Up, Un, F);
}

// Full WENO5 flux from the same face state (matches fast::convective_flux).
// With label != nullptr also returns the weight-deviation label.
template <int S>
__host__ __device__ __forceinline__ void weno_flux(const FaceState &fs, const double st[6][5], double F[5], double *label)
{
    double W[5][6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        double t[5];
        characteristic<S>(fs, st[m], t);
#pragma unroll
        for (int r = 0; r < 5; ++r) W[r][m] = t[r];
    }
    double h[5];
    double dev = 0.0;
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        const double lp = fs.lp[r], ln = fs.ln[r];
        h[r] = fast::weno5(lp * W[r][0], lp * W[r][1], lp * W[r][2], lp * W[r][3], lp * W[r][4]) +
               fast::weno5(ln * W[r][5], ln * W[r][4], ln * W[r][3], ln * W[r][2], ln * W[r][1]);
        if (label) {
            dev = fmax(dev, weno5_weight_deviation(lp * W[r][0], lp * W[r][1], lp * W[r][2], lp * W[r][3],
                                                   lp * W[r][4]));
            dev = fmax(dev, weno5_weight_deviation(ln * W[r][5], ln * W[r][4], ln * W[r][3], ln * W[r][2],
                                                   ln * W[r][1]));
        }
    }
    if (label) *label = dev;
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
// transposed: H^T = W^T X^T, so the MFMA result layout of layer 1 (lane
// (j, half) holds hidden units 8g + 4 half + (0..3), g = 0..3, of sample j)
// is exactly the B-operand layout layer 2 needs; no LDS round trip.  The two
// 32-sample column blocks are faces 0..31 and 32..63.
//
// A fragments: lane (h, half) holds W^T[h][8 ks + 4 half + t] = W[8 ks + 4 half + t][h].
struct MlpWave {
    bf16x4 w1t[2];
    bf16x4 w2t[4];
};

__device__ __forceinline__ void mlp_wave_fragments(const uint16_t *W1, const uint16_t *W2, MlpWave &f)
{
    const int l = __lane_id();
    const int h = l & 31, kh = 4 * (l >> 5);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int t = 0; t < 4; ++t) f.w1t[ks][t] = (short)W1[(8 * ks + kh + t) * kHidden + h];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int t = 0; t < 4; ++t) f.w2t[ks][t] = (short)W2[(8 * ks + kh + t) * kHidden + h];
}

__device__ __forceinline__ uint32_t pack_bf16(float a, float b)
{
    return (uint32_t)f32_to_bf16(a) | ((uint32_t)f32_to_bf16(b) << 16);
}

__device__ __forceinline__ bf16x4 frag_of(uint32_t lo, uint32_t hi)
{
    bf16x4 v;
    v[0] = (short)(lo & 0xffffu);
    v[1] = (short)(lo >> 16);
    v[2] = (short)(hi & 0xffffu);
    v[3] = (short)(hi >> 16);
    return v;
}

// Logit of this lane's face.  Every lane of the wavefront must call it with
// the whole wavefront active.  bias points to 96 FP32 values (b1[32], b2[32],
// w3[32]) that every lane can read (LDS or global).
__device__ __forceinline__ float mlp_logit_wave(const float x[kFeat], const MlpWave &f, const float *bias, float b3)
{
    const int l = __lane_id();
    const bool upper = l >= 32;
    // own features packed as bf16 pairs: P0 = f0..3, P1 = f4..7, P2 = f8..11
    const uint32_t p0a = pack_bf16(x[0], x[1]), p0b = pack_bf16(x[2], x[3]);
    const uint32_t p1a = pack_bf16(x[4], x[5]), p1b = pack_bf16(x[6], x[7]);
    const uint32_t p2a = pack_bf16(x[8], x[9]), p2b = pack_bf16(x[10], x[11]);
    // lower lanes need P0, P2 of their upper partner; upper lanes need P1 of their lower partner
    const uint32_t s0 = upper ? p0a : p1a, s1 = upper ? p0b : p1b;
    const uint32_t s2 = upper ? p2a : 0u, s3 = upper ? p2b : 0u;
    const uint32_t r0 = (uint32_t)__shfl_xor((int)s0, 32), r1 = (uint32_t)__shfl_xor((int)s1, 32);
    const uint32_t r2 = (uint32_t)__shfl_xor((int)s2, 32), r3 = (uint32_t)__shfl_xor((int)s3, 32);
    bf16x4 xb[2][2];
    const bf16x4 zero = frag_of(0u, 0u);
    if (!upper) {
        xb[0][0] = frag_of(p0a, p0b);
        xb[0][1] = frag_of(p2a, p2b);
        xb[1][0] = frag_of(r0, r1);
        xb[1][1] = frag_of(r2, r3);
    } else {
        xb[0][0] = frag_of(r0, r1);
        xb[0][1] = zero;
        xb[1][0] = frag_of(p1a, p1b);
        xb[1][1] = zero;
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
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(f.w1t[0], xb[c][0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(f.w1t[1], xb[c][1], acc, 0, 0, 0);
        bf16x4 hb[4];
#pragma unroll
        for (int g = 0; g < 4; ++g)
            hb[g] = frag_of(pack_bf16(fmaxf(acc[4 * g], 0.0f), fmaxf(acc[4 * g + 1], 0.0f)),
                            pack_bf16(fmaxf(acc[4 * g + 2], 0.0f), fmaxf(acc[4 * g + 3], 0.0f)));
        floatx16 acc2;
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc2[4 * g + t] = bias[32 + 8 * g + kh + t];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) acc2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(f.w2t[ks], hb[ks], acc2, 0, 0, 0);
        float s = 0.0f;
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int t = 0; t < 4; ++t) s += fmaxf(acc2[4 * g + t], 0.0f) * bias[64 + 8 * g + kh + t];
        part[c] = s;
    }
    // sample j of block c: halves in lanes j and j + 32
    const float mine = upper ? part[1] : part[0];
    const float other = __shfl_xor(upper ? part[0] : part[1], 32);
    return b3 + (upper ? other + mine : mine + other);
}

}  // namespace artra
