/*
 * artra_oracle.c - CPU restatement of ArtraCFD's fluid-update hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the device
 * path: only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it, and nothing in the product calls it.
 *
 * It restates, on the solver's own node array (Node AoS, Space/Model/Time of
 * artra_types.h), the functions of the hot path in the reference's
 * operation order, so that built with -O2 -ffp-contract=off on x86-64 it
 * reproduces the reference bit for bit:
 *
 *   oracle_compute_time_step      solve.c:108-153
 *   oracle_march                  solve.c:48-107 (dt rectification, no output)
 *   oracle_evolve_fluid_dynamics  fluid_dynamics.c:55-274 (with the gravity
 *                                 source operator, source_term.c:21-39)
 *   oracle_compute_fhat           convective_flux.c:48-124, cfd_commons.c:79-246
 *   oracle_weno5 / oracle_weno3   weno5.c:35-62 / weno3.c:35-55
 *   oracle_compute_fvhat          diffusive_flux.c:46-312, cfd_commons.c:288-313
 *   oracle_treat_boundary         boundary_treatment.c:28-192
 *   immersed-boundary fill        immersed_boundary.c:270-487,
 *                                 computational_geometry.c:449-783
 *   oracle_integrate_surface_force solid_dynamics.c:50-171
 *
 * Pinning: every function is compared with the compiled reference
 * (tests/golden fixtures made by oracle/ref_harness.c) and with closed-form
 * properties in tests/test_oracle.py.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "artra_types.h"

#define IDX(k, j, i, ny, nx) (((k) * (ny) + (j)) * (nx) + (i))

static const int kDir[DIMS][DIMS] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

static inline double Sq(double x) { return x * x; }

/* ------------------------------------------------------------ WENO */
void oracle_weno5(const double F[5][DIMU], double out[DIMU])
{
    static const double C[3] = {1.0 / 10.0, 6.0 / 10.0, 3.0 / 10.0};
    const double eps = 1.0e-6;
    for (int r = 0; r < DIMU; ++r) {
        const double f0 = F[0][r], f1 = F[1][r], f2 = F[2][r], f3 = F[3][r], f4 = F[4][r];
        const double IS0 = (13.0 / 12.0) * Sq(f0 - 2.0 * f1 + f2) + (1.0 / 4.0) * Sq(f0 - 4.0 * f1 + 3.0 * f2);
        const double IS1 = (13.0 / 12.0) * Sq(f1 - 2.0 * f2 + f3) + (1.0 / 4.0) * Sq(f1 - f3);
        const double IS2 = (13.0 / 12.0) * Sq(f2 - 2.0 * f3 + f4) + (1.0 / 4.0) * Sq(3.0 * f2 - 4.0 * f3 + f4);
        const double a0 = C[0] / Sq(eps + IS0);
        const double a1 = C[1] / Sq(eps + IS1);
        const double a2 = C[2] / Sq(eps + IS2);
        const double w0 = a0 / (a0 + a1 + a2);
        const double w1 = a1 / (a0 + a1 + a2);
        const double w2 = a2 / (a0 + a1 + a2);
        const double q0 = (1.0 / 6.0) * (2.0 * f0 - 7.0 * f1 + 11.0 * f2);
        const double q1 = (1.0 / 6.0) * (-f1 + 5.0 * f2 + 2.0 * f3);
        const double q2 = (1.0 / 6.0) * (2.0 * f2 + 5.0 * f3 - f4);
        out[r] = w0 * q0 + w1 * q1 + w2 * q2;
    }
}

void oracle_weno3(const double F[3][DIMU], double out[DIMU])
{
    static const double C[2] = {1.0 / 3.0, 2.0 / 3.0};
    const double eps = 1.0e-6;
    for (int r = 0; r < DIMU; ++r) {
        const double f0 = F[0][r], f1 = F[1][r], f2 = F[2][r];
        const double IS0 = Sq(f1 - f0);
        const double IS1 = Sq(f2 - f1);
        const double a0 = C[0] / Sq(eps + IS0);
        const double a1 = C[1] / Sq(eps + IS1);
        const double w0 = a0 / (a0 + a1);
        const double w1 = a1 / (a0 + a1);
        const double q0 = (1.0 / 2.0) * (-f0 + 3.0 * f1);
        const double q1 = (1.0 / 2.0) * (f1 + f2);
        out[r] = w0 * q0 + w1 * q1;
    }
}

/* ------------------------------------------------ face eigen-system */
typedef struct {
    double u, v, w, hT, c;
} FaceAverage;

static FaceAverage Average(int roe, double gamma, const double *UL, const double *UR)
{
    const double rhoL = UL[0];
    const double uL = UL[1] / UL[0], vL = UL[2] / UL[0], wL = UL[3] / UL[0];
    const double hTL = (UL[4] / UL[0]) * gamma - 0.5 * (uL * uL + vL * vL + wL * wL) * (gamma - 1.0);
    const double rhoR = UR[0];
    const double uR = UR[1] / UR[0], vR = UR[2] / UR[0], wR = UR[3] / UR[0];
    const double hTR = (UR[4] / UR[0]) * gamma - 0.5 * (uR * uR + vR * vR + wR * wR) * (gamma - 1.0);
    FaceAverage a;
    if (roe) {
        const double D = sqrt(rhoR / rhoL);
        a.u = (uL + D * uR) / (1.0 + D);
        a.v = (vL + D * vR) / (1.0 + D);
        a.w = (wL + D * wR) / (1.0 + D);
        a.hT = (hTL + D * hTR) / (1.0 + D);
    } else {
        a.u = 0.5 * (uL + uR);
        a.v = 0.5 * (vL + vR);
        a.w = 0.5 * (wL + wR);
        a.hT = 0.5 * (hTL + hTR);
    }
    a.c = sqrt((gamma - 1.0) * (a.hT - 0.5 * (a.u * a.u + a.v * a.v + a.w * a.w)));
    return a;
}

/* left eigenvectors along s: characteristic = L * conservative
 * (cfd_commons.c:149-205, entries and product order as written there) */
static void LeftVectors(int s, double gamma, const FaceAverage *a, double L[DIMU][DIMU])
{
    const double u = a->u, v = a->v, w = a->w, c = a->c;
    const double q = 0.5 * (u * u + v * v + w * w);
    const double b = (gamma - 1.0) / (2.0 * c * c);
    const double d = 1.0 / (2.0 * c);
    if (X == s) {
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
    } else if (Y == s) {
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

/* right eigenvectors along s: conservative = R * characteristic */
static void RightVectors(int s, const FaceAverage *a, double R[DIMU][DIMU])
{
    const double u = a->u, v = a->v, w = a->w, hT = a->hT, c = a->c;
    const double q = 0.5 * (u * u + v * v + w * w);
    const double vel[DIMS] = {u, v, w};
    const double un = vel[s];
    memset(R, 0, sizeof(double) * DIMU * DIMU);
    /* acoustic columns 0 and 4 */
    R[0][0] = 1.0;
    R[0][4] = 1.0;
    for (int t = 0; t < DIMS; ++t) {
        R[1 + t][0] = (t == s) ? vel[t] - c : vel[t];
        R[1 + t][4] = (t == s) ? vel[t] + c : vel[t];
    }
    R[4][0] = hT - un * c;
    R[4][4] = hT + un * c;
    /* column 1 + s: entropy wave; the other two: shear waves */
    for (int t = 0; t < DIMS; ++t) {
        const int col = 1 + t;
        if (t == s) {
            R[0][col] = 1.0;
            for (int m = 0; m < DIMS; ++m) R[1 + m][col] = (m == s) ? vel[m] : 0.0;
            R[4][col] = vel[s] * vel[s] - q;
        } else {
            R[0][col] = 0.0;
            for (int m = 0; m < DIMS; ++m) R[1 + m][col] = (m == t) ? 1.0 : 0.0;
            R[4][col] = vel[t];
        }
    }
}

static void SplitEigenvalues(int sw, const double lam[DIMU], double lp[DIMU], double ln[DIMU])
{
    if (sw) {
        const double eps = 1.0e-3;
        for (int r = 0; r < DIMU; ++r) {
            lp[r] = 0.5 * (lam[r] + sqrt(lam[r] * lam[r] + eps * eps));
            ln[r] = 0.5 * (lam[r] - sqrt(lam[r] * lam[r] + eps * eps));
        }
    } else {
        const double star = fabs(lam[2]) + lam[4] - lam[2];
        for (int r = 0; r < DIMU; ++r) {
            lp[r] = 0.5 * (lam[r] + star);
            ln[r] = 0.5 * (lam[r] - star);
        }
    }
}

/*
 * Numerical convective flux at the face between node (k,j,i) and its
 * neighbour along s, from the level-tn states of nodes i+sL .. i+sR.
 */
void oracle_compute_fhat(int tn, int s, int k, int j, int i, const int partn[DIMS], const Node *node,
                         const Model *model, double Fhat[DIMU])
{
    const int *h = kDir[s];
    const int idxL = IDX(k, j, i, partn[Y], partn[X]);
    const int idxR = IDX(k + h[Z], j + h[Y], i + h[X], partn[Y], partn[X]);
    const FaceAverage a = Average(model->jacobMean, model->gamma, node[idxL].U[tn], node[idxR].U[tn]);
    const double vel[DIMS] = {a.u, a.v, a.w};
    const double lam[DIMU] = {vel[s] - a.c, vel[s], vel[s], vel[s], vel[s] + a.c};
    double L[DIMU][DIMU], R[DIMU][DIMU], lp[DIMU], ln[DIMU];
    LeftVectors(s, model->gamma, &a, L);
    RightVectors(s, &a, R);
    SplitEigenvalues(model->fluxSplit, lam, lp, ln);
    const int sL = model->sL, sR = model->sR, tot = sR - sL;
    double W[6][DIMU];
    for (int n = sL, m = 0; n <= sR; ++n, ++m) {
        const double *U = node[IDX(k + n * h[Z], j + n * h[Y], i + n * h[X], partn[Y], partn[X])].U[tn];
        for (int r = 0; r < DIMU; ++r) {
            W[m][r] = 0.0;
            for (int c = 0; c < DIMU; ++c) W[m][r] = W[m][r] + L[r][c] * U[c];
        }
    }
    double HP[5][DIMU], HN[5][DIMU], hp[DIMU], hn[DIMU];
    for (int m = 0; m < tot; ++m) {
        for (int r = 0; r < DIMU; ++r) {
            HP[m][r] = lp[r] * W[m][r];
            HN[m][r] = ln[r] * W[tot - m][r];
        }
    }
    if (WENOFIVE == model->sScheme) {
        oracle_weno5((const double (*)[DIMU])HP, hp);
        oracle_weno5((const double (*)[DIMU])HN, hn);
    } else {
        oracle_weno3((const double (*)[DIMU])HP, hp);
        oracle_weno3((const double (*)[DIMU])HN, hn);
    }
    for (int r = 0; r < DIMU; ++r) {
        Fhat[r] = 0.0;
        for (int c = 0; c < DIMU; ++c) Fhat[r] = Fhat[r] + R[r][c] * (hp[c] + hn[c]);
    }
}

/* --------------------------------------------------------- viscous flux */
static double Temperature(double cv, const double *U)
{
    return (U[4] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]) / (U[0] * cv);
}

/*
 * Central viscous and heat flux at the face between (k,j,i) and its
 * neighbour along s.  Tangential derivatives average the two nodes of the
 * face over their +-1 neighbours along the other axes.
 */
void oracle_compute_fvhat(int tn, int s, int k, int j, int i, const int partn[DIMS], const double dd[DIMS],
                          const Node *node, const Model *model, double Fv[DIMU])
{
    if (0.0 >= model->refMu) {
        memset(Fv, 0, DIMU * sizeof(double));
        return;
    }
    const int *h = kDir[s];
#define NODE(dk, dj, di) (node[IDX(k + (dk), j + (dj), i + (di), partn[Y], partn[X])].U[tn])
#define VEL(U, c) ((U)[1 + (c)] / (U)[0])
    /* a = first tangential axis, b = second (X: y,z; Y: x,z; Z: x,y) */
    const int ta = (X == s) ? Y : X;
    const int tb = (Z == s) ? Y : Z;
    const int *ha = kDir[ta], *hb = kDir[tb];
    const double *U0 = NODE(0, 0, 0);
    const double *UE = NODE(h[Z], h[Y], h[X]);
    const double *Ua0 = NODE(-ha[Z], -ha[Y], -ha[X]);
    const double *Ua1 = NODE(ha[Z], ha[Y], ha[X]);
    const double *Ub0 = NODE(-hb[Z], -hb[Y], -hb[X]);
    const double *Ub1 = NODE(hb[Z], hb[Y], hb[X]);
    const double *UEa0 = NODE(h[Z] - ha[Z], h[Y] - ha[Y], h[X] - ha[X]);
    const double *UEa1 = NODE(h[Z] + ha[Z], h[Y] + ha[Y], h[X] + ha[X]);
    const double *UEb0 = NODE(h[Z] - hb[Z], h[Y] - hb[Y], h[X] - hb[X]);
    const double *UEb1 = NODE(h[Z] + hb[Z], h[Y] + hb[Y], h[X] + hb[X]);
    const double vel0[DIMS] = {VEL(U0, 0), VEL(U0, 1), VEL(U0, 2)};
    const double velE[DIMS] = {VEL(UE, 0), VEL(UE, 1), VEL(UE, 2)};
    const double T0 = Temperature(model->cv, U0), TE = Temperature(model->cv, UE);
    /* gradient tensor g[m][n] = d vel_m / d x_n */
    double g[DIMS][DIMS];
    for (int m = 0; m < DIMS; ++m) g[m][s] = (velE[m] - vel0[m]) * dd[s];
    for (int m = 0; m < DIMS; ++m) {
        g[m][ta] = 0.25 * (VEL(Ua1, m) + VEL(UEa1, m) - VEL(Ua0, m) - VEL(UEa0, m)) * dd[ta];
        g[m][tb] = 0.25 * (VEL(Ub1, m) + VEL(UEb1, m) - VEL(Ub0, m) - VEL(UEb0, m)) * dd[tb];
    }
    const double dT = (TE - T0) * dd[s];
    const double uh = 0.5 * (vel0[0] + velE[0]);
    const double vh = 0.5 * (vel0[1] + velE[1]);
    const double wh = 0.5 * (vel0[2] + velE[2]);
    const double Th = 0.5 * (T0 + TE);
    const double Ts = Th * model->refT;
    const double mu = model->refMu * (1.458e-6 * pow(Ts, 1.5) / (Ts + 110.4));
    const double heatK = model->gamma * model->cv * mu / 0.71;
    /* divergence in the reference's term order for each direction */
    double divV;
    if (X == s) divV = g[0][0] + g[1][1] + g[2][2];
    else if (Y == s) divV = g[0][0] + g[1][1] + g[2][2];
    else divV = g[0][0] + g[1][1] + g[2][2];
    Fv[0] = 0.0;
    for (int m = 0; m < DIMS; ++m) {
        if (m == s) Fv[1 + m] = mu * (g[s][s] + g[s][s] - (2.0 / 3.0) * divV);
        else Fv[1 + m] = mu * (g[m][s] + g[s][m]);
    }
    /* X: (du_dy + dv_dx), (du_dz + dw_dx); Y: (dv_dx + du_dy), (dv_dz + dw_dy); Z: (dw_dx + du_dz), (dw_dy + dv_dz) */
    if (X == s) {
        Fv[2] = mu * (g[0][1] + g[1][0]);
        Fv[3] = mu * (g[0][2] + g[2][0]);
    } else if (Y == s) {
        Fv[1] = mu * (g[1][0] + g[0][1]);
        Fv[3] = mu * (g[1][2] + g[2][1]);
    } else {
        Fv[1] = mu * (g[2][0] + g[0][2]);
        Fv[2] = mu * (g[2][1] + g[1][2]);
    }
    Fv[4] = heatK * dT + Fv[1] * uh + Fv[2] * vh + Fv[3] * wh;
#undef NODE
#undef VEL
}

/* ------------------------------------------------------ state algebra */
static void MapPrimitive(double gamma, double gasR, const double *U, double *Uo)
{
    Uo[0] = U[0];
    Uo[1] = U[1] / U[0];
    Uo[2] = U[2] / U[0];
    Uo[3] = U[3] / U[0];
    Uo[4] = (U[4] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]) * (gamma - 1.0);
    Uo[5] = Uo[4] / (Uo[0] * gasR);
}

static void MapConservative(double gamma, const double *Uo, double *U)
{
    U[0] = Uo[0];
    U[1] = Uo[0] * Uo[1];
    U[2] = Uo[0] * Uo[2];
    U[3] = Uo[0] * Uo[3];
    U[4] = 0.5 * Uo[0] * (Uo[1] * Uo[1] + Uo[2] * Uo[2] + Uo[3] * Uo[3]) + Uo[4] / (gamma - 1.0);
}

void oracle_map_primitive(double gamma, double gasR, const double *U, double *Uo) { MapPrimitive(gamma, gasR, U, Uo); }
void oracle_map_conservative(double gamma, const double *Uo, double *U) { MapConservative(gamma, Uo, U); }

static void MethodOfImage(const double *UoI, const double *UoO, double *UoG)
{
    UoG[1] = UoO[1] + UoO[1] - UoI[1];
    UoG[2] = UoO[2] + UoO[2] - UoI[2];
    UoG[3] = UoO[3] + UoO[3] - UoI[3];
    UoG[4] = UoI[4];
    UoG[5] = UoI[5];
}

/* -------------------------------------------------- geometry queries */
static double Dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double Dist2(const double *a, const double *b)
{
    const double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    return Dot3(d, d);
}

double oracle_point_triangle_distance(const double p[3], const double v0[3], const double e01[3],
                                      const double e02[3], double para[3])
{
    const double D[3] = {v0[0] - p[0], v0[1] - p[1], v0[2] - p[2]};
    const double a = Dot3(e01, e01), b = Dot3(e01, e02), c = Dot3(e02, e02);
    const double d = Dot3(e01, D), e = Dot3(e02, D), f = Dot3(D, D);
    const double det = a * c - b * b;
    double s = b * e - c * d, t = b * d - a * e, dist = 0.0;
    if (s + t <= det) {
        if (s < 0.0) {
            if (t < 0.0) {
                if (d < 0.0) {
                    t = 0.0;
                    if (-d >= a) { s = 1.0; dist = a + 2.0 * d + f; }
                    else { s = -d / a; dist = d * s + f; }
                } else {
                    s = 0.0;
                    if (e >= 0.0) { t = 0.0; dist = f; }
                    else if (-e >= c) { t = 1.0; dist = c + 2.0 * e + f; }
                    else { t = -e / c; dist = e * t + f; }
                }
            } else {
                s = 0.0;
                if (e >= 0.0) { t = 0.0; dist = f; }
                else if (-e >= c) { t = 1.0; dist = c + 2.0 * e + f; }
                else { t = -e / c; dist = e * t + f; }
            }
        } else if (t < 0.0) {
            t = 0.0;
            if (d >= 0.0) { s = 0.0; dist = f; }
            else if (-d >= a) { s = 1.0; dist = a + 2.0 * d + f; }
            else { s = -d / a; dist = d * s + f; }
        } else {
            s = s / det;
            t = t / det;
            dist = s * (a * s + b * t + 2.0 * d) + t * (b * s + c * t + 2.0 * e) + f;
        }
    } else if (s < 0.0) {
        const double t0 = b + d, t1 = c + e;
        if (t1 > t0) {
            const double num = t1 - t0, den = a - 2.0 * b + c;
            if (num >= den) { s = 1.0; t = 0.0; dist = a + 2.0 * d + f; }
            else { s = num / den; t = 1.0 - s; dist = s * (a * s + b * t + 2.0 * d) + t * (b * s + c * t + 2.0 * e) + f; }
        } else {
            s = 0.0;
            if (t1 <= 0.0) { t = 1.0; dist = c + 2.0 * e + f; }
            else if (e >= 0.0) { t = 0.0; dist = f; }
            else { t = -e / c; dist = e * t + f; }
        }
    } else if (t < 0.0) {
        const double t0 = b + e, t1 = a + d;
        if (t1 > t0) {
            const double num = t1 - t0, den = a - 2.0 * b + c;
            if (num >= den) { t = 1.0; s = 0.0; dist = c + 2.0 * e + f; }
            else { t = num / den; s = 1.0 - t; dist = s * (a * s + b * t + 2.0 * d) + t * (b * s + c * t + 2.0 * e) + f; }
        } else {
            t = 0.0;
            if (t1 <= 0.0) { s = 1.0; dist = a + 2.0 * d + f; }
            else if (d >= 0.0) { s = 0.0; dist = f; }
            else { s = -d / a; dist = d * s + f; }
        }
    } else {
        const double num = c + e - b - d;
        if (num <= 0.0) { s = 0.0; t = 1.0; dist = c + 2.0 * e + f; }
        else {
            const double den = a - 2.0 * b + c;
            if (num >= den) { s = 1.0; t = 0.0; dist = a + 2.0 * d + f; }
            else { s = num / den; t = 1.0 - s; dist = s * (a * s + b * t + 2.0 * d) + t * (b * s + c * t + 2.0 * e) + f; }
        }
    }
    para[0] = 1.0 - s - t;
    para[1] = s;
    para[2] = t;
    return (dist < 0.0) ? 0.0 : dist;
}

static int EdgeIndex(const Polyhedron *poly, int a, int b)
{
    const int hi = (a > b) ? a : b, lo = (a > b) ? b : a;
    int i = 0, j = poly->edgeN - 1;
    while (i <= j) {
        const int k = (i + j) / 2;
        if (hi == poly->e[k][0] && lo == poly->e[k][1]) return k;
        if (hi > poly->e[k][0] || (hi == poly->e[k][0] && lo > poly->e[k][1])) i = k + 1;
        else j = k - 1;
    }
    abort();
}

static void GeometricData(const double *p, int fid, const Polyhedron *poly, double *pO, double *pI, double *N)
{
    if (0 >= poly->faceN) {
        for (int s = 0; s < 3; ++s) N[s] = p[s] - poly->O[s];
        double dist = sqrt(Dot3(N, N));
        for (int s = 0; s < 3; ++s) N[s] = N[s] / dist;
        dist = poly->r - dist;
        for (int s = 0; s < 3; ++s) pO[s] = p[s] + dist * N[s];
    } else {
        const int *v = poly->f[fid];
        double v0[3], v1[3], v2[3], e01[3], e02[3], para[3];
        for (int s = 0; s < 3; ++s) {
            v0[s] = poly->v[v[0]][s];
            v1[s] = poly->v[v[1]][s];
            v2[s] = poly->v[v[2]][s];
            e01[s] = v1[s] - v0[s];
            e02[s] = v2[s] - v0[s];
        }
        oracle_point_triangle_distance(p, v0, e01, e02, para);
        const double *nn;
        if (0.0 == para[1] && 0.0 == para[2]) {
            for (int s = 0; s < 3; ++s) pO[s] = v0[s];
            nn = poly->Nv[v[0]];
        } else if (0.0 == para[1] && 1.0 == para[2]) {
            for (int s = 0; s < 3; ++s) pO[s] = v2[s];
            nn = poly->Nv[v[2]];
        } else if (0.0 == para[1]) {
            for (int s = 0; s < 3; ++s) pO[s] = v0[s] + para[2] * e02[s];
            nn = poly->Ne[EdgeIndex(poly, v[0], v[2])];
        } else if (1.0 == para[1]) {
            for (int s = 0; s < 3; ++s) pO[s] = v1[s];
            nn = poly->Nv[v[1]];
        } else if (0.0 == para[2]) {
            for (int s = 0; s < 3; ++s) pO[s] = v0[s] + para[1] * e01[s];
            nn = poly->Ne[EdgeIndex(poly, v[0], v[1])];
        } else if (0.0 == para[0]) {
            for (int s = 0; s < 3; ++s) pO[s] = v0[s] + para[1] * e01[s] + para[2] * e02[s];
            nn = poly->Ne[EdgeIndex(poly, v[1], v[2])];
        } else {
            for (int s = 0; s < 3; ++s) pO[s] = v0[s] + para[1] * e01[s] + para[2] * e02[s];
            nn = poly->Nf[fid];
        }
        for (int s = 0; s < 3; ++s) N[s] = nn[s];
    }
    for (int s = 0; s < 3; ++s) pI[s] = pO[s] + pO[s] - p[s];
}

/* ------------------------------------------------- immersed boundary */
static int MapNodeOf(double s, double smin, double dds, int ng) { return (int)((s - smin) * dds + 0.5) + ng; }
static double MapPointOf(int n, double smin, double ds, int ng) { return smin + (n - ng) * ds; }
static int InPIN(const Partition *part, int k, int j, int i)
{
    return part->ns[PIN][Z][MIN] <= k && part->ns[PIN][Z][MAX] > k && part->ns[PIN][Y][MIN] <= j &&
           part->ns[PIN][Y][MAX] > j && part->ns[PIN][X][MIN] <= i && part->ns[PIN][X][MAX] > i;
}

static void Weigh(const double *Uoh, double tiny, double w, double *wsum, double *Uo)
{
    if (tiny > w) w = tiny;
    w = 1.0 / w;
    for (int n = 0; n < DIMUo; ++n) Uo[n] = Uo[n] + Uoh[n] * w;
    *wsum = *wsum + w;
}

/* layer < 0: any node of domain did; layer >= 1: nodes of did on that ghost layer */
static double Idw(int tn, const int *c, const double *p, int h, int layer, int did, const Space *space,
                  const Model *model, double *Uo)
{
    const Partition *part = &space->part;
    double wsum = 0.0;
    memset(Uo, 0, DIMUo * sizeof(double));
    for (int r = h, tally = 0; 0 == tally; ++r) {
        for (int kh = -r; kh <= r; ++kh)
            for (int jh = -r; jh <= r; ++jh)
                for (int ih = -r; ih <= r; ++ih) {
                    const int nh[3] = {c[0] + ih, c[1] + jh, c[2] + kh};
                    if (!InPIN(part, nh[2], nh[1], nh[0])) continue;
                    const Node *nd = space->node + IDX(nh[2], nh[1], nh[0], part->n[Y], part->n[X]);
                    if (did != nd->did) continue;
                    if (layer >= 1 && layer != nd->gst) continue;
                    ++tally;
                    double ph[3], Uoh[DIMUo];
                    for (int s = 0; s < 3; ++s) ph[s] = MapPointOf(nh[s], part->domain[s][MIN], part->d[s], part->ng[s]);
                    MapPrimitive(model->gamma, model->gasR, nd->U[tn], Uoh);
                    Weigh(Uoh, part->tinyL, Dist2(p, ph), &wsum, Uo);
                }
    }
    return wsum;
}

static void Cross3(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

static void Tangents(const double *N, double *Ta, double *Tb)
{
    int mark = 2;
    if (fabs(N[mark]) > fabs(N[1])) mark = 1;
    if (fabs(N[mark]) > fabs(N[0])) mark = 0;
    if (0 == mark) { Ta[0] = 0.0; Ta[1] = -N[2]; Ta[2] = N[1]; }
    else if (1 == mark) { Ta[0] = N[2]; Ta[1] = 0.0; Ta[2] = -N[0]; }
    else { Ta[0] = -N[1]; Ta[1] = N[0]; Ta[2] = 0.0; }
    const double nrm = sqrt(Dot3(Ta, Ta));
    for (int s = 0; s < 3; ++s) Ta[s] = Ta[s] / nrm;
    Cross3(Ta, N, Tb);
}

static void Reconstruct(int tn, const int *nI, const double *pI, const Polyhedron *poly, const Space *space,
                        const Model *model, const double *pO, const double *N, double *UoO, double *Uo)
{
    double wsum = Idw(tn, nI, pI, 2, -1, 0, space, model, Uo);
    const double weight = 1.0 / wsum;
    const double r[3] = {pO[0] - poly->O[0], pO[1] - poly->O[1], pO[2] - poly->O[2]};
    double Vs[3];
    Cross3(poly->W[TO], r, Vs);
    for (int s = 0; s < 3; ++s) Vs[s] = poly->V[TO][s] + Vs[s];
    if (0.0 < poly->cf) {
        UoO[1] = Vs[0]; UoO[2] = Vs[1]; UoO[3] = Vs[2];
    } else {
        const double V[3] = {Uo[1] * weight, Uo[2] * weight, Uo[3] * weight};
        double Ta[3], Tb[3];
        Tangents(N, Ta, Tb);
        const double rhs[3] = {Dot3(Vs, N), Dot3(V, Ta), Dot3(V, Tb)};
        for (int s = 0; s < 3; ++s) UoO[1 + s] = N[s] * rhs[0] + Ta[s] * rhs[1] + Tb[s] * rhs[2];
    }
    UoO[4] = Uo[4] * weight;
    UoO[5] = (0.0 > poly->T) ? Uo[5] * weight : poly->T;
    Weigh(UoO, space->part.tinyL, Dist2(pI, pO), &wsum, Uo);
    for (int n = 0; n < DIMUo; ++n) Uo[n] = Uo[n] / wsum;
}

static void ImmersedBoundary(int tn, Space *space, const Model *model)
{
    const Partition *part = &space->part;
    const Geometry *geo = &space->geo;
    for (int n = 0; n < geo->totN; ++n) {
        const Polyhedron *poly = geo->poly + n;
        int box[3][2];
        for (int s = 0; s < 3; ++s) {
            const int lo = part->ns[PIN][s][MIN], hi = part->ns[PIN][s][MAX];
            int a = MapNodeOf(poly->box[s][MIN], part->domain[s][MIN], part->dd[s], part->ng[s]);
            int b = MapNodeOf(poly->box[s][MAX], part->domain[s][MIN], part->dd[s], part->ng[s]);
            a = (a < lo) ? lo : a; a = (a > hi - 1) ? hi - 1 : a;
            b = (b < lo) ? lo : b; b = (b > hi - 1) ? hi - 1 : b;
            box[s][0] = a;
            box[s][1] = b + 1;
        }
        for (int r = 1; r <= part->gl; ++r)
            for (int k = box[2][0]; k < box[2][1]; ++k)
                for (int j = box[1][0]; j < box[1][1]; ++j)
                    for (int i = box[0][0]; i < box[0][1]; ++i) {
                        Node *nd = space->node + IDX(k, j, i, part->n[Y], part->n[X]);
                        if (r != nd->gst || n + 1 != nd->did) continue;
                        const double pG[3] = {MapPointOf(i, part->domain[0][MIN], part->d[0], part->ng[0]),
                                              MapPointOf(j, part->domain[1][MIN], part->d[1], part->ng[1]),
                                              MapPointOf(k, part->domain[2][MIN], part->d[2], part->ng[2])};
                        double UoG[DIMUo] = {0.0}, UoO[DIMUo] = {0.0}, UoI[DIMUo];
                        if (model->ibmLayer >= r) {
                            double pO[3], pI[3], N[3];
                            GeometricData(pG, nd->fid, poly, pO, pI, N);
                            int nI[3];
                            for (int s = 0; s < 3; ++s) nI[s] = MapNodeOf(pI[s], part->domain[s][MIN], part->dd[s], part->ng[s]);
                            Reconstruct(tn, nI, pI, poly, space, model, pO, N, UoO, UoI);
                            MethodOfImage(UoI, UoO, UoG);
                        } else {
                            const int nG[3] = {i, j, k};
                            const double wsum = Idw(tn, nG, pG, 1, r - 1, n + 1, space, model, UoG);
                            for (int q = 0; q < DIMUo; ++q) UoG[q] = UoG[q] / wsum;
                        }
                        UoG[0] = UoG[4] / (UoG[5] * model->gasR);
                        MapConservative(model->gamma, UoG, nd->U[tn]);
                    }
    }
}

/* --------------------------------------------------- external boundary */
static void ApplyBox(int p, int r, int box[3][2], int tn, Space *space, const Model *model)
{
    const Partition *part = &space->part;
    Node *node = space->node;
    const double *given = part->varBC[p];
    const int *N = part->N[p];
    const int LN[3] = {part->m[0] * N[0], part->m[1] * N[1], part->m[2] * N[2]};
    const int type = part->typeBC[p];
    for (int k = box[2][0]; k < box[2][1]; ++k)
        for (int j = box[1][0]; j < box[1][1]; ++j)
            for (int i = box[0][0]; i < box[0][1]; ++i) {
                double *U = node[IDX(k, j, i, part->n[Y], part->n[X])].U[tn];
                double UoO[DIMUo] = {0.0}, UoI[DIMUo] = {0.0}, UoG[DIMUo] = {0.0}, Uoh[DIMUo] = {0.0};
                if (0 != r) {
                    if (SLIPWALL == type || NOSLIPWALL == type) {
                        MapPrimitive(model->gamma, model->gasR,
                                     node[IDX(k - r * N[2], j - r * N[1], i - r * N[0], part->n[Y], part->n[X])].U[tn], UoO);
                        MapPrimitive(model->gamma, model->gasR,
                                     node[IDX(k - 2 * r * N[2], j - 2 * r * N[1], i - 2 * r * N[0], part->n[Y], part->n[X])].U[tn],
                                     UoI);
                        MethodOfImage(UoI, UoO, UoG);
                        UoG[0] = UoG[4] / (UoG[5] * model->gasR);
                        MapConservative(model->gamma, UoG, U);
                    } else if (PERIODIC == type) {
                        memcpy(U, node[IDX(k - LN[2], j - LN[1], i - LN[0], part->n[Y], part->n[X])].U[tn], DIMU * sizeof(double));
                    } else {
                        memcpy(U, node[IDX(k - N[2], j - N[1], i - N[0], part->n[Y], part->n[X])].U[tn], DIMU * sizeof(double));
                    }
                    continue;
                }
                const double *Uh = node[IDX(k - N[2], j - N[1], i - N[0], part->n[Y], part->n[X])].U[tn];
                switch (type) {
                    case INFLOW:
                        MapConservative(model->gamma, given, U);
                        break;
                    case OUTFLOW:
                        memcpy(U, Uh, DIMU * sizeof(double));
                        break;
                    case SLIPWALL:
                    case NOSLIPWALL:
                        MapPrimitive(model->gamma, model->gasR, Uh, Uoh);
                        if (SLIPWALL == type) {
                            UoO[1] = (!N[0]) * Uoh[1];
                            UoO[2] = (!N[1]) * Uoh[2];
                            UoO[3] = (!N[2]) * Uoh[3];
                        } else {
                            UoO[1] = 0.0; UoO[2] = 0.0; UoO[3] = 0.0;
                        }
                        UoO[4] = Uoh[4];
                        UoO[5] = (0.0 > given[5]) ? Uoh[5] : given[5];
                        UoO[0] = UoO[4] / (UoO[5] * model->gasR);
                        MapConservative(model->gamma, UoO, U);
                        break;
                    default:
                        break;
                }
            }
}

void oracle_treat_boundary(int tn, Space *space, const Model *model)
{
    ImmersedBoundary(tn, space, model);
    const Partition *part = &space->part;
    const int *ng = part->ng;
    int R = ng[0];
    if (ng[1] > R) R = ng[1];
    if (ng[2] > R) R = ng[2];
    for (int r = 0; r <= R; ++r)
        for (int p = PWB; p <= PBB; ++p) {
            const int *N = part->N[p];
            int box[3][2];
            for (int s = 0; s < 3; ++s) {
                const int reach = (r < ng[s]) ? r : ng[s];
                box[s][0] = part->ns[p][s][MIN] + reach * (N[s] - !N[s]);
                box[s][1] = part->ns[p][s][MAX] + reach * (N[s] + !N[s]) - (ng[s] < r) * (!!N[s]);
            }
            if (box[0][0] >= box[0][1] || box[1][0] >= box[1][1] || box[2][0] >= box[2][1]) continue;
            ApplyBox(p, r, box, tn, space, model);
        }
}

/* --------------------------------------------------------- time stepping */
static void Sweep(double dt, double a, double b, int to, int tn, int tm, int p, Space *space, const Model *model)
{
    const Partition *part = &space->part;
    Node *node = space->node;
    const int partn[3] = {part->n[0], part->n[1], part->n[2]};
    const double dd[3] = {part->dd[0], part->dd[1], part->dd[2]};
    const double rr[3] = {dt * dd[0], dt * dd[1], dt * dd[2]};
    double FR[DIMU], FL[DIMU], FvR[DIMU], FvL[DIMU], Phi[DIMU];
    const int s0 = (DIMS == p) ? 0 : p, s1 = (DIMS == p) ? DIMS : p + 1;
    for (int s = s0; s < s1; ++s) {
        const int *h = kDir[s];
        for (int ks = part->np[s][2][MIN]; ks < part->np[s][2][MAX]; ++ks)
            for (int js = part->np[s][1][MIN]; js < part->np[s][1][MAX]; ++js) {
                int have = 0;
                for (int is = part->np[s][0][MIN]; is < part->np[s][0][MAX]; ++is) {
                    int i, j, k;
                    if (X == s) { i = is; j = js; k = ks; }
                    else if (Y == s) { i = js; j = is; k = ks; }
                    else { i = js; j = ks; k = is; }
                    Node *nd = node + IDX(k, j, i, partn[Y], partn[X]);
                    if (0 != nd->did) {
                        have = 0;
                        continue;
                    }
                    if (have) {
                        memcpy(FL, FR, sizeof FL);
                        memcpy(FvL, FvR, sizeof FvL);
                    } else {
                        oracle_compute_fhat(tn, s, k - h[Z], j - h[Y], i - h[X], partn, node, model, FL);
                        oracle_compute_fvhat(tn, s, k - h[Z], j - h[Y], i - h[X], partn, dd, node, model, FvL);
                        have = 1;
                    }
                    oracle_compute_fhat(tn, s, k, j, i, partn, node, model, FR);
                    oracle_compute_fvhat(tn, s, k, j, i, partn, dd, node, model, FvR);
                    for (int n = 0; n < DIMU; ++n) Phi[n] = FL[n] - FR[n] + FvR[n] - FvL[n];
                    double *Um = nd->U[tm];
                    const double *Uo = nd->U[to], *Un = nd->U[tn];
                    if (OPTBYOPT == model->multidim && X != s) {
                        for (int n = 0; n < DIMU; ++n) Um[n] = Um[n] + b * rr[s] * Phi[n];
                    } else {
                        for (int n = 0; n < DIMU; ++n) Um[n] = a * Uo[n] + b * (Un[n] + rr[s] * Phi[n]);
                    }
                }
            }
    }
}

static void Stages(double dt, int p, Space *space, const Model *model)
{
    if (0 == model->tScheme) {
        Sweep(dt, 0.0, 1.0, TO, TO, TN, p, space, model);
        oracle_treat_boundary(TN, space, model);
        Sweep(dt, 1.0 / 2.0, 1.0 / 2.0, TO, TN, TO, p, space, model);
        oracle_treat_boundary(TO, space, model);
    } else {
        Sweep(dt, 0.0, 1.0, TO, TO, TN, p, space, model);
        oracle_treat_boundary(TN, space, model);
        Sweep(dt, 3.0 / 4.0, 1.0 / 4.0, TO, TN, TM, p, space, model);
        oracle_treat_boundary(TM, space, model);
        Sweep(dt, 1.0 / 3.0, 2.0 / 3.0, TO, TM, TO, p, space, model);
        oracle_treat_boundary(TO, space, model);
    }
}

/* Gravity source operator (source_term.c:21-39; LLLU's PHI branch,
 * fluid_dynamics.c:177-214): every interior fluid node of the x-sweep
 * ranges takes Um = a Uo + b (Un + dt Phi(Un)), Phi = (0, rho g, rho g . u),
 * with the stage's dt itself as the factor (no 1/ds) and never the unsplit
 * accumulation. */
static void SourceSweep(double dt, double a, double b, int to, int tn, int tm, Space *space, const Model *model)
{
    const Partition *part = &space->part;
    Node *node = space->node;
    for (int k = part->np[X][Z][MIN]; k < part->np[X][Z][MAX]; ++k)
        for (int j = part->np[X][Y][MIN]; j < part->np[X][Y][MAX]; ++j)
            for (int i = part->np[X][X][MIN]; i < part->np[X][X][MAX]; ++i) {
                Node *nd = node + IDX(k, j, i, part->n[Y], part->n[X]);
                if (0 != nd->did) continue;
                const double *Un = nd->U[tn];
                const double V[3] = {Un[1] / Un[0], Un[2] / Un[0], Un[3] / Un[0]};
                const double fb[3] = {Un[0] * model->g[X], Un[0] * model->g[Y], Un[0] * model->g[Z]};
                const double Phi[DIMU] = {0.0, fb[X], fb[Y], fb[Z], fb[X] * V[X] + fb[Y] * V[Y] + fb[Z] * V[Z]};
                double *Um = nd->U[tm];
                const double *Uo = nd->U[to];
                for (int n = 0; n < DIMU; ++n) Um[n] = a * Uo[n] + b * (Un[n] + dt * Phi[n]);
            }
}

static void SourceStages(double dt, Space *space, const Model *model)
{
    if (0 == model->tScheme) {
        SourceSweep(dt, 0.0, 1.0, TO, TO, TN, space, model);
        oracle_treat_boundary(TN, space, model);
        SourceSweep(dt, 1.0 / 2.0, 1.0 / 2.0, TO, TN, TO, space, model);
        oracle_treat_boundary(TO, space, model);
    } else {
        SourceSweep(dt, 0.0, 1.0, TO, TO, TN, space, model);
        oracle_treat_boundary(TN, space, model);
        SourceSweep(dt, 3.0 / 4.0, 1.0 / 4.0, TO, TN, TM, space, model);
        oracle_treat_boundary(TM, space, model);
        SourceSweep(dt, 1.0 / 3.0, 2.0 / 3.0, TO, TM, TO, space, model);
        oracle_treat_boundary(TO, space, model);
    }
}

static void FluidOperators(double h, Space *space, const Model *model)
{
    if (OPTBYOPT == model->multidim) {
        Stages(h, DIMS, space, model);
        Stages(h, DIMS, space, model);
        return;
    }
    static const int order[9][6] = {
        /* COLLAPSEN */ {Z, Y, X, X, Y, Z},
        /* COLLAPSEX */ {Z, Y, Y, Z, -1, -1},
        /* COLLAPSEY */ {Z, X, X, Z, -1, -1},
        /* COLLAPSEZ */ {Y, X, X, Y, -1, -1},
        {-1}, /* 4 unused */
        /* COLLAPSEXY */ {Z, Z, -1, -1, -1, -1},
        {-1}, /* 6 unused */
        /* COLLAPSEXZ */ {Y, Y, -1, -1, -1, -1},
        /* COLLAPSEYZ */ {X, X, -1, -1, -1, -1}};
    const int c = space->part.collapse;
    if (c < 0 || c > 8) return;
    for (int q = 0; q < 6 && order[c][q] >= 0; ++q) Stages(h, order[c][q], space, model);
}

/* EvolveFluidDynamics (fluid_dynamics.c:55-116): the gravity source
 * half-steps (57-59, 112-114) around the split or unsplit operators. */
void oracle_evolve_fluid_dynamics(double dt, Space *space, const Model *model)
{
    const double h = 0.5 * dt;
    if (0 != model->sState) SourceStages(h, space, model);
    FluidOperators(h, space, model);
    if (0 != model->sState) SourceStages(h, space, model);
}

double oracle_compute_time_step(const Time *time, const Space *space, const Model *model)
{
    const Partition *part = &space->part;
    const Geometry *geo = &space->geo;
    double Vmax[3] = {0.0, 0.0, 0.0};
    for (int n = 0; n < geo->totN; ++n) {
        const Polyhedron *poly = geo->poly + n;
        const double V[3] = {fabs(poly->V[TO][0]) + fmax(fabs(poly->W[TO][1]), fabs(poly->W[TO][2])) * poly->r,
                             fabs(poly->V[TO][1]) + fmax(fabs(poly->W[TO][2]), fabs(poly->W[TO][0])) * poly->r,
                             fabs(poly->V[TO][2]) + fmax(fabs(poly->W[TO][0]), fabs(poly->W[TO][1])) * poly->r};
        for (int s = 0; s < 3; ++s) if (Vmax[s] < V[s]) Vmax[s] = V[s];
    }
    for (int k = part->ns[PIN][2][MIN]; k < part->ns[PIN][2][MAX]; ++k)
        for (int j = part->ns[PIN][1][MIN]; j < part->ns[PIN][1][MAX]; ++j)
            for (int i = part->ns[PIN][0][MIN]; i < part->ns[PIN][0][MAX]; ++i) {
                const Node *nd = space->node + IDX(k, j, i, part->n[Y], part->n[X]);
                if (0 != nd->did) continue;
                double Uo[DIMUo];
                MapPrimitive(model->gamma, model->gasR, nd->U[TO], Uo);
                const double c = sqrt(model->gamma * model->gasR * Uo[5]);
                for (int s = 0; s < 3; ++s) {
                    const double V = fabs(Uo[s + 1]) + c;
                    if (Vmax[s] < V) Vmax[s] = V;
                }
            }
    const double dx = part->d[0] / Vmax[0], dy = part->d[1] / Vmax[1], dz = part->d[2] / Vmax[2];
    const double m = (dy < dz) ? dy : dz;
    return time->numCFL * ((dx < m) ? dx : m);
}

/* Time loop of EvolveSolution without output; returns steps taken and the
 * dt of every step in dts (capacity max_steps). */
int oracle_march(Time *time, Space *space, const Model *model, int max_steps, double *dts)
{
    const double tmInt = (0x7fffffff == time->dataW[PROSD]) ? time->end : time->end / (double)time->dataW[PROSD];
    double rcInt = 0.0;
    int done = 0;
    while (time->now < time->end && time->stepC < time->stepN && done < max_steps) {
        ++time->stepC;
        double dt = oracle_compute_time_step(time, space, model);
        if (rcInt + dt > tmInt) {
            dt = tmInt - rcInt;
            rcInt = 0.0;
        } else {
            rcInt = rcInt + dt;
        }
        time->now = time->now + dt;
        if (time->now > time->end) {
            dt = time->end - (time->now - dt);
            time->now = time->end;
        }
        oracle_evolve_fluid_dynamics(dt, space, model);
        if (dts) dts[done] = dt;
        ++done;
    }
    return done;
}

/* ------------------------------------------------------ surface force */
/* IntegrateSurfaceForce (solid_dynamics.c:50-171) on level TO: per body with
 * state <= 0, the pressure force, the wall shear of a no-slip body in a
 * viscous flow and their torque about the body centre, summed over the
 * body's layer-2 ghost nodes in (k, j, i) order inside its bounding box,
 * scaled by -area / (interfacial layer-2 nodes); an equilibrium pressure
 * (variance below FLT_EPSILON^2 of the squared mean) gives zero. */
void oracle_integrate_surface_force(Space *space, const Model *model)
{
    const Partition *part = &space->part;
    const Node *node = space->node;
    Geometry *geo = &space->geo;
    const double percent = FLT_EPSILON * FLT_EPSILON;
    for (int n = 0; n < geo->totN; ++n) {
        Polyhedron *poly = geo->poly + n;
        if (0 < poly->state) continue;
        for (int s = 0; s < DIMS; ++s) poly->Fp[s] = poly->Fv[s] = poly->Tt[s] = 0.0;
        double fvar[3] = {0.0, 0.0, 0.0};
        int lidN = 0, gstN = 0;
        int box[DIMS][LIMIT];
        for (int s = 0; s < DIMS; ++s) {
            const int lo = part->ns[PIN][s][MIN], hi = part->ns[PIN][s][MAX];
            for (int e = 0; e < LIMIT; ++e) {
                int q = MapNodeOf(poly->box[s][e], part->domain[s][MIN], part->dd[s], part->ng[s]);
                q = (q < lo) ? lo : q;
                q = (q > hi - 1) ? hi - 1 : q;
                box[s][e] = q + e;
            }
        }
        for (int k = box[Z][MIN]; k < box[Z][MAX]; ++k) {
            for (int j = box[Y][MIN]; j < box[Y][MAX]; ++j) {
                for (int i = box[X][MIN]; i < box[X][MAX]; ++i) {
                    const Node *nd = node + IDX(k, j, i, part->n[Y], part->n[X]);
                    if (2 == nd->lid && n + 1 == nd->did) ++lidN;
                    if (2 != nd->gst || n + 1 != nd->did) continue;
                    ++gstN;
                    const double pG[3] = {MapPointOf(i, part->domain[X][MIN], part->d[X], part->ng[X]),
                                          MapPointOf(j, part->domain[Y][MIN], part->d[Y], part->ng[Y]),
                                          MapPointOf(k, part->domain[Z][MIN], part->d[Z], part->ng[Z])};
                    double pO[3], pI[3], N[3], Uo[DIMUo], Fp[3], Fv[3], Fs[3], Tt[3];
                    GeometricData(pG, nd->fid, poly, pO, pI, N);
                    const double r[3] = {pO[0] - poly->O[0], pO[1] - poly->O[1], pO[2] - poly->O[2]};
                    MapPrimitive(model->gamma, model->gasR, nd->U[TO], Uo);
                    for (int s = 0; s < 3; ++s) Fp[s] = Uo[4] * N[s];
                    if (1 == gstN) fvar[0] = Uo[4];
                    fvar[1] = fvar[1] + Uo[4] - fvar[0];
                    fvar[2] = fvar[2] + (Uo[4] - fvar[0]) * (Uo[4] - fvar[0]);
                    if (0.0 < model->refMu && 0.0 < poly->cf) {
                        const double Ts = Uo[5] * model->refT;
                        const double mu = model->refMu * (1.458e-6 * pow(Ts, 1.5) / (Ts + 110.4));
                        double V[3];
                        Cross3(poly->W[TO], r, V);
                        for (int s = 0; s < 3; ++s) V[s] = Uo[1 + s] - (poly->V[TO][s] + V[s]);
                        const double Vn = Dot3(V, N);
                        const double dist = sqrt(Dist2(pG, pO));
                        for (int s = 0; s < 3; ++s) Fv[s] = mu * (V[s] - Vn * N[s]) / dist;
                    } else {
                        Fv[0] = Fv[1] = Fv[2] = 0.0;
                    }
                    for (int s = 0; s < 3; ++s) Fs[s] = Fp[s] + Fv[s];
                    Cross3(r, Fs, Tt);
                    for (int s = 0; s < 3; ++s) {
                        poly->Fp[s] = poly->Fp[s] + Fp[s];
                        poly->Fv[s] = poly->Fv[s] + Fv[s];
                        poly->Tt[s] = poly->Tt[s] + Tt[s];
                    }
                }
            }
        }
        if (0 == lidN || 0 == gstN) continue;
        double ds = poly->area / lidN;
        fvar[2] = (fvar[2] - fvar[1] * fvar[1] / gstN) / gstN;
        fvar[1] = fvar[1] / gstN + fvar[0];
        if (percent * fvar[1] * fvar[1] > fvar[2]) ds = 0.0;
        for (int s = 0; s < 3; ++s) {
            poly->Fp[s] = -poly->Fp[s] * ds;
            poly->Fv[s] = -poly->Fv[s] * ds;
            poly->Tt[s] = -poly->Tt[s] * ds;
        }
    }
}
