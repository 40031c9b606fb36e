/*
 * ref_harness.c - test-infrastructure driver around the reference solver.
 *
 * This file is part of the parity oracle, never of the product.  It is
 * linked against objects compiled from the reference sources where they lie
 * (see oracle/Makefile), and only in a container that holds those sources.
 * It replaces the reference's main() (main.c:24-39) with a main program that
 *   - runs Preprocess (preprocess.c:28-42) and InitializeComputeDomain
 *     (initialization.c:38-56) on the case in the working directory,
 *   - advances with the reference's own ComputeTimeStep (solve.c:108-153) and
 *     EvolveFluidDynamics (fluid_dynamics.c:55-116) under the dt rectification
 *     of EvolveSolution (solve.c:65-78), without the field output,
 *   - writes full-precision FP64 dumps of the padded node array, and
 *   - writes per-function input/output vectors for the flux building blocks,
 *   - or (mode "solve") runs the reference's whole Solve with all of its
 *     output files, for the probe, surface-force and writer fixtures.
 *
 * Built with -DARTRA_DROPIN it is instead linked against the reference
 * objects minus the ones libartra_hip replaces (see oracle/Makefile,
 * target dropin) plus libartra_hip.so: the reference's own driver then runs
 * on the GPU through the drop-in symbols (only the "run" and "layout"
 * modes; the per-function vectors need the replaced objects).
 *
 * Dump layout ("ARTRAST1"): int32 n[3], ng[3], steps, 0; double now, dt;
 * int32 did/fid/lid/gst[N]; double U[level TO][5][N] (component-major).
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "commons.h"
#include "preprocess.h"
#include "solve.h"
#include "initialization.h"
#include "fluid_dynamics.h"
#include "solid_dynamics.h"
#include "boundary_treatment.h"
#include "convective_flux.h"
#include "diffusive_flux.h"
#include "weno.h"
#include "cfd_commons.h"
#include "computational_geometry.h"
#include <stddef.h>
#include "abi_fields.h"

/* made visible by compiling solve.c with -Dstatic= */
extern Real ComputeTimeStep(const Time *, const Space *, const Model *);

static void Die(const char *msg)
{
    fprintf(stderr, "ref_harness: %s\n", msg);
    exit(2);
}

static void DumpState(const char *fname, const Space *space, int steps, Real now, Real dt)
{
    const Partition *part = &(space->part);
    const long N = (long)part->n[X] * part->n[Y] * part->n[Z];
    FILE *fp = fopen(fname, "wb");
    if (!fp) Die("cannot open dump file");
    const char magic[8] = {'A', 'R', 'T', 'R', 'A', 'S', 'T', '1'};
    fwrite(magic, 1, 8, fp);
    int32_t hdr[8] = {part->n[X], part->n[Y], part->n[Z], part->ng[X], part->ng[Y], part->ng[Z], steps, 0};
    fwrite(hdr, sizeof(int32_t), 8, fp);
    double t2[2] = {now, dt};
    fwrite(t2, sizeof(double), 2, fp);
    int32_t *iv = malloc(N * sizeof(int32_t));
    double *dv = malloc(N * sizeof(double));
    for (int f = 0; f < 4; ++f) {
        for (long q = 0; q < N; ++q) {
            const Node *nd = space->node + q;
            iv[q] = (0 == f) ? nd->did : (1 == f) ? nd->fid : (2 == f) ? nd->lid : nd->gst;
        }
        fwrite(iv, sizeof(int32_t), N, fp);
    }
    for (int c = 0; c < DIMU; ++c) {
        for (long q = 0; q < N; ++q) {
            dv[q] = space->node[q].U[TO][c];
        }
        fwrite(dv, sizeof(double), N, fp);
    }
    free(iv);
    free(dv);
    fclose(fp);
}

static double WallSeconds(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

/*
 * Time loop of EvolveSolution (solve.c:48-107) minus output, with the solid
 * dynamics half-steps where bodies move (psi != 0).  maxSteps caps the loop
 * in addition to the case's own stepN.
 */
static int March(Time *time, Space *space, const Model *model, int maxSteps, FILE *dtLog, double *elapsed)
{
    const Real zero = 0.0;
    const Real tmInt = (INT_MAX == time->dataW[PROSD]) ? time->end : time->end / (Real)(time->dataW[PROSD]);
    Real rcInt = zero;
    Real dt = zero;
    int done = 0;
    double t0 = WallSeconds();
    while ((time->now < time->end) && (time->stepC < time->stepN) && (done < maxSteps)) {
        ++(time->stepC);
        dt = ComputeTimeStep(time, space, model);
        if (rcInt + dt > tmInt) {
            dt = tmInt - rcInt;
            rcInt = zero;
        } else {
            rcInt = rcInt + dt;
        }
        time->now = time->now + dt;
        if (time->now > time->end) {
            dt = time->end - (time->now - dt);
            time->now = time->end;
        }
        if (0 != model->psi) EvolveSolidDynamics(time->now, 0.5 * dt, space, model);
        EvolveFluidDynamics(dt, space, model);
        if (0 != model->psi) EvolveSolidDynamics(time->now, 0.5 * dt, space, model);
        if (dtLog) fwrite(&dt, sizeof(double), 1, dtLog);
        ++done;
    }
    *elapsed = WallSeconds() - t0;
    return done;
}

static int RunCase(int argc, char **argv)
{
    if (argc < 4) Die("usage: run <max_steps|0> <out_prefix>");
    int maxSteps = atoi(argv[2]);
    if (maxSteps <= 0) maxSteps = INT_MAX;
    const char *prefix = argv[3];
    char fname[1024];
    Time time = {0};
    Space space = {0};
    Model model = {0};
    space.part.proc[X] = space.part.proc[Y] = space.part.proc[Z] = 1;
    space.part.procN = 1;
    FILE *out = stdout;
    stdout = fopen("/dev/null", "w");
    Preprocess(&time, &space, &model);
    InitializeComputeDomain(&time, &space, &model);
    snprintf(fname, sizeof fname, "%s_init.bin", prefix);
    DumpState(fname, &space, 0, time.now, 0.0);
    snprintf(fname, sizeof fname, "%s_dt.bin", prefix);
    FILE *dtLog = fopen(fname, "wb");
    double elapsed = 0.0;
    int steps = March(&time, &space, &model, maxSteps, dtLog, &elapsed);
    fclose(dtLog);
    snprintf(fname, sizeof fname, "%s_final.bin", prefix);
    DumpState(fname, &space, steps, time.now, 0.0);
    fclose(stdout);
    stdout = out;
    long fluid = 0;
    const Partition *part = &space.part;
    for (int k = part->ns[PIN][Z][MIN]; k < part->ns[PIN][Z][MAX]; ++k)
        for (int j = part->ns[PIN][Y][MIN]; j < part->ns[PIN][Y][MAX]; ++j)
            for (int i = part->ns[PIN][X][MIN]; i < part->ns[PIN][X][MAX]; ++i)
                fluid += (0 == space.node[IndexNode(k, j, i, part->n[Y], part->n[X])].did);
    printf("{\"steps\": %d, \"now\": %.17g, \"seconds\": %.6f, \"fluid_nodes\": %ld}\n",
           steps, time.now, elapsed, fluid);
    return 0;
}

/*
 * The reference's whole solver on the case in the working directory:
 * Preprocess and Solve (solve.c:38-46), i.e. initialisation and
 * EvolveSolution with every output it makes (probes, surface force, field
 * and geometry files, artracfd.log) written into the working directory.
 * Afterwards the final state, each body's surface force (Fp, Fv, Tt: 9
 * doubles per body, as the last IntegrateSurfaceForce of the loop left
 * them) and each body's motion state (O, V, W, at, ar of level TO: 15
 * doubles per body) are dumped.
 */
static int SolveCase(int argc, char **argv)
{
    if (argc < 3) Die("usage: solve <out_prefix>");
    const char *prefix = argv[2];
    char fname[1024];
    Time time = {0};
    Space space = {0};
    Model model = {0};
    space.part.proc[X] = space.part.proc[Y] = space.part.proc[Z] = 1;
    space.part.procN = 1;
    FILE *out = stdout;
    stdout = fopen("/dev/null", "w");
    Preprocess(&time, &space, &model);
    Solve(&time, &space, &model);
    snprintf(fname, sizeof fname, "%s_final.bin", prefix);
    DumpState(fname, &space, time.stepC, time.now, 0.0);
    snprintf(fname, sizeof fname, "%s_force.bin", prefix);
    FILE *fp = fopen(fname, "wb");
    if (!fp) Die("cannot open force file");
    for (int n = 0; n < space.geo.totN; ++n) {
        const Polyhedron *poly = space.geo.poly + n;
        fwrite(poly->Fp, sizeof(double), DIMS, fp);
        fwrite(poly->Fv, sizeof(double), DIMS, fp);
        fwrite(poly->Tt, sizeof(double), DIMS, fp);
    }
    fclose(fp);
    snprintf(fname, sizeof fname, "%s_bodies.bin", prefix);
    fp = fopen(fname, "wb");
    if (!fp) Die("cannot open body file");
    for (int n = 0; n < space.geo.totN; ++n) {
        const Polyhedron *poly = space.geo.poly + n;
        fwrite(poly->O, sizeof(double), DIMS, fp);
        fwrite(poly->V[TO], sizeof(double), DIMS, fp);
        fwrite(poly->W[TO], sizeof(double), DIMS, fp);
        fwrite(poly->at[TO], sizeof(double), DIMS, fp);
        fwrite(poly->ar[TO], sizeof(double), DIMS, fp);
    }
    fclose(fp);
    fclose(stdout);
    stdout = out;
    printf("{\"steps\": %d, \"now\": %.17g, \"bodies\": %d}\n", time.stepC, time.now, space.geo.totN);
    return 0;
}

#ifndef ARTRA_DROPIN
/* splitmix64: deterministic inputs for the per-function vectors */
static uint64_t rngState = 0;
static uint64_t NextU64(void)
{
    uint64_t z = (rngState += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double Uniform(double a, double b)
{
    return a + (b - a) * ((NextU64() >> 11) * (1.0 / 9007199254740992.0));
}

/* A positive-pressure random conservative state. */
static void RandomState(double gamma, double U[DIMU])
{
    const double rho = Uniform(0.1, 5.0);
    const double u = Uniform(-2.0, 2.0), v = Uniform(-2.0, 2.0), w = Uniform(-2.0, 2.0);
    const double p = Uniform(0.05, 10.0);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = rho * w;
    U[4] = 0.5 * rho * (u * u + v * v + w * w) + p / (gamma - 1.0);
}

static void WriteArray(const char *prefix, const char *name, const void *data, size_t bytes)
{
    char fname[1024];
    snprintf(fname, sizeof fname, "%s_%s.bin", prefix, name);
    FILE *fp = fopen(fname, "wb");
    if (!fp) Die("cannot open vector file");
    fwrite(data, 1, bytes, fp);
    fclose(fp);
}

/*
 * Per-function vectors.  Outputs, all little-endian:
 *   weno5_in [cnt][5][5], weno5_out [cnt][5]; weno3_in [cnt][3][5], weno3_out [cnt][5]
 *   grid_U [5][N] (random states on an n^3 grid, node-major U[q][c] written as [q][5])
 *   fhat_out [8 variants][3 dirs][F faces][5] for the faces listed in fhat_faces;
 *     variant bits: 0 sScheme (WENO3/WENO5), 1 jacobMean, 2 fluxSplit
 *   fvhat_out [3 dirs][F faces][5]
 *   prim_in [cnt][5], prim_out [cnt][6], cons_out [cnt][5] (MapConservative of prim_out)
 *   ptd_in [cnt][12] (p, v0, e01, e02), ptd_out [cnt][4] (dist2, para[3])
 */
static int Vectors(int argc, char **argv)
{
    if (argc < 5) Die("usage: vectors <out_prefix> <seed> <count>");
    const char *prefix = argv[2];
    rngState = strtoull(argv[3], NULL, 10);
    const int cnt = atoi(argv[4]);
    const double gamma = 1.4;
    /* WENO5 / WENO3: a mix of smooth, noisy and jump stencils */
    double *in5 = malloc(sizeof(double) * cnt * 25), *out5 = malloc(sizeof(double) * cnt * 5);
    double *in3 = malloc(sizeof(double) * cnt * 15), *out3 = malloc(sizeof(double) * cnt * 5);
    for (int n = 0; n < cnt; ++n) {
        const int kind = n % 4;
        for (int m = 0; m < 5; ++m) {
            for (int r = 0; r < 5; ++r) {
                double x = Uniform(-3.0, 3.0);
                if (0 == kind) x = 0.3 * r + 0.1 * m * (r + 1); /* linear */
                if (1 == kind) x = sin(0.7 * m + r) * (r + 1); /* smooth */
                if (2 == kind) x = (m < 2 + (r % 3)) ? 1.0 + r : -0.5 * r; /* jump */
                in5[(n * 5 + m) * 5 + r] = x;
            }
        }
        WENO5((Real (*)[DIMU])(in5 + n * 25), out5 + n * 5);
        for (int m = 0; m < 3; ++m) {
            for (int r = 0; r < 5; ++r) {
                in3[(n * 3 + m) * 5 + r] = in5[(n * 5 + m + 1) * 5 + r];
            }
        }
        WENO3((Real (*)[DIMU])(in3 + n * 15), out3 + n * 5);
    }
    WriteArray(prefix, "weno5_in", in5, sizeof(double) * cnt * 25);
    WriteArray(prefix, "weno5_out", out5, sizeof(double) * cnt * 5);
    WriteArray(prefix, "weno3_in", in3, sizeof(double) * cnt * 15);
    WriteArray(prefix, "weno3_out", out3, sizeof(double) * cnt * 5);
    /* ComputeFhat / ComputeFvhat on a random 10^3 grid */
    const int G = 10;
    const int partn[DIMS] = {G, G, G};
    Node *node = calloc((size_t)G * G * G, sizeof(Node));
    double *gridU = malloc(sizeof(double) * G * G * G * 5);
    for (int q = 0; q < G * G * G; ++q) {
        RandomState(gamma, node[q].U[TO]);
        for (int c = 0; c < 5; ++c) gridU[q * 5 + c] = node[q].U[TO][c];
    }
    WriteArray(prefix, "grid_U", gridU, sizeof(double) * G * G * G * 5);
    /* faces (k,j,i) with the full WENO5 and viscous stencils inside the grid */
    int nf = 0;
    int faces[512][3];
    for (int k = 3; k < G - 3; ++k)
        for (int j = 3; j < G - 3; ++j)
            for (int i = 3; i < G - 3; ++i)
                if (nf < 512) { faces[nf][0] = k; faces[nf][1] = j; faces[nf][2] = i; ++nf; }
    int32_t fcount = nf;
    WriteArray(prefix, "fhat_nfaces", &fcount, sizeof fcount);
    WriteArray(prefix, "fhat_faces", faces, sizeof(int) * 3 * nf);
    double *fout = malloc(sizeof(double) * 8 * 3 * nf * 5);
    Model model = {0};
    model.gamma = gamma;
    for (int var = 0; var < 8; ++var) {
        model.sScheme = var & 1;
        model.jacobMean = (var >> 1) & 1;
        model.fluxSplit = (var >> 2) & 1;
        model.sL = (WENOFIVE == model.sScheme) ? -2 : -1;
        model.sR = (WENOFIVE == model.sScheme) ? 3 : 2;
        for (int s = 0; s < DIMS; ++s) {
            for (int f = 0; f < nf; ++f) {
                double *o = fout + ((var * 3 + s) * nf + f) * 5;
                ComputeFhat(TO, s, faces[f][0], faces[f][1], faces[f][2], partn, node, &model, o);
            }
        }
    }
    WriteArray(prefix, "fhat_out", fout, sizeof(double) * 8 * 3 * nf * 5);
    double *fvout = malloc(sizeof(double) * 3 * nf * 5);
    model.refMu = 0.37;
    model.refT = 1.7;
    model.cv = 2.5;
    const Real dd[DIMS] = {1.0 / 0.1, 1.0 / 0.07, 1.0 / 0.13};
    for (int s = 0; s < DIMS; ++s) {
        for (int f = 0; f < nf; ++f) {
            ComputeFvhat(TO, s, faces[f][0], faces[f][1], faces[f][2], partn, dd, node, &model,
                         fvout + (s * nf + f) * 5);
        }
    }
    WriteArray(prefix, "fvhat_out", fvout, sizeof(double) * 3 * nf * 5);
    /* primitive <-> conservative maps */
    double *pin = malloc(sizeof(double) * cnt * 5), *pout = malloc(sizeof(double) * cnt * 6);
    double *cout = malloc(sizeof(double) * cnt * 5);
    const double gasR = 1.0 / (gamma * 0.1 * 0.1);
    for (int n = 0; n < cnt; ++n) {
        RandomState(gamma, pin + n * 5);
        MapPrimitive(gamma, gasR, pin + n * 5, pout + n * 6);
        MapConservative(gamma, pout + n * 6, cout + n * 5);
    }
    WriteArray(prefix, "prim_in", pin, sizeof(double) * cnt * 5);
    WriteArray(prefix, "prim_out", pout, sizeof(double) * cnt * 6);
    WriteArray(prefix, "cons_out", cout, sizeof(double) * cnt * 5);
    /* point-triangle distance across all Voronoi regions */
    double *tin = malloc(sizeof(double) * cnt * 12), *tout = malloc(sizeof(double) * cnt * 4);
    for (int n = 0; n < cnt; ++n) {
        double *t = tin + n * 12;
        for (int q = 0; q < 12; ++q) t[q] = Uniform(-1.0, 1.0);
        tout[n * 4] = PointTriangleDistance(t, t + 3, t + 6, t + 9, tout + n * 4 + 1);
    }
    WriteArray(prefix, "ptd_in", tin, sizeof(double) * cnt * 12);
    WriteArray(prefix, "ptd_out", tout, sizeof(double) * cnt * 4);
    printf("{\"count\": %d, \"faces\": %d}\n", cnt, nf);
    return 0;
}

#endif /* ARTRA_DROPIN */

/* sizeof and offsetof of the structures of the drop-in ABI (commons.h) */
static int Layout(void)
{
    ABI_LIST
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) Die("usage: artra_ref run|solve|vectors|layout ...");
    if (0 == strcmp(argv[1], "run")) return RunCase(argc, argv);
    if (0 == strcmp(argv[1], "solve")) return SolveCase(argc, argv);
#ifndef ARTRA_DROPIN
    if (0 == strcmp(argv[1], "vectors")) return Vectors(argc, argv);
#endif
    if (0 == strcmp(argv[1], "layout")) return Layout();
    Die("unknown mode");
    return 2;
}
