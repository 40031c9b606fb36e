/*
 * capi.c - the exported C-ABI of libartra_hip (declared in artra_hip.h).
 *
 * The resident solver keeps the field on the device between steps.  The
 * drop-in symbols EvolveFluidDynamics and TreatBoundary keep one solver per
 * node array, created on first use, and copy level TO (or tn) in and out
 * around every call so that the caller's node array stays authoritative.
 */
#define _POSIX_C_SOURCE 200809L
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "artra_device.h"
#include "artra_hip.h"
#include "artra_host.h"
#include "solver.h"

/* The drop-in symbols take the caller's structures as the reference lays
 * them out (commons.h:128-268, x86-64 with gcc): sizes and a few offsets as
 * the reference's own headers give them, so that a layout change fails the
 * build instead of a caller. */
#include <stddef.h>
_Static_assert(sizeof(Node) == 136 && offsetof(Node, U) == 16, "Node layout differs from commons.h:128-134");
_Static_assert(sizeof(Partition) == 1032 && offsetof(Partition, ns) == 104 && offsetof(Partition, typeBC) == 912,
               "Partition layout differs from commons.h:136-159");
_Static_assert(sizeof(Polyhedron) == 568 && offsetof(Polyhedron, V) == 128 && offsetof(Polyhedron, Fp) == 344,
               "Polyhedron layout differs from commons.h:174-205");
_Static_assert(sizeof(Geometry) == 32 && sizeof(Space) == 1072 && offsetof(Space, geo) == 8 &&
               offsetof(Space, part) == 40, "Space/Geometry layout differs from commons.h:207-228");
_Static_assert(sizeof(Time) == 104 && offsetof(Time, end) == 64, "Time layout differs from commons.h:230-243");
_Static_assert(sizeof(Model) == 152 && offsetof(Model, gamma) == 64 && offsetof(Model, g) == 120,
               "Model layout differs from commons.h:245-268");


struct ArtraCase {
    Time time;
    Space space;
    Model model;
};

static int DevError(void)
{
    return artra_set_error(1, "%s", artra_dev_error());
}

int artra_dev_failed(void)
{
    return DevError();
}

const char *artra_hip_error(void)
{
    return artra_error_text();
}

/* -------------------------------------------------------- node array I/O */
static void NodesToField(const Space *space, int tn, double *field, long N)
{
    const Node *node = space->node;
    for (long q = 0; q < N; ++q) {
        for (int c = 0; c < DIMU; ++c) field[c * N + q] = node[q].U[tn][c];
    }
}

static void FieldToNodes(Space *space, int tn, const double *field, long N)
{
    Node *node = space->node;
    for (long q = 0; q < N; ++q) {
        for (int c = 0; c < DIMU; ++c) node[q].U[tn][c] = field[c * N + q];
    }
}

static int Supported(const Space *space, const Model *model)
{
    if (model->refMu > 0.0 && OPTSPLIT == model->multidim && COLLAPSEN != space->part.collapse) {
        return artra_set_error(2, "viscous flux on a collapsed split grid reads outside the grid");
    }
    return 0;
}

/* Box groups for the full fill and for each axis's faces, plus the resolved
 * gather when the boundary types allow it. */
static int SetBoundaryTables(ArtraHipSolver *s)
{
    ArtraBcBox boxes[64], axisBoxes[64];
    int starts[65];
    const int nbox = artra_bc_boxes(s->space, boxes, 64);
    if (nbox < 0) return artra_set_error(1, "boundary box table overflow");
    for (int mode = 0; mode < 4; ++mode) {
        int n = 0;
        for (int b = 0; b < nbox; ++b) {
            const int axis = (boxes[b].face - PWB) / 2;
            if (0 == mode || axis == mode - 1) axisBoxes[n++] = boxes[b];
        }
        const int ngroup = artra_bc_groups(s->space, axisBoxes, n, starts, 64);
        if (ngroup < 0) return artra_set_error(1, "boundary group table overflow");
        if (artra_dev_set_boundary(s->dev, mode, axisBoxes, n, starts, ngroup)) return DevError();
    }
    ArtraBcGather gather;
    if (artra_bc_resolve(s->space, boxes, nbox, &gather)) return 1;
    int rc = artra_dev_set_gather(s->dev, &gather);
    const int resolved = gather.resolved;
    artra_free_bc_gather(&gather);
    if (rc) return DevError();
    /* walls keep the full fill ordered; the fills between the stages of a
     * sweep along an axis whose faces resolve become gathers */
    for (int a = 0; a < DIMS && !resolved; ++a) {
        int n = 0;
        for (int b = 0; b < nbox; ++b) {
            if ((boxes[b].face - PWB) / 2 == a) axisBoxes[n++] = boxes[b];
        }
        if (artra_bc_resolve_axis(s->space, axisBoxes, n, a, &gather)) return 1;
        rc = artra_dev_set_axis_gather(s->dev, a, &gather);
        artra_free_bc_gather(&gather);
        if (rc) return DevError();
    }
    return 0;
}

int artra_hip_attach(Space *space, const Model *model, ArtraHipSolver **out)
{
    *out = NULL;
    if (Supported(space, model)) return 2;
    const Partition *part = &space->part;
    ArtraHipSolver *s = calloc(1, sizeof *s);
    if (NULL == s) return artra_set_error(1, "out of host memory");
    s->space = space;
    s->model = model;
    s->N = (long)part->n[X] * part->n[Y] * part->n[Z];
    s->staging = malloc((size_t)5 * s->N * sizeof(double));
    s->fluid = malloc((size_t)s->N);
    if (NULL == s->staging || NULL == s->fluid) {
        artra_hip_detach(s);
        return artra_set_error(1, "out of host memory for %ld nodes", s->N);
    }
    uint8_t *fluid = s->fluid;
    for (long q = 0; q < s->N; ++q) fluid[q] = (0 == space->node[q].did);
    ArtraDevParams par;
    memset(&par, 0, sizeof par);
    for (int d = 0; d < DIMS; ++d) {
        par.n[d] = part->n[d];
        par.ng[d] = part->ng[d];
        par.m[d] = part->m[d];
        par.pin[d][MIN] = part->ns[PIN][d][MIN];
        par.pin[d][MAX] = part->ns[PIN][d][MAX];
        par.d[d] = part->d[d];
        par.dd[d] = part->dd[d];
    }
    par.collapse = part->collapse;
    par.multidim = model->multidim;
    par.tScheme = model->tScheme;
    par.sScheme = model->sScheme;
    par.jacobMean = model->jacobMean;
    par.fluxSplit = model->fluxSplit;
    par.gamma = model->gamma;
    par.gasR = model->gasR;
    par.cv = model->cv;
    par.refMu = model->refMu;
    par.refT = model->refT;
    par.sState = model->sState;
    for (int d = 0; d < DIMS; ++d) par.g[d] = model->g[d];
    for (int p = 0; p < NBC; ++p) {
        par.typeBC[p] = part->typeBC[p];
        for (int d = 0; d < DIMS; ++d) par.normal[p][d] = part->N[p][d];
        for (int v = 0; v < VARBC; ++v) par.varBC[p][v] = part->varBC[p][v];
    }
    const int rc = artra_dev_create(&par, fluid, &s->dev);
    if (rc) {
        artra_hip_detach(s);
        return DevError();
    }
    if (SetBoundaryTables(s)) {
        artra_hip_detach(s);
        return 1;
    }
    if (artra_build_ghost_table(space, model, &s->ghosts)) {
        artra_hip_detach(s);
        return 1;
    }
    if (artra_dev_set_ghosts(s->dev, &s->ghosts) || artra_dev_poison(s->dev)) {
        artra_hip_detach(s);
        return DevError();
    }
    if (artra_hip_upload(s, TO)) {
        artra_hip_detach(s);
        return 1;
    }
    if (artra_build_force_table(s)) {
        artra_hip_detach(s);
        return 1;
    }
    if (0 != model->psi && (artra_scope_init(space, &s->scope) || artra_geo_device_init(s))) {
        artra_hip_detach(s);
        return 1;
    }
    const char *exact = getenv("ARTRA_EXACT");
    if (exact && exact[0] == '1') artra_dev_set_exact(s->dev, 1);
    *out = s;
    return 0;
}

int artra_hip_detach(ArtraHipSolver *s)
{
    if (NULL == s) return 0;
    artra_dev_destroy(s->dev);
    artra_free_ghost_table(&s->ghosts);
    free(s->staging);
    free(s->force);
    free(s->fluid);
    artra_scope_free(&s->scope);
    free(s);
    return 0;
}

int artra_hip_treat_boundary(ArtraHipSolver *s, int tn)
{
    return artra_dev_treat_boundary(s->dev, tn) ? DevError() : 0;
}

int artra_hip_treat_immersed_boundary(ArtraHipSolver *s, int tn)
{
    return artra_dev_treat_immersed(s->dev, tn) ? DevError() : 0;
}

/* Body speeds of the CFL condition (solve.c:121-131): per axis the largest
 * |V_s| + max(|W_t1|, |W_t2|) r over the polyhedra. */
static void BodySpeedFloor(const ArtraHipSolver *s, double floor[DIMS])
{
    const Geometry *geo = &s->space->geo;
    floor[X] = floor[Y] = floor[Z] = 0.0;
    for (int n = 0; n < geo->totN; ++n) {
        const Polyhedron *p = geo->poly + n;
        const double V[DIMS] = {fabs(p->V[TO][X]) + fmax(fabs(p->W[TO][Y]), fabs(p->W[TO][Z])) * p->r,
                                fabs(p->V[TO][Y]) + fmax(fabs(p->W[TO][Z]), fabs(p->W[TO][X])) * p->r,
                                fabs(p->V[TO][Z]) + fmax(fabs(p->W[TO][X]), fabs(p->W[TO][Y])) * p->r};
        for (int d = 0; d < DIMS; ++d) {
            if (floor[d] < V[d]) floor[d] = V[d];
        }
    }
}

int artra_hip_compute_dt(ArtraHipSolver *s, const Time *time, Real *dt)
{
    const Partition *part = &s->space->part;
    double floor[DIMS];
    BodySpeedFloor(s, floor);
    double vmax[DIMS];
    if (artra_dev_max_speed(s->dev, floor, vmax)) return DevError();
    const double tx = part->d[X] / vmax[X], ty = part->d[Y] / vmax[Y], tz = part->d[Z] / vmax[Z];
    const double tyz = (ty < tz) ? ty : tz;
    *dt = time->numCFL * ((tx < tyz) ? tx : tyz);
    return 0;
}

int artra_hip_evolve(ArtraHipSolver *s, Real dt)
{
    return artra_dev_evolve(s->dev, dt) ? DevError() : 0;
}

/* The device march's clock for the plain time loop: no output cadence
 * (the intervals are infinite, so the march only stops at the loop's end). */
static void MarchClock(const Time *time, Real rcInt, Real tmInt, double clock[ARTRA_DEV_MARCH_CLOCK])
{
    const double base[6] = {time->now, time->end, rcInt, tmInt, (double)time->stepC, (double)time->stepN};
    for (int q = 0; q < 6; ++q) clock[q] = base[q];
    clock[6] = 0.0;
    for (int n = 0; n < NPROBE; ++n) {
        clock[7 + n] = 0.0;
        clock[12 + n] = INFINITY;
    }
}

/* Steps the device may take in one call without passing the end time: the
 * remaining time over the last step, less a margin of two steps for a
 * growing dt (a step past the end on the device is a no-op anyway). */
static int ChunkSteps(const Time *time, Real lastDt, int left)
{
    int k = ARTRA_DEV_MARCH_CHUNK;
    if (k > left) k = left;
    if (time->stepN - time->stepC < k) k = time->stepN - time->stepC;
    if (lastDt > 0.0) {
        const double safe = (time->end - time->now) / lastDt - 2.0;
        if (safe < (double)k) k = (safe >= 1.0) ? (int)safe : 1;
    } else {
        k = 1;
    }
    return (k < 1) ? 1 : k;
}

/* The time loop with moving bodies: the host takes every step, the solid
 * half-steps needing the surface force and moving the geometry between the
 * fluid steps (solve.c:65-89). */
static int MarchMoving(ArtraHipSolver *s, Time *time, Real tmInt, int max_steps, int *steps_done, Real *dts)
{
    int done = 0, rc = 0;
    while (time->now < time->end && time->stepC < time->stepN && done < max_steps) {
        Real dt = 0.0;
        if ((rc = artra_moving_step(s, time, tmInt, &s->rcInt, &dt))) break;
        if (dts) dts[done] = dt;
        ++done;
    }
    if (steps_done) *steps_done = done;
    return rc ? rc : artra_hip_synchronize(s);
}

int artra_hip_march(ArtraHipSolver *s, Time *time, int max_steps, int *steps_done, Real *dts)
{
    const Real tmInt = (INT_MAX == time->dataW[PROSD]) ? time->end : time->end / (Real)(time->dataW[PROSD]);
    if (0 != s->model->psi) return MarchMoving(s, time, tmInt, max_steps, steps_done, dts);
    double floor[DIMS];
    BodySpeedFloor(s, floor);
    double rec[ARTRA_DEV_MARCH_CHUNK];
    Real lastDt = 0.0;
    int done = 0;
    /* the time loop of solve.c:59-107 with the dt of every step computed on
     * the device (bit-identical to artra_hip_compute_dt and the loop's
     * rectification), in chunks that the host checks between */
    while (time->now < time->end && time->stepC < time->stepN && done < max_steps) {
        const int k = ChunkSteps(time, lastDt, max_steps - done);
        double clock[ARTRA_DEV_MARCH_CLOCK];
        MarchClock(time, s->rcInt, tmInt, clock);
        int adv = 0;
        if (artra_dev_march_steps(s->dev, clock, floor, time->numCFL, k, rec, &adv)) return DevError();
        time->now = clock[0];
        s->rcInt = clock[2];
        time->stepC += adv;
        for (int i = 0; i < adv; ++i) {
            if (dts) dts[done + i] = rec[i];
        }
        if (adv > 0) lastDt = rec[adv - 1];
        done += adv;
        if (adv < k) break;          /* the clock stopped inside the chunk */
    }
    if (steps_done) *steps_done = done;
    return artra_hip_synchronize(s);
}

int artra_hip_march_noalloc(ArtraHipSolver *s, Time *time, int max_steps, int *steps_done, Real *dts)
{
    if (steps_done) *steps_done = 0;
    if (max_steps < 0 || max_steps > ARTRA_NOALLOC_MAX_STEPS) {
        return artra_set_error(1, "artra_hip_march_noalloc: max_steps %d outside 0..%d", max_steps,
                               ARTRA_NOALLOC_MAX_STEPS);
    }
    if (0 != s->model->psi) {
        return artra_set_error(1, "artra_hip_march_noalloc: moving bodies (phase interaction %d) rebuild "
                               "device tables every step; use artra_hip_march", s->model->psi);
    }
    if (!(time->now < time->end && time->stepC < time->stepN) || max_steps == 0) return 0;
    const Real tmInt = (INT_MAX == time->dataW[PROSD]) ? time->end : time->end / (Real)(time->dataW[PROSD]);
    double floor[DIMS];
    BodySpeedFloor(s, floor);
    int k = max_steps;
    if (time->stepN - time->stepC < k) k = time->stepN - time->stepC;
    /* every step in one device call: steps past the end time are no-ops */
    double rec[ARTRA_NOALLOC_MAX_STEPS];
    double clock[ARTRA_DEV_MARCH_CLOCK];
    MarchClock(time, s->rcInt, tmInt, clock);
    int adv = 0;
    if (artra_dev_march_noalloc(s->dev, clock, floor, time->numCFL, k, rec, &adv)) return DevError();
    time->now = clock[0];
    s->rcInt = clock[2];
    time->stepC += adv;
    for (int i = 0; i < adv; ++i) {
        if (dts) dts[i] = rec[i];
    }
    if (steps_done) *steps_done = adv;
    return 0;
}

int artra_hip_download(ArtraHipSolver *s, int tn)
{
    if (artra_dev_download(s->dev, tn, s->staging)) return DevError();
    FieldToNodes(s->space, tn, s->staging, s->N);
    return 0;
}

int artra_hip_upload(ArtraHipSolver *s, int tn)
{
    NodesToField(s->space, tn, s->staging, s->N);
    return artra_dev_upload(s->dev, tn, s->staging) ? DevError() : 0;
}

int artra_hip_read_field(ArtraHipSolver *s, int tn, Real *field)
{
    return artra_dev_download(s->dev, tn, field) ? DevError() : 0;
}

int artra_hip_write_field(ArtraHipSolver *s, int tn, const Real *field)
{
    return artra_dev_upload(s->dev, tn, field) ? DevError() : 0;
}

int artra_hip_synchronize(ArtraHipSolver *s)
{
    return artra_dev_sync(s->dev) ? DevError() : 0;
}

int artra_hip_set_timing(ArtraHipSolver *s, int on)
{
    return artra_dev_set_timing(s->dev, on) ? DevError() : 0;
}

int artra_hip_get_timing(ArtraHipSolver *s, double ms[4], long launches[4])
{
    return artra_dev_get_timing(s->dev, ms, launches) ? DevError() : 0;
}

int artra_hip_set_tuning(ArtraHipSolver *s, int march_segment)
{
    return artra_dev_set_tuning(s->dev, march_segment, 0) ? DevError() : 0;
}

int artra_hip_set_fill_options(ArtraHipSolver *s, int use_gather, int partial)
{
    return artra_dev_set_fill_options(s->dev, use_gather, partial) ? DevError() : 0;
}

int artra_hip_set_graph(ArtraHipSolver *s, int on)
{
    return artra_dev_set_graph(s->dev, on) ? DevError() : 0;
}

int artra_hip_set_arithmetic(ArtraHipSolver *s, int exact)
{
    return artra_dev_set_exact(s->dev, exact) ? DevError() : 0;
}

int artra_hip_copy_probe(long bytes, int iters, double *gbps)
{
    return artra_dev_copy_probe(bytes, iters, gbps) ? DevError() : 0;
}

/* ------------------------------------------- learned troubled-face indicator */
static int MlpError(void)
{
    return artra_set_error(1, "%s", artra_mlp_error());
}

int artra_hip_face_count(ArtraHipSolver *s, int axis, long *count)
{
    if (axis < 0 || axis > 2) return artra_set_error(1, "bad axis %d", axis);
    *count = artra_dev_face_count(s->dev, axis);
    return 0;
}

int artra_hip_sample_faces(ArtraHipSolver *s, int axis, float *features, float *labels)
{
    if (axis < 0 || axis > 2) return artra_set_error(1, "bad axis %d", axis);
    return artra_dev_sample_faces(s->dev, axis, features, labels) ? DevError() : 0;
}

int artra_hip_set_indicator(ArtraHipSolver *s, const float *params, float logit_threshold)
{
    return artra_dev_set_indicator(s->dev, params, logit_threshold) ? DevError() : 0;
}

int artra_hip_hybrid_counts(ArtraHipSolver *s, unsigned long long counts[ARTRA_HYBRID_TALLIES], int reset)
{
    return artra_dev_hybrid_counts(s->dev, counts, reset) ? DevError() : 0;
}

int artra_hip_set_indicator_options(ArtraHipSolver *s, int every_stage, float smooth_variation)
{
    return artra_dev_set_indicator_options(s->dev, every_stage, smooth_variation) ? DevError() : 0;
}

int artra_hip_set_fused(ArtraHipSolver *s, int fused)
{
    return artra_dev_set_fused(s->dev, fused) ? DevError() : 0;
}

int artra_hip_fused_axes(ArtraHipSolver *s, int *axes)
{
    *axes = artra_dev_fused_axes(s->dev);
    return 0;
}

int artra_mlp_train(float *params, const float *features, const float *labels, long n,
                    const ArtraMlpTraining *options, float *loss_history)
{
    if (n < options->batch || options->batch <= 0 || options->batch % 64) {
        return artra_set_error(1, "training needs at least one batch of a multiple of 64 samples");
    }
    return artra_mlp_train_dev(params, features, labels, n, options, loss_history) ? MlpError() : 0;
}

int artra_mlp_gradient(const float *params, const float *features, const float *labels, long n, float tau,
                       float wpos, float *grad, float *loss)
{
    if (n <= 0) return artra_set_error(1, "no samples");
    return artra_mlp_gradient_dev(params, features, labels, n, tau, wpos, grad, loss) ? MlpError() : 0;
}

int artra_mlp_logits(const float *params, const float *features, long n, float *logits)
{
    if (n <= 0) return artra_set_error(1, "no samples");
    return artra_mlp_infer_dev(params, features, n, logits) ? MlpError() : 0;
}

/* ------------------------------------------------------------- drop-in */
static struct {
    Space *space;
    Node *node;
    const Model *model;
    ArtraHipSolver *solver;
} gDropIn;

static ArtraHipSolver *DropInSolver(Space *space, const Model *model)
{
    if (gDropIn.solver && (gDropIn.space != space || gDropIn.node != space->node || gDropIn.model != model)) {
        artra_hip_detach(gDropIn.solver);
        gDropIn.solver = NULL;
    }
    if (NULL == gDropIn.solver) {
        if (artra_hip_attach(space, model, &gDropIn.solver)) {
            fprintf(stderr, "error: %s\n", artra_hip_error());
            exit(EXIT_FAILURE);
        }
        gDropIn.space = space;
        gDropIn.node = space->node;
        gDropIn.model = model;
    }
    return gDropIn.solver;
}

void EvolveFluidDynamics(const Real dt, Space *space, const Model *model)
{
    ArtraHipSolver *s = DropInSolver(space, model);
    if (artra_hip_upload(s, TO) || artra_hip_evolve(s, dt) || artra_hip_download(s, TO)) {
        fprintf(stderr, "error: %s\n", artra_hip_error());
        exit(EXIT_FAILURE);
    }
}

void TreatBoundary(const int tn, Space *space, const Model *model)
{
    ArtraHipSolver *s = DropInSolver(space, model);
    if (artra_hip_upload(s, tn) || artra_hip_treat_boundary(s, tn) || artra_hip_download(s, tn)) {
        fprintf(stderr, "error: %s\n", artra_hip_error());
        exit(EXIT_FAILURE);
    }
}

void TreatImmersedBoundary(const int tn, Space *space, const Model *model)
{
    ArtraHipSolver *s = DropInSolver(space, model);
    if (artra_hip_upload(s, tn) || artra_hip_treat_immersed_boundary(s, tn) || artra_hip_download(s, tn)) {
        fprintf(stderr, "error: %s\n", artra_hip_error());
        exit(EXIT_FAILURE);
    }
}

void ComputeGeometricField(Space *space, const Model *model)
{
    ArtraGhostTable joins;
    memset(&joins, 0, sizeof joins);
    int rc = artra_compute_geometric_field(space, model, &joins, NULL);
    /* the nodes a moving body uncovered are reconstructed in the caller's
     * node array, which is authoritative between drop-in calls */
    if (!rc) artra_apply_joins(space, model, &joins, TO);
    artra_free_ghost_table(&joins);
    /* the cached device solver's tables follow the new flags */
    if (!rc && gDropIn.solver && gDropIn.space == space && gDropIn.node == space->node) {
        rc = artra_hip_refresh_geometry(gDropIn.solver, NULL);
    }
    if (rc) {
        fprintf(stderr, "error: %s\n", artra_hip_error());
        exit(EXIT_FAILURE);
    }
}

/* ------------------------------------------------------ case preparation */
int artra_case_open(const char *dir, ArtraCase **out)
{
    *out = NULL;
    ArtraCase *c = calloc(1, sizeof *c);
    if (NULL == c) return artra_set_error(1, "out of host memory");
    const int rc = artra_prepare_case(dir, &c->time, &c->space, &c->model);
    if (rc) {
        artra_release_case(&c->space, &c->model, &c->time);
        free(c);
        return rc;
    }
    *out = c;
    return 0;
}

void artra_case_close(ArtraCase *c)
{
    if (NULL == c) return;
    if (gDropIn.space == &c->space) {
        artra_hip_detach(gDropIn.solver);
        memset(&gDropIn, 0, sizeof gDropIn);
    }
    artra_release_case(&c->space, &c->model, &c->time);
    free(c);
}

Space *artra_case_space(ArtraCase *c) { return &c->space; }
Model *artra_case_model(ArtraCase *c) { return &c->model; }
Time *artra_case_time(ArtraCase *c) { return &c->time; }

int artra_case_grid(const ArtraCase *c, int n[3], int ng[3], int pin[6])
{
    const Partition *part = &c->space.part;
    for (int d = 0; d < DIMS; ++d) {
        n[d] = part->n[d];
        ng[d] = part->ng[d];
        pin[2 * d] = part->ns[PIN][d][MIN];
        pin[2 * d + 1] = part->ns[PIN][d][MAX];
    }
    return 0;
}

int artra_case_get_field(const ArtraCase *c, int tn, Real *field)
{
    const Partition *part = &c->space.part;
    NodesToField(&c->space, tn, field, (long)part->n[X] * part->n[Y] * part->n[Z]);
    return 0;
}

int artra_case_set_field(ArtraCase *c, int tn, const Real *field)
{
    const Partition *part = &c->space.part;
    FieldToNodes(&c->space, tn, field, (long)part->n[X] * part->n[Y] * part->n[Z]);
    return 0;
}

int artra_case_get_flags(const ArtraCase *c, int *flags)
{
    const Partition *part = &c->space.part;
    const long N = (long)part->n[X] * part->n[Y] * part->n[Z];
    for (long q = 0; q < N; ++q) {
        const Node *nd = c->space.node + q;
        flags[q] = nd->did;
        flags[N + q] = nd->fid;
        flags[2 * N + q] = nd->lid;
        flags[3 * N + q] = nd->gst;
    }
    return 0;
}

int artra_calc_expression(const char *text, Real x, Real y, Real z, Real *ans)
{
    return artra_calc(text, 0.0, x, y, z, ans);
}

/* ------------------------------------------------------------ batched cases */
struct ArtraHipBatch {
    ArtraBatch *b;
    int count;
    ArtraHipSolver **solvers;
    Time **times;
};

/* The batch clock of every case from its Time and solver: now, end, the
 * time since the last output instant and the output interval (solve.c:57-64),
 * the step count and the step limit. */
enum { kClock = 6 };

static void FillClock(ArtraHipSolver *const *solvers, Time *const *times, int count, double *clock)
{
    for (int c = 0; c < count; ++c) {
        const Time *t = times[c];
        double *ck = clock + kClock * c;
        ck[0] = t->now;
        ck[1] = t->end;
        ck[2] = solvers[c]->rcInt;
        ck[3] = (INT_MAX == t->dataW[PROSD]) ? t->end : t->end / (Real)(t->dataW[PROSD]);
        ck[4] = (double)t->stepC;
        ck[5] = (double)t->stepN;
    }
}

int artra_hip_batch_create(ArtraHipSolver *const *solvers, Time *const *times, int count, ArtraHipBatch **out)
{
    *out = NULL;
    if (count < 1) return artra_set_error(1, "a batch needs at least one case");
    ArtraHipBatch *hb = calloc(1, sizeof *hb);
    ArtraDevice **devs = malloc(count * sizeof *devs);
    double *clock = malloc(kClock * count * sizeof *clock);
    double *cfl = malloc(count * sizeof *cfl);
    if (hb) {
        hb->solvers = malloc(count * sizeof *hb->solvers);
        hb->times = malloc(count * sizeof *hb->times);
    }
    if (!hb || !devs || !clock || !cfl || !hb->solvers || !hb->times) {
        free(devs); free(clock); free(cfl);
        if (hb) { free(hb->solvers); free(hb->times); free(hb); }
        return artra_set_error(1, "out of host memory");
    }
    int rc = 0;
    for (int c = 0; c < count && !rc; ++c) {
        if (solvers[c]->space->geo.totN > 0) rc = artra_set_error(1, "batched case %d has bodies; batches take fluid-only cases", c);
        devs[c] = solvers[c]->dev;
        cfl[c] = times[c]->numCFL;
        hb->solvers[c] = solvers[c];
        hb->times[c] = times[c];
    }
    FillClock(solvers, times, count, clock);
    if (!rc && artra_dev_batch_create(devs, count, clock, cfl, &hb->b)) rc = DevError();
    free(devs);
    free(clock);
    free(cfl);
    if (rc) {
        free(hb->solvers);
        free(hb->times);
        free(hb);
        return rc;
    }
    hb->count = count;
    *out = hb;
    return 0;
}

int artra_hip_batch_destroy(ArtraHipBatch *hb)
{
    if (!hb) return 0;
    artra_dev_batch_destroy(hb->b);
    free(hb->solvers);
    free(hb->times);
    free(hb);
    return 0;
}

int artra_hip_batch_tune(ArtraHipBatch *hb, int reps, int *threads, int *grid)
{
    return artra_dev_batch_tune(hb->b, reps, threads, grid) ? DevError() : 0;
}

int artra_hip_batch_set_threads(ArtraHipBatch *hb, const int *threads)
{
    return artra_dev_batch_set_threads(hb->b, threads) ? DevError() : 0;
}

int artra_hip_batch_set_indicator(ArtraHipBatch *hb, const float *params, float logit_threshold)
{
    return artra_dev_batch_set_indicator(hb->b, params, logit_threshold) ? DevError() : 0;
}

int artra_hip_batch_counts(ArtraHipBatch *hb, unsigned long long counts[ARTRA_HYBRID_TALLIES], int reset)
{
    return artra_dev_batch_counts(hb->b, counts, reset) ? DevError() : 0;
}

int artra_hip_batch_march(ArtraHipBatch *hb, int steps, int use_graph, double *ms)
{
    double *clock = malloc(kClock * hb->count * sizeof *clock);
    if (!clock) return artra_set_error(1, "out of host memory");
    /* the clocks as the cases' Time structures hold them now (a case may have
     * been marched on its own since the last batch march) */
    FillClock(hb->solvers, hb->times, hb->count, clock);
    int rc = artra_dev_batch_set_clock(hb->b, clock) || artra_dev_batch_march(hb->b, steps, use_graph, ms) ||
             artra_dev_batch_clock(hb->b, clock);
    if (rc) {
        free(clock);
        return DevError();
    }
    for (int c = 0; c < hb->count; ++c) {
        const double *ck = clock + kClock * c;
        hb->times[c]->now = ck[0];
        hb->solvers[c]->rcInt = ck[2];
        hb->times[c]->stepC = (int)ck[4];
    }
    free(clock);
    return 0;
}

int artra_case_clock(const ArtraCase *c, double *now, int *step)
{
    *now = c->time.now;
    *step = c->time.stepC;
    return 0;
}

int artra_case_forces(const ArtraCase *c, Real *forces)
{
    const Geometry *geo = &c->space.geo;
    for (int n = 0; forces && n < geo->totN; ++n) {
        const Polyhedron *p = geo->poly + n;
        for (int d = 0; d < DIMS; ++d) {
            forces[9 * n + d] = p->Fp[d];
            forces[9 * n + 3 + d] = p->Fv[d];
            forces[9 * n + 6 + d] = p->Tt[d];
        }
    }
    return geo->totN;
}

int artra_case_bodies(const ArtraCase *c, Real *state)
{
    const Geometry *geo = &c->space.geo;
    for (int n = 0; state && n < geo->totN; ++n) {
        const Polyhedron *p = geo->poly + n;
        const Real *parts[5] = {p->O, p->V[TO], p->W[TO], p->at[TO], p->ar[TO]};
        for (int q = 0; q < 5; ++q) {
            for (int d = 0; d < DIMS; ++d) state[15 * n + 3 * q + d] = parts[q][d];
        }
    }
    return geo->totN;
}
