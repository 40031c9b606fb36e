/*
 * setup.c - grid parameters, index partition, storage and the initial
 * field of a case, as the solver prepares them before its time loop.
 *
 *   artra_compute_parameters   cfd_parameters.c:28-186
 *   artra_partition_domain     domain_partition.c:19-150
 *   artra_allocate             preprocess.c:47-59
 *   artra_initialize_domain    initialization.c:38-162 (field, geometry,
 *                              classification); the first boundary treatment
 *                              is done on the device at attach time
 */
#define _POSIX_C_SOURCE 200809L
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "artra_host.h"

int artra_compute_parameters(Time *time, Space *space, Model *model)
{
    Partition *part = &space->part;
    Geometry *geo = &space->geo;
    /* collapse flags: an axis with one cell is collapsed (cfd_parameters.c:50-59) */
    part->collapse = COLLAPSEN;
    if (1 == part->m[Z]) part->collapse = COLLAPSEZ;
    if (1 == part->m[Y]) part->collapse = 2 * part->collapse + COLLAPSEY;
    if (1 == part->m[X]) part->collapse = 2 * part->collapse + COLLAPSEX;
    if (COLLAPSEXYZ == part->collapse) return artra_set_error(1, "all three axes are collapsed");
    /* stencil offsets and ghost layers of the scheme */
    if (WENOTHREE == model->sScheme) {
        model->sL = -1; model->sR = 2; part->gl = 2;
    } else {
        model->sL = -2; model->sR = 3; part->gl = 3;
    }
    for (int s = 0; s < DIMS; ++s) part->ng[s] = part->gl - 1;
    if (OPTSPLIT == model->multidim) {
        const int c = part->collapse;
        if (COLLAPSEX == c || COLLAPSEXY == c || COLLAPSEXZ == c) part->ng[X] = 0;
        if (COLLAPSEY == c || COLLAPSEXY == c || COLLAPSEYZ == c) part->ng[Y] = 0;
        if (COLLAPSEZ == c || COLLAPSEXZ == c || COLLAPSEYZ == c) part->ng[Z] = 0;
    }
    for (int s = 0, p = PWB; s < DIMS; ++s, p += 2) {
        if (PERIODIC == part->typeBC[p]) part->ng[s] = part->gl;
    }
    for (int s = 0; s < DIMS; ++s) {
        if (part->m[s] < 2) part->m[s] = 2;
        part->n[s] = part->m[s] + 1 + 2 * part->ng[s];
    }
    /* nondimensionalisation (cfd_parameters.c:129-186) */
    for (int s = 0; s < DIMS; ++s) {
        part->domain[s][MAX] = part->domain[s][MAX] / model->refL;
        part->domain[s][MIN] = part->domain[s][MIN] / model->refL;
        part->d[s] = (part->domain[s][MAX] - part->domain[s][MIN]) / (Real)(part->m[s]);
        part->dd[s] = 1.0 / part->d[s];
    }
    const Real dmin = (part->d[X] < part->d[Y]) ? ((part->d[X] < part->d[Z]) ? part->d[X] : part->d[Z])
                                                : ((part->d[Y] < part->d[Z]) ? part->d[Y] : part->d[Z]);
    part->tinyL = 1.0e-6 * dmin;
    part->tinyL = part->tinyL * part->tinyL;
    time->end = time->end * model->refV / model->refL;
    if (0 >= time->stepN) time->stepN = INT_MAX;
    time->dataC = time->restart;
    for (int n = 0; n < NPROBE; ++n) {
        if (0 >= time->dataN[n]) time->dataN[n] = 0;
        if (0 >= time->dataW[n]) time->dataW[n] = INT_MAX;
    }
    if (0 >= geo->sphN) geo->sphN = 0;
    if (0 >= geo->stlN) geo->stlN = 0;
    geo->totN = geo->sphN + geo->stlN;
    if (0 >= model->ibmLayer) model->ibmLayer = INT_MAX;
    model->gamma = 1.4;
    model->gasR = 287.058;
    for (int s = 0; s < DIMS; ++s) model->g[s] = model->g[s] * model->refL / (model->refV * model->refV);
    model->sState = model->gState;
    model->refMa = model->refV / sqrt(model->gamma * model->gasR * model->refT);
    model->refMu = model->refMu / (model->refRho * model->refV * model->refL);
    model->gasR = 1.0 / (model->gamma * model->refMa * model->refMa);
    model->cv = model->gasR / (model->gamma - 1.0);
    return 0;
}

void artra_partition_domain(Space *space)
{
    Partition *part = &space->part;
    static const int normal[NBC][DIMS] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0},
                                          {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    memcpy(part->N, normal, sizeof normal);
    for (int s = 0, q = PWB; s < DIMS; ++s, q += 2) {
        const int periodic = (PERIODIC == part->typeBC[q]);
        part->ns[PIN][s][MIN] = part->ng[s] + (periodic ? 0 : 1);
        part->ns[PIN][s][MAX] = part->n[s] - part->ng[s] - (periodic ? 0 : 1);
        for (int p = PWB; p <= PBB; ++p) {
            part->ns[p][s][MIN] = part->ng[s];
            part->ns[p][s][MAX] = part->n[s] - part->ng[s];
        }
        for (int p = PWG; p <= PBG; ++p) {
            part->ns[p][s][MIN] = 0;
            part->ns[p][s][MAX] = part->n[s];
        }
        part->ns[PHY][s][MIN] = part->ng[s];
        part->ns[PHY][s][MAX] = part->n[s] - part->ng[s];
        part->ns[PAL][s][MIN] = 0;
        part->ns[PAL][s][MAX] = part->n[s];
    }
    /* each boundary box is one node layer thick along its own axis */
    for (int s = 0; s < DIMS; ++s) {
        part->ns[PWB + 2 * s][s][MAX] = part->ns[PWB + 2 * s][s][MIN] + 1;
        part->ns[PEB + 2 * s][s][MIN] = part->ns[PEB + 2 * s][s][MAX] - 1;
        part->ns[PWG + 2 * s][s][MAX] = part->ng[s];
        part->ns[PEG + 2 * s][s][MIN] = part->n[s] - part->ng[s];
    }
    /* sweep ranges: np[s][0] is the sweep axis, then the two others */
    static const int order[DIMS][DIMS] = {{X, Y, Z}, {Y, X, Z}, {Z, X, Y}};
    for (int s = 0; s < DIMS; ++s) {
        for (int a = 0; a < DIMS; ++a) {
            part->np[s][a][MIN] = part->ns[PIN][order[s][a]][MIN];
            part->np[s][a][MAX] = part->ns[PIN][order[s][a]][MAX];
        }
    }
    /* interfacial search path: 6 face neighbours, 18 at distance 2 (12 edge
     * diagonals + 6 face steps of 2), 6 face steps of 3 (domain_partition.c:130-148) */
    static const int path[PATHN][DIMS] = {
        {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
        {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0},
        {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
        {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1},
        {-2, 0, 0}, {2, 0, 0}, {0, -2, 0}, {0, 2, 0}, {0, 0, -2}, {0, 0, 2},
        {-3, 0, 0}, {3, 0, 0}, {0, -3, 0}, {0, 3, 0}, {0, 0, -3}, {0, 0, 3}};
    memcpy(part->path, path, sizeof path);
    part->pathSep[1] = 6;
    part->pathSep[2] = 24;
    part->pathSep[3] = 30;
    part->pathSep[0] = part->pathSep[part->gl];
}

int artra_allocate(Space *space, Model *model)
{
    Partition *part = &space->part;
    Geometry *geo = &space->geo;
    const size_t total = (size_t)part->n[X] * part->n[Y] * part->n[Z];
    if (total > (size_t)INT_MAX) return artra_set_error(1, "grid of %zu nodes exceeds the int index range", total);
    space->node = calloc(total, sizeof(*space->node));
    if (NULL == space->node) return artra_set_error(1, "out of host memory for %zu nodes", total);
    if (geo->totN) {
        geo->col = calloc(geo->totN, sizeof(*geo->col));
        geo->poly = calloc(geo->totN, sizeof(*geo->poly));
        if (!geo->col || !geo->poly) return artra_set_error(1, "out of host memory");
    }
    model->mat = calloc(1, sizeof(*model->mat));
    return (NULL == model->mat) ? artra_set_error(1, "out of host memory") : 0;
}

/* Region test of one initialiser at point pc (initialization.c:117-157). */
static int InRegion(int type, const Real pos[POSIC], const Real pc[DIMS])
{
    const Real p1[DIMS] = {pos[0], pos[1], pos[2]};
    const Real p2[DIMS] = {pos[3], pos[4], pos[5]};
    const Real r = pos[6];
    const Real P1P2[DIMS] = {p2[X] - p1[X], p2[Y] - p1[Y], p2[Z] - p1[Z]};
    const Real l2 = artra_dot(P1P2, P1P2);
    Real P1Pc[DIMS] = {pc[X] - p1[X], pc[Y] - p1[Y], pc[Z] - p1[Z]};
    Real proj;
    switch (type) {
        case ICGLOBAL:
            return 1;
        case ICPLANE:
            return 0.0 <= artra_dot(P1Pc, p2);
        case ICSPHERE:
            return r * r >= artra_dot(P1Pc, P1Pc);
        case ICBOX:
            P1Pc[X] = P1Pc[X] * (pc[X] - p2[X]);
            P1Pc[Y] = P1Pc[Y] * (pc[Y] - p2[Y]);
            P1Pc[Z] = P1Pc[Z] * (pc[Z] - p2[Z]);
            return (0.0 >= P1Pc[X]) && (0.0 >= P1Pc[Y]) && (0.0 >= P1Pc[Z]);
        case ICCYLINDER:
            proj = artra_dot(P1Pc, P1P2);
            if ((0.0 > proj) || (l2 < proj)) return 0;
            proj = artra_dot(P1Pc, P1Pc) - proj * proj / l2;
            return r * r >= proj;
        default:
            return 0;
    }
}

/* Initial field (initialization.c:68-162).  Every node starts with the
 * 0x01 byte pattern in all three levels and flags -1; interior nodes get
 * flags 0 and the initialisers applied in order. */
static int InitializeField(Space *space, const Model *model)
{
    const Partition *part = &space->part;
    ArtraExpr *expr[NIC][VARIC] = {{NULL}};
    for (int n = 0; n < part->nIC; ++n) {
        for (int v = 0; v < VARIC; ++v) {
            if (artra_expr_compile(part->varIC[n][v], &expr[n][v])) return 1;
        }
    }
    const long total = (long)part->n[X] * part->n[Y] * part->n[Z];
#pragma omp parallel for schedule(static)
    for (long q = 0; q < total; ++q) {
        Node *nd = space->node + q;
        nd->did = nd->fid = nd->lid = nd->gst = NONE;
        memset(nd->U, 1, sizeof nd->U);
        const int i = (int)(q % part->n[X]);
        const int j = (int)((q / part->n[X]) % part->n[Y]);
        const int k = (int)(q / ((long)part->n[X] * part->n[Y]));
        if (!artra_in_box(k, j, i, part->ns[PIN])) continue;
        nd->did = nd->fid = nd->lid = nd->gst = 0;
        const Real pc[DIMS] = {artra_map_point(i, part->domain[X][MIN], part->d[X], part->ng[X]),
                               artra_map_point(j, part->domain[Y][MIN], part->d[Y], part->ng[Y]),
                               artra_map_point(k, part->domain[Z][MIN], part->d[Z], part->ng[Z])};
        for (int n = 0; n < part->nIC; ++n) {
            /* the five expressions share one answer register, starting at 0 */
            Real ans = 0.0;
            Real Uo[DIMUo] = {0.0};
            for (int v = 0; v < VARIC; ++v) {
                artra_expr_eval(expr[n][v], 0.0, pc[X], pc[Y], pc[Z], &ans);
                Uo[v] = ans;
            }
            if (InRegion(part->typeIC[n], part->posIC[n], pc)) {
                artra_map_conservative(model->gamma, Uo, nd->U[TO]);
            }
        }
    }
    for (int n = 0; n < part->nIC; ++n) {
        for (int v = 0; v < VARIC; ++v) artra_expr_free(expr[n][v]);
    }
    return 0;
}

/* Body state lines of artracfd.geo (data_stream.c:154-188). */
static int ReadPolyState(FILE *fp, int pm, int pn, Geometry *geo)
{
    char buf[STR];
    for (int n = pm; n < pn; ++n) {
        Polyhedron *poly = geo->poly + n;
        if (NULL == fgets(buf, sizeof buf, fp)) return artra_set_error(1, "artracfd.geo: truncated state");
        if (16 != sscanf(buf, "%lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %d",
                         &poly->O[X], &poly->O[Y], &poly->O[Z], &poly->r, &poly->V[TO][X], &poly->V[TO][Y],
                         &poly->V[TO][Z], &poly->W[TO][X], &poly->W[TO][Y], &poly->W[TO][Z], &poly->rho,
                         &poly->T, &poly->cf, &poly->area, &poly->volume, &poly->mid)) {
            return artra_set_error(1, "artracfd.geo: bad state line '%s'", buf);
        }
        if (NULL == fgets(buf, sizeof buf, fp)) return artra_set_error(1, "artracfd.geo: truncated state");
        if (16 != sscanf(buf, "%lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg",
                         &poly->at[TO][X], &poly->at[TO][Y], &poly->at[TO][Z], &poly->ar[TO][X], &poly->ar[TO][Y],
                         &poly->ar[TO][Z], &poly->at[TN][X], &poly->at[TN][Y], &poly->at[TN][Z], &poly->g[X],
                         &poly->g[Y], &poly->g[Z], &poly->ar[TN][X], &poly->ar[TN][Y], &poly->ar[TN][Z],
                         &poly->to)) {
            return artra_set_error(1, "artracfd.geo: bad state line '%s'", buf);
        }
        if (0.0 >= poly->to) poly->to = FLT_MAX;
        if (geo->sphN > n) {
            poly->faceN = 0;
            poly->facet = NULL;
        }
        for (int s = 0; s < DIMS; ++s) {
            poly->V[TN][s] = poly->V[TO][s];
            poly->W[TN][s] = poly->W[TO][s];
        }
    }
    return 0;
}

/* Geometry sections of artracfd.geo (initialization.c:163-209). */
static int InitializeGeometry(const char *dir, Geometry *geo)
{
    char path[4096];
    snprintf(path, sizeof path, "%s/artracfd.geo", dir);
    FILE *fp = fopen(path, "r");
    if (NULL == fp) return artra_set_error(1, "failed to open file: %s", path);
    char buf[STR];
    int rc = 0;
    while (0 == rc && NULL != fgets(buf, sizeof buf, fp)) {
        artra_parse_line(buf);
        if (0 == strcmp(buf, "sphere state begin")) {
            rc = ReadPolyState(fp, 0, geo->sphN, geo);
        } else if (0 == strcmp(buf, "polyhedron geometry begin")) {
            for (int n = geo->sphN; n < geo->totN && 0 == rc; ++n) {
                char name[STR];
                if (NULL == fgets(buf, sizeof buf, fp) || 1 != sscanf(buf, "%199s", name)) {
                    rc = artra_set_error(1, "artracfd.geo: missing STL file name");
                    break;
                }
                char stl[4400];
                snprintf(stl, sizeof stl, "%s/%s", dir, name);
                rc = artra_read_stl(stl, geo->poly + n);
                if (!rc) rc = artra_convert_polyhedron(geo->poly + n);
            }
        } else if (0 == strcmp(buf, "polyhedron state begin")) {
            rc = ReadPolyState(fp, geo->sphN, geo->totN, geo);
        } else if (0 == strcmp(buf, "polyhedron transform begin")) {
            for (int n = geo->sphN; n < geo->totN && 0 == rc; ++n) {
                Real sc[DIMS], an[DIMS], of[DIMS];
                if (NULL == fgets(buf, sizeof buf, fp) ||
                    9 != sscanf(buf, "%lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg, %lg", sc + X, sc + Y, sc + Z,
                                an + X, an + Y, an + Z, of + X, of + Y, of + Z)) {
                    rc = artra_set_error(1, "artracfd.geo: bad transform line");
                    break;
                }
                if (1.0 == sc[X] && 1.0 == sc[Y] && 1.0 == sc[Z] && 0.0 == an[X] && 0.0 == an[Y] &&
                    0.0 == an[Z] && 0.0 == of[X] && 0.0 == of[Y] && 0.0 == of[Z]) {
                    continue;
                }
                artra_transform_polyhedron(geo->poly[n].O, sc, an, of, geo->poly + n);
            }
        }
    }
    fclose(fp);
    return rc;
}

/* Stationary and force-free bodies (initialization.c:235-256). */
void artra_identify_geometry_state(Geometry *geo)
{
    for (int n = 0; n < geo->totN; ++n) {
        Polyhedron *poly = geo->poly + n;
        if (1.0e10 < poly->rho) {
            poly->state = 2;
            if (0.0 == poly->V[TO][X] && 0.0 == poly->V[TO][Y] && 0.0 == poly->V[TO][Z] &&
                0.0 == poly->W[TO][X] && 0.0 == poly->W[TO][Y] && 0.0 == poly->W[TO][Z] &&
                0.0 == poly->at[TN][X] && 0.0 == poly->at[TN][Y] && 0.0 == poly->at[TN][Z] &&
                0.0 == poly->g[X] && 0.0 == poly->g[Y] && 0.0 == poly->g[Z] &&
                0.0 == poly->ar[TN][X] && 0.0 == poly->ar[TN][Y] && 0.0 == poly->ar[TN][Z] &&
                1.0e36 < poly->rho) {
                poly->state = 1;
            }
        }
    }
}

int artra_initialize_domain(const char *dir, Time *time, Space *space, Model *model)
{
    (void)time;
    if (InitializeField(space, model)) return 1;
    if (InitializeGeometry(dir, &space->geo)) return 1;
    artra_compute_geometry_parameters(space->part.collapse, &space->geo);
    if (artra_compute_geometric_field(space, model, NULL, NULL)) return 1;
    artra_identify_geometry_state(&space->geo);
    return 0;
}

int artra_prepare_case(const char *dir, Time *time, Space *space, Model *model)
{
    memset(time, 0, sizeof *time);
    memset(space, 0, sizeof *space);
    memset(model, 0, sizeof *model);
    if (artra_load_case(dir, time, space, model)) return 1;
    if (artra_compute_parameters(time, space, model)) return 1;
    artra_partition_domain(space);
    if (artra_allocate(space, model)) return 1;
    return artra_initialize_domain(dir, time, space, model);
}

void artra_release_case(Space *space, Model *model, Time *time)
{
    Partition *part = &space->part;
    free(space->node);
    for (int n = 0; n < space->geo.totN; ++n) {
        Polyhedron *poly = space->geo.poly + n;
        free(poly->f); free(poly->Nf); free(poly->e); free(poly->Ne); free(poly->v); free(poly->Nv);
        free(poly->facet);
    }
    free(space->geo.poly);
    free(space->geo.col);
    free(part->typeBC); free(part->N); free(part->varBC); free(part->typeIC); free(part->posIC);
    free(part->varIC);
    free(model->mat);
    free(time->pp);
    free(time->lp);
    memset(space, 0, sizeof *space);
    memset(model, 0, sizeof *model);
    memset(time, 0, sizeof *time);
}
