/*
 * output.c - what the solver loop writes, read from the device-resident
 * field without downloading it: the surface force of the bodies
 * (solid_dynamics.c:50-171) as a device reduction over their layer-2 ghost
 * nodes, and the point, line, body-conformal (curve) and surface-force
 * probes (data_probe.c:22-197) from device gathers of the probed nodes
 * only.  The files are those the reference writes, line for line: the same
 * names, headers, node selection and "%.6g" formatting.
 */
#define _POSIX_C_SOURCE 200809L
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "artra_hip.h"
#include "solver.h"

/* Node-space box of a body: its bounding box mapped to nodes and confined
 * to the range [lo, hi) of each axis (solid_dynamics.c:89-92,
 * data_probe.c:149-152). */
static void BodyBox(const Partition *part, const Polyhedron *poly, int range, int box[DIMS][LIMIT])
{
    for (int s = 0; s < DIMS; ++s) {
        const int lo = part->ns[range][s][MIN], hi = part->ns[range][s][MAX];
        box[s][MIN] = artra_confine(artra_map_node(poly->box[s][MIN], part->domain[s][MIN], part->dd[s], part->ng[s]),
                                    lo, hi);
        box[s][MAX] = artra_confine(artra_map_node(poly->box[s][MAX], part->domain[s][MIN], part->dd[s], part->ng[s]),
                                    lo, hi) + 1;
    }
}

static void NodePoint(const Partition *part, int k, int j, int i, Real p[DIMS])
{
    p[X] = artra_map_point(i, part->domain[X][MIN], part->d[X], part->ng[X]);
    p[Y] = artra_map_point(j, part->domain[Y][MIN], part->d[Y], part->ng[Y]);
    p[Z] = artra_map_point(k, part->domain[Z][MIN], part->d[Z], part->ng[Z]);
}

/* ------------------------------------------------------------ surface force */
int artra_build_force_table(ArtraHipSolver *s)
{
    const Space *space = s->space;
    const Partition *part = &space->part;
    const Geometry *geo = &space->geo;
    free(s->force);
    s->force = NULL;
    s->forceBodies = 0;
    int nb = 0;
    for (int n = 0; n < geo->totN; ++n) nb += (0 >= geo->poly[n].state);
    if (0 == nb) return artra_dev_set_force_table(s->dev, NULL, 0, NULL, 0) ? artra_dev_failed() : 0;
    s->force = calloc((size_t)nb, sizeof *s->force);
    ArtraForceBody *bodies = calloc((size_t)nb, sizeof *bodies);
    int cap = 1024, count = 0;
    ArtraForceNode *nodes = malloc((size_t)cap * sizeof *nodes);
    if (!s->force || !bodies || !nodes) {
        free(bodies);
        free(nodes);
        return artra_set_error(1, "out of host memory for the surface-force table");
    }
    int b = 0;
    for (int n = 0; n < geo->totN; ++n) {
        const Polyhedron *poly = geo->poly + n;
        if (0 < poly->state) continue;
        int box[DIMS][LIMIT];
        BodyBox(part, poly, PIN, box);
        ArtraForceBodyInfo *info = s->force + b;
        info->body = n;
        bodies[b].first = count;
        bodies[b].viscous = (0.0 < s->model->refMu) && (0.0 < poly->cf);
        for (int d = 0; d < DIMS; ++d) {
            bodies[b].V[d] = poly->V[TO][d];
            bodies[b].W[d] = poly->W[TO][d];
        }
        for (int k = box[Z][MIN]; k < box[Z][MAX]; ++k) {
            for (int j = box[Y][MIN]; j < box[Y][MAX]; ++j) {
                for (int i = box[X][MIN]; i < box[X][MAX]; ++i) {
                    const int idx = artra_index(k, j, i, part->n[Y], part->n[X]);
                    const Node *nd = space->node + idx;
                    if (2 == nd->lid && n + 1 == nd->did) ++info->lidN;
                    if (2 != nd->gst || n + 1 != nd->did) continue;
                    if (count == cap) {
                        cap *= 2;
                        ArtraForceNode *grown = realloc(nodes, (size_t)cap * sizeof *nodes);
                        if (!grown) {
                            free(bodies);
                            free(nodes);
                            return artra_set_error(1, "out of host memory for the surface-force table");
                        }
                        nodes = grown;
                    }
                    Real pG[DIMS], pO[DIMS], pI[DIMS], N[DIMS];
                    NodePoint(part, k, j, i, pG);
                    artra_compute_geometric_data(pG, nd->fid, poly, pO, pI, N);
                    ArtraForceNode *e = nodes + count++;
                    e->node = idx;
                    e->body = b;
                    for (int d = 0; d < DIMS; ++d) {
                        e->N[d] = N[d];
                        e->r[d] = pO[d] - poly->O[d];
                    }
                    e->dist = sqrt(artra_dist2(pG, pO));
                    ++info->gstN;
                }
            }
        }
        bodies[b].count = count - bodies[b].first;
        ++b;
    }
    s->forceBodies = nb;
    const int rc = artra_dev_set_force_table(s->dev, nodes, count, bodies, nb);
    free(bodies);
    free(nodes);
    return rc ? artra_dev_failed() : 0;
}

/* IntegrateSurfaceForce (solid_dynamics.c:50-171): the device sums each
 * body's terms in the reference's node order; the calibration by the
 * interfacial area per node and the equilibrium test run here. */
int artra_hip_surface_force(ArtraHipSolver *s)
{
    if (0 == s->forceBodies) return 0;
    double *sums = malloc((size_t)s->forceBodies * 12 * sizeof(double));
    if (!sums) return artra_set_error(1, "out of host memory");
    if (artra_dev_surface_force(s->dev, sums)) {
        free(sums);
        return artra_dev_failed();
    }
    const Real percent = FLT_EPSILON * FLT_EPSILON;
    for (int b = 0; b < s->forceBodies; ++b) {
        const ArtraForceBodyInfo *info = s->force + b;
        Polyhedron *poly = s->space->geo.poly + info->body;
        const double *o = sums + (size_t)b * 12;
        for (int d = 0; d < DIMS; ++d) {
            poly->Fp[d] = o[d];
            poly->Fv[d] = o[3 + d];
            poly->Tt[d] = o[6 + d];
        }
        if (0 == info->lidN || 0 == info->gstN) continue;
        const int gstN = info->gstN;
        Real ds = poly->area / info->lidN;
        Real fvar[3] = {o[9], o[10], o[11]};
        fvar[2] = (fvar[2] - fvar[1] * fvar[1] / gstN) / gstN;
        fvar[1] = fvar[1] / gstN + fvar[0];
        if (percent * fvar[1] * fvar[1] > fvar[2]) ds = 0.0;
        for (int d = 0; d < DIMS; ++d) {
            poly->Fp[d] = -poly->Fp[d] * ds;
            poly->Fv[d] = -poly->Fv[d] * ds;
            poly->Tt[d] = -poly->Tt[d] * ds;
        }
    }
    free(sums);
    return 0;
}

/* ------------------------------------------------------------------- probes */
static FILE *OpenFile(const char *name, const char *mode)
{
    FILE *fp = fopen(name, mode);
    if (!fp) artra_set_error(1, "cannot open %s", name);
    return fp;
}

/* Primitive values of level TO at the listed nodes (prim[count][6]). */
static int Gather(ArtraHipSolver *s, const int *nodes, int count, double *prim)
{
    return artra_dev_gather_primitive(s->dev, nodes, count, prim) ? artra_dev_failed() : 0;
}

/* Node of a point, confined to the physical range (data_probe.c:46-49). */
static int ProbeNode(const Partition *part, const Real p[DIMS], int ijk[DIMS])
{
    for (int d = 0; d < DIMS; ++d) {
        ijk[d] = artra_confine(artra_map_node(p[d], part->domain[d][MIN], part->dd[d], part->ng[d]),
                               part->ns[PHY][d][MIN], part->ns[PHY][d][MAX]);
    }
    return artra_index(ijk[Z], ijk[Y], ijk[X], part->n[Y], part->n[X]);
}

/* WritePointProbeData (data_probe.c:22-57). */
static int PointProbes(ArtraHipSolver *s, const Time *time)
{
    const int np = time->dataN[PROPT];
    if (0 >= np) return 0;
    const Partition *part = &s->space->part;
    int *nodes = malloc((size_t)np * sizeof(int));
    double *prim = malloc((size_t)np * 6 * sizeof(double));
    int rc = (!nodes || !prim) ? artra_set_error(1, "out of host memory") : 0;
    for (int n = 0; !rc && n < np; ++n) {
        int ijk[DIMS];
        nodes[n] = ProbeNode(part, time->pp[n], ijk);
    }
    if (!rc) rc = Gather(s, nodes, np, prim);
    for (int n = 0; !rc && n < np; ++n) {
        char name[64];
        snprintf(name, sizeof name, "%s%03d.csv", "point_probe_", n + 1);
        FILE *fp = OpenFile(name, "a");
        if (!fp) {
            rc = 1;
            break;
        }
        if (0 == time->stepC) fprintf(fp, "# time, rho, u, v, w, p, T\n");
        const double *Uo = prim + (size_t)n * 6;
        fprintf(fp, "%.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g\n", time->now, Uo[0], Uo[1], Uo[2], Uo[3], Uo[4], Uo[5]);
        fclose(fp);
    }
    free(nodes);
    free(prim);
    return rc;
}

/* WriteLineProbeData (data_probe.c:58-115): the samples p1 + m dl, m =
 * 0..stepN, at their nearest physical nodes, a node repeated by the next
 * sample written once. */
static int LineProbes(ArtraHipSolver *s, const Time *time)
{
    const int nl = time->dataN[PROLN];
    if (0 >= nl) return 0;
    const Partition *part = &s->space->part;
    int rc = 0;
    for (int n = 0; !rc && n < nl; ++n) {
        const Real *lp = time->lp[n];
        const Real p1[DIMS] = {lp[0], lp[1], lp[2]};
        const Real p2[DIMS] = {lp[3], lp[4], lp[5]};
        const int stepN = ((int)(lp[6] - 1) > 1) ? (int)(lp[6] - 1) : 1;
        const Real dl[DIMS] = {(p2[X] - p1[X]) / (Real)stepN, (p2[Y] - p1[Y]) / (Real)stepN,
                               (p2[Z] - p1[Z]) / (Real)stepN};
        int *nodes = malloc((size_t)(stepN + 1) * sizeof(int));
        int (*ijk)[DIMS] = malloc((size_t)(stepN + 1) * sizeof *ijk);
        double *prim = malloc((size_t)(stepN + 1) * 6 * sizeof(double));
        if (!nodes || !ijk || !prim) rc = artra_set_error(1, "out of host memory");
        int count = 0, last = -1;
        for (int m = 0; !rc && m <= stepN; ++m) {
            const Real p[DIMS] = {p1[X] + m * dl[X], p1[Y] + m * dl[Y], p1[Z] + m * dl[Z]};
            int q[DIMS];
            const int idx = ProbeNode(part, p, q);
            if (idx == last) continue;
            last = idx;
            nodes[count] = idx;
            for (int d = 0; d < DIMS; ++d) ijk[count][d] = q[d];
            ++count;
        }
        if (!rc) rc = Gather(s, nodes, count, prim);
        if (!rc) {
            char name[64];
            snprintf(name, sizeof name, "%s%03d_%05d.csv", "line_probe_", n + 1, time->stepC);
            FILE *fp = OpenFile(name, "w");
            if (!fp) {
                rc = 1;
            } else {
                fprintf(fp, "# x, y, z, rho, u, v, w, p, T <time=%.6g>\n", time->now);
                for (int q = 0; q < count; ++q) {
                    Real p[DIMS];
                    NodePoint(part, ijk[q][Z], ijk[q][Y], ijk[q][X], p);
                    const double *Uo = prim + (size_t)q * 6;
                    fprintf(fp, "%.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g\n", p[X], p[Y], p[Z], Uo[0],
                            Uo[1], Uo[2], Uo[3], Uo[4], Uo[5]);
                }
                fclose(fp);
            }
        }
        free(nodes);
        free(ijk);
        free(prim);
    }
    return rc;
}

/* WriteCurveProbeData (data_probe.c:116-173): per body, its layer-1 ghost
 * nodes inside its bounding box confined to the physical range, with the
 * boundary point and normal of each. */
static int CurveProbes(ArtraHipSolver *s, const Time *time)
{
    if (0 == time->dataN[PROCV]) return 0;
    const Space *space = s->space;
    const Partition *part = &space->part;
    const Geometry *geo = &space->geo;
    int rc = 0;
    for (int n = 0; !rc && n < geo->totN; ++n) {
        const Polyhedron *poly = geo->poly + n;
        int box[DIMS][LIMIT];
        BodyBox(part, poly, PHY, box);
        const long span = (long)(box[X][MAX] - box[X][MIN]) * (box[Y][MAX] - box[Y][MIN]) * (box[Z][MAX] - box[Z][MIN]);
        const long cap = (span > 0) ? span : 1;
        int *nodes = malloc((size_t)cap * sizeof(int));
        Real (*geom)[2 * DIMS] = malloc((size_t)cap * sizeof *geom);
        double *prim = malloc((size_t)cap * 6 * sizeof(double));
        if (!nodes || !geom || !prim) rc = artra_set_error(1, "out of host memory");
        int count = 0;
        for (int k = box[Z][MIN]; !rc && k < box[Z][MAX]; ++k) {
            for (int j = box[Y][MIN]; j < box[Y][MAX]; ++j) {
                for (int i = box[X][MIN]; i < box[X][MAX]; ++i) {
                    const int idx = artra_index(k, j, i, part->n[Y], part->n[X]);
                    const Node *nd = space->node + idx;
                    if (1 != nd->gst || n + 1 != nd->did) continue;
                    Real pG[DIMS], pO[DIMS], pI[DIMS], N[DIMS];
                    NodePoint(part, k, j, i, pG);
                    artra_compute_geometric_data(pG, nd->fid, poly, pO, pI, N);
                    nodes[count] = idx;
                    for (int d = 0; d < DIMS; ++d) {
                        geom[count][d] = pO[d];
                        geom[count][DIMS + d] = N[d];
                    }
                    ++count;
                }
            }
        }
        if (!rc) rc = Gather(s, nodes, count, prim);
        if (!rc) {
            char name[64];
            snprintf(name, sizeof name, "%s%03d_%05d.csv", "curve_probe_", n + 1, time->stepC);
            FILE *fp = OpenFile(name, "w");
            if (!fp) {
                rc = 1;
            } else {
                fprintf(fp, "# x, y, z, Nx, Ny, Nz, rho, u, v, w, p, T <time=%.6g>\n", time->now);
                for (int q = 0; q < count; ++q) {
                    const Real *g = geom[q];
                    const double *Uo = prim + (size_t)q * 6;
                    fprintf(fp, "%.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g\n", g[0], g[1],
                            g[2], g[3], g[4], g[5], Uo[0], Uo[1], Uo[2], Uo[3], Uo[4], Uo[5]);
                }
                fclose(fp);
            }
        }
        free(nodes);
        free(geom);
        free(prim);
    }
    return rc;
}

/* WriteSurfaceForceData (data_probe.c:174-197): the bodies' current
 * surface force as the last artra_hip_surface_force left it. */
static int ForceProbes(ArtraHipSolver *s, const Time *time)
{
    if (0 == time->dataN[PROFC]) return 0;
    const Geometry *geo = &s->space->geo;
    for (int n = 0; n < geo->totN; ++n) {
        const Polyhedron *poly = geo->poly + n;
        char name[64];
        snprintf(name, sizeof name, "%s%03d.csv", "surface_force_", n + 1);
        FILE *fp = OpenFile(name, "a");
        if (!fp) return 1;
        if (0 == time->stepC) {
            fprintf(fp, "# time, Fpx, Fpy, Fpz, Fvx, Fvy, Fvz, Ttx, Tty, Ttz <model.mid=%d>\n", s->model->mid);
        }
        fprintf(fp, "%.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g\n", time->now, poly->Fp[X],
                poly->Fp[Y], poly->Fp[Z], poly->Fv[X], poly->Fv[Y], poly->Fv[Z], poly->Tt[X], poly->Tt[Y],
                poly->Tt[Z]);
        fclose(fp);
    }
    return 0;
}

int artra_hip_write_probe_data(ArtraHipSolver *s, const Time *time, int kind)
{
    switch (kind) {
    case PROPT:
        return PointProbes(s, time);
    case PROLN:
        return LineProbes(s, time);
    case PROCV:
        return CurveProbes(s, time);
    case PROFC:
        return ForceProbes(s, time);
    default:
        return artra_set_error(1, "artra_hip_write_probe_data: kind %d is not a probe (0..3)", kind);
    }
}
