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
#include <limits.h>
#include <math.h>
#include <stddef.h>
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
void artra_force_body_box(const Partition *part, const Polyhedron *poly, int box[DIMS][LIMIT])
{
    BodyBox(part, poly, PIN, box);
}

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

/* ------------------------------------------------------------ field output */
/* The field files are formatted from one download of level TO; the
 * "%.6g" text of each array is formatted in parallel chunks (the bulk of
 * the writer's time) and written in order. */
enum { PV_RHO, PV_U, PV_V, PV_W, PV_P, PV_T, PV_DID, PV_FID, PV_LID, PV_GST, PV_VEL, PV_POINTS };

static void NodeValues(const ArtraHipSolver *s, const double *U, int k, int j, int i, int what, double out[3])
{
    const Space *space = s->space;
    const Partition *part = &space->part;
    const long N = s->N;
    const int idx = artra_index(k, j, i, part->n[Y], part->n[X]);
    const double U0 = U[idx], U1 = U[N + idx], U2 = U[2 * N + idx], U3 = U[3 * N + idx], U4 = U[4 * N + idx];
    const Node *nd = space->node + idx;
    switch (what) {
    case PV_RHO: out[0] = U0; break;
    case PV_U: out[0] = U1 / U0; break;
    case PV_V: out[0] = U2 / U0; break;
    case PV_W: out[0] = U3 / U0; break;
    case PV_P: out[0] = (U4 - 0.5 * (U1 * U1 + U2 * U2 + U3 * U3) / U0) * (s->model->gamma - 1.0); break;
    case PV_T: out[0] = (U4 - 0.5 * (U1 * U1 + U2 * U2 + U3 * U3) / U0) / (U0 * s->model->cv); break;
    case PV_DID: out[0] = nd->did; break;
    case PV_FID: out[0] = nd->fid; break;
    case PV_LID: out[0] = nd->lid; break;
    case PV_GST: out[0] = nd->gst; break;
    case PV_VEL:
        out[0] = U1 / U0;
        out[1] = U2 / U0;
        out[2] = U3 / U0;
        break;
    default:
        NodePoint(part, k, j, i, out);
        break;
    }
}

/* Every node of the output range in (k, j, i) order as "%.6g " text (three
 * values per node for the vector and the points). */
static int WriteRangeValues(FILE *fp, const ArtraHipSolver *s, const double *U, int range, int what)
{
    const Partition *part = &s->space->part;
    const int x0 = part->ns[range][X][MIN], nx = part->ns[range][X][MAX] - x0;
    const int y0 = part->ns[range][Y][MIN], ny = part->ns[range][Y][MAX] - y0;
    const int z0 = part->ns[range][Z][MIN], nz = part->ns[range][Z][MAX] - z0;
    const long total = (long)nx * ny * nz;
    const int per = (what >= PV_VEL) ? 3 : 1;
    const long chunk = 1L << 15;
    const long nchunk = (total + chunk - 1) / chunk;
    char **text = calloc((size_t)(nchunk > 0 ? nchunk : 1), sizeof(char *));
    size_t *len = calloc((size_t)(nchunk > 0 ? nchunk : 1), sizeof(size_t));
    if (!text || !len) {
        free(text);
        free(len);
        return artra_set_error(1, "out of host memory");
    }
    int failed = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (long c = 0; c < nchunk; ++c) {
        const long lo = c * chunk, hi = (lo + chunk < total) ? lo + chunk : total;
        const size_t cap = (size_t)(hi - lo) * per * 16 + 1;
        char *buf = malloc(cap);
        if (!buf) {
            failed = 1;
            continue;
        }
        size_t at = 0;
        for (long t = lo; t < hi; ++t) {
            const int i = x0 + (int)(t % nx), j = y0 + (int)((t / nx) % ny), k = z0 + (int)(t / ((long)nx * ny));
            double v[3];
            NodeValues(s, U, k, j, i, what, v);
            for (int q = 0; q < per; ++q) at += (size_t)snprintf(buf + at, cap - at, "%.6g ", v[q]);
        }
        text[c] = buf;
        len[c] = at;
    }
    for (long c = 0; c < nchunk; ++c) {
        if (!failed && text[c]) fwrite(text[c], 1, len[c], fp);
        free(text[c]);
    }
    free(text);
    free(len);
    return failed ? artra_set_error(1, "out of host memory") : 0;
}

/* The line of the collection file where the next data set is appended:
 * the first line reading "</Collection>" once normalised
 * (commons.c:167-182). */
static void SeekLine(FILE *fp, const char *line)
{
    char str[256];
    fpos_t pos;
    fgetpos(fp, &pos);
    while (fgets(str, sizeof str, fp)) {
        artra_parse_line(str);
        if (0 == strcmp(str, line)) {
            fsetpos(fp, &pos);
            break;
        }
        fgetpos(fp, &pos);
    }
}

/* The transient collection file of one data set and its per-output case
 * file (paraview_writer.c:56-97). */
static int PvCaseFiles(const Time *time, const char *rname, const char *bname, const char *fext)
{
    char name[160];
    FILE *fp;
    if (0 == time->stepC) {
        snprintf(name, sizeof name, "%s.pvd", rname);
        if (!(fp = OpenFile(name, "w"))) return 1;
        fprintf(fp, "<?xml version=\"1.0\"?>\n");
        fprintf(fp, "<VTKFile type=\"Collection\" version=\"1.0\" byte_order=\"%s\">\n", "LittleEndian");
        fprintf(fp, "  <Collection>\n");
        fprintf(fp, "  </Collection>\n");
        fprintf(fp, "</VTKFile>\n");
        fclose(fp);
    }
    snprintf(name, sizeof name, "%s.pvd", bname);
    if (!(fp = OpenFile(name, "w"))) return 1;
    fprintf(fp, "<?xml version=\"1.0\"?>\n");
    fprintf(fp, "<VTKFile type=\"Collection\" version=\"1.0\" byte_order=\"%s\">\n", "LittleEndian");
    fprintf(fp, "  <Collection>\n");
    fprintf(fp, "    <DataSet timestep=\"%.6g\" group=\"\" part=\"0\"\n", time->now);
    fprintf(fp, "             file=\"%s%s\"/>\n", bname, fext);
    fprintf(fp, "  </Collection>\n");
    fprintf(fp, "</VTKFile>\n");
    fprintf(fp, "<!--\n");
    fprintf(fp, "  Time %.6g\n", time->now);
    fprintf(fp, "  Step %d\n", time->stepC);
    fprintf(fp, "-->\n");
    fclose(fp);
    snprintf(name, sizeof name, "%s.pvd", rname);
    if (!(fp = OpenFile(name, "r+"))) return 1;
    SeekLine(fp, "</Collection>");
    fprintf(fp, "    <DataSet timestep=\"%.6g\" group=\"\" part=\"0\"\n", time->now);
    fprintf(fp, "             file=\"%s%s\"/>\n", bname, fext);
    fprintf(fp, "  </Collection>\n");
    fprintf(fp, "</VTKFile>\n");
    fclose(fp);
    return 0;
}

/* WriteStructuredDataParaview (paraview_writer.c:31-197): the output range
 * (ghost layers of the physical boundary included) as a VTK structured
 * grid with ten scalars, the velocity vector and the node coordinates. */
static int PvStructured(ArtraHipSolver *s, const Time *time)
{
    static const char *sca[10] = {"rho", "u", "v", "w", "p", "T", "did", "fid", "lid", "gst"};
    char bname[96], name[128];
    snprintf(bname, sizeof bname, "%s%05d", "field", time->dataC);
    if (PvCaseFiles(time, "field", bname, ".vts")) return 1;
    if (artra_dev_download(s->dev, TO, s->staging)) return artra_dev_failed();
    const Partition *part = &s->space->part;
    const int ne[DIMS] = {part->ns[PIO][X][MAX] - part->ns[PIO][X][MIN] - 1,
                          part->ns[PIO][Y][MAX] - part->ns[PIO][Y][MIN] - 1,
                          part->ns[PIO][Z][MAX] - part->ns[PIO][Z][MIN] - 1};
    snprintf(name, sizeof name, "%s%s", bname, ".vts");
    FILE *fp = OpenFile(name, "w");
    if (!fp) return 1;
    int rc = 0;
    fprintf(fp, "<?xml version=\"1.0\"?>\n");
    fprintf(fp, "<VTKFile type=\"StructuredGrid\" version=\"1.0\" byte_order=\"%s\">\n", "LittleEndian");
    fprintf(fp, "  <StructuredGrid WholeExtent=\"%d %d %d %d %d %d\">\n", 0, ne[X], 0, ne[Y], 0, ne[Z]);
    fprintf(fp, "    <Piece Extent=\"%d %d %d %d %d %d\">\n", 0, ne[X], 0, ne[Y], 0, ne[Z]);
    fprintf(fp, "      <PointData>\n");
    for (int q = 0; q < 10 && !rc; ++q) {
        fprintf(fp, "        <DataArray type=\"%s\" Name=\"%s\" format=\"ascii\">\n", "Float32", sca[q]);
        fprintf(fp, "          ");
        rc = WriteRangeValues(fp, s, s->staging, PIO, q);
        fprintf(fp, "\n        </DataArray>\n");
    }
    if (!rc) {
        fprintf(fp, "        <DataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"3\" format=\"ascii\">\n", "Float32",
                "Vel");
        fprintf(fp, "          ");
        rc = WriteRangeValues(fp, s, s->staging, PIO, PV_VEL);
        fprintf(fp, "\n        </DataArray>\n");
    }
    if (!rc) {
        fprintf(fp, "      </PointData>\n");
        fprintf(fp, "      <CellData>\n");
        fprintf(fp, "      </CellData>\n");
        fprintf(fp, "      <Points>\n");
        fprintf(fp, "        <DataArray type=\"%s\" Name=\"points\" NumberOfComponents=\"3\" format=\"ascii\">\n",
                "Float32");
        fprintf(fp, "          ");
        rc = WriteRangeValues(fp, s, s->staging, PIO, PV_POINTS);
        fprintf(fp, "\n        </DataArray>\n");
        fprintf(fp, "      </Points>\n");
        fprintf(fp, "    </Piece>\n");
        fprintf(fp, "  </StructuredGrid>\n");
        fprintf(fp, "</VTKFile>\n");
    }
    fclose(fp);
    return rc;
}

/* WritePolyStateData (data_stream.c:131-152). */
static void PolyState(FILE *fp, int pm, int pn, const Geometry *geo)
{
    for (int n = pm; n < pn; ++n) {
        const Polyhedron *poly = geo->poly + n;
        fprintf(fp, "  %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %d\n",
                poly->O[X], poly->O[Y], poly->O[Z], poly->r, poly->V[TO][X], poly->V[TO][Y], poly->V[TO][Z],
                poly->W[TO][X], poly->W[TO][Y], poly->W[TO][Z], poly->rho, poly->T, poly->cf, poly->area,
                poly->volume, poly->mid);
        fprintf(fp, "  %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g, %.6g\n",
                poly->at[TO][X], poly->at[TO][Y], poly->at[TO][Z], poly->ar[TO][X], poly->ar[TO][Y], poly->ar[TO][Z],
                poly->at[TN][X], poly->at[TN][Y], poly->at[TN][Z], poly->g[X], poly->g[Y], poly->g[Z],
                poly->ar[TN][X], poly->ar[TN][Y], poly->ar[TN][Z], poly->to);
    }
}

/* WritePolyDataParaview for the analytical spheres (paraview_writer.c:216-303). */
static int PvSpheres(const Time *time, const Geometry *geo)
{
    char bname[96], name[128];
    snprintf(bname, sizeof bname, "%s%05d", "geo_sph", time->dataC);
    if (PvCaseFiles(time, "geo_sph", bname, ".vtp")) return 1;
    snprintf(name, sizeof name, "%s%s", bname, ".vtp");
    FILE *fp = OpenFile(name, "w");
    if (!fp) return 1;
    const int pm = 0, pn = geo->sphN;
    fprintf(fp, "<?xml version=\"1.0\"?>\n");
    fprintf(fp, "<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"%s\">\n", "LittleEndian");
    fprintf(fp, "  <PolyData>\n");
    fprintf(fp, "    <Piece NumberOfPoints=\"%d\" NumberOfVerts=\"1\" NumberOfPolys=\"0\">\n", pn - pm);
    fprintf(fp, "      <PointData>\n");
    static const char *sca[2] = {"r", "did"};
    for (int q = 0; q < 2; ++q) {
        fprintf(fp, "        <DataArray type=\"%s\" Name=\"%s\" format=\"ascii\">\n", "Float32", sca[q]);
        fprintf(fp, "          ");
        for (int n = pm; n < pn; ++n) fprintf(fp, "%.6g ", (0 == q) ? geo->poly[n].r : (Real)(n + 1));
        fprintf(fp, "\n        </DataArray>\n");
    }
    fprintf(fp, "        <DataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"3\" format=\"ascii\">\n", "Float32",
            "Vel");
    fprintf(fp, "          ");
    for (int n = pm; n < pn; ++n) {
        const Real *V = geo->poly[n].V[TO];
        fprintf(fp, "%.6g %.6g %.6g ", V[X], V[Y], V[Z]);
    }
    fprintf(fp, "\n        </DataArray>\n");
    fprintf(fp, "      </PointData>\n");
    fprintf(fp, "      <CellData>\n");
    fprintf(fp, "      </CellData>\n");
    fprintf(fp, "      <Points>\n");
    fprintf(fp, "        <DataArray type=\"%s\" Name=\"points\" NumberOfComponents=\"3\" format=\"ascii\">\n", "Float32");
    fprintf(fp, "          ");
    for (int n = pm; n < pn; ++n) {
        const Real *O = geo->poly[n].O;
        fprintf(fp, "%.6g %.6g %.6g ", O[X], O[Y], O[Z]);
    }
    fprintf(fp, "\n        </DataArray>\n");
    fprintf(fp, "      </Points>\n");
    fprintf(fp, "      <Verts>\n");
    fprintf(fp, "        <DataArray type=\"%s\" Name=\"connectivity\" format=\"ascii\">\n", "Int32");
    fprintf(fp, "          ");
    for (int n = pm; n < pn; ++n) fprintf(fp, "%d ", n);
    fprintf(fp, "\n        </DataArray>\n");
    fprintf(fp, "        <DataArray type=\"%s\" Name=\"offsets\" format=\"ascii\">\n", "Int32");
    fprintf(fp, "          ");
    fprintf(fp, "%d ", pn - pm);
    fprintf(fp, "\n        </DataArray>\n");
    fprintf(fp, "      </Verts>\n");
    fprintf(fp, "      <Polys>\n");
    fprintf(fp, "      </Polys>\n");
    fprintf(fp, "    </Piece>\n");
    fprintf(fp, "  </PolyData>\n");
    fprintf(fp, "</VTKFile>\n");
    fprintf(fp, "<!--\n");
    PolyState(fp, pm, pn, geo);
    fprintf(fp, "-->\n");
    fclose(fp);
    return 0;
}

/* WritePolyDataParaview for the triangulated bodies (paraview_writer.c:304-398). */
static int PvPolygons(const Time *time, const Geometry *geo)
{
    char bname[96], name[128];
    snprintf(bname, sizeof bname, "%s%05d", "geo_stl", time->dataC);
    if (PvCaseFiles(time, "geo_stl", bname, ".vtp")) return 1;
    snprintf(name, sizeof name, "%s%s", bname, ".vtp");
    FILE *fp = OpenFile(name, "w");
    if (!fp) return 1;
    const int pm = geo->sphN, pn = geo->totN;
    fprintf(fp, "<?xml version=\"1.0\"?>\n");
    fprintf(fp, "<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"%s\">\n", "LittleEndian");
    fprintf(fp, "  <PolyData>\n");
    for (int m = pm; m < pn; ++m) {
        const Polyhedron *poly = geo->poly + m;
        fprintf(fp, "    <Piece NumberOfPoints=\"%d\" NumberOfVerts=\"0\" NumberOfPolys=\"%d\">\n", poly->vertN,
                poly->faceN);
        fprintf(fp, "      <!--\n");
        fprintf(fp, "        vertN = %d\n", poly->vertN);
        fprintf(fp, "        edgeN = %d\n", poly->edgeN);
        fprintf(fp, "        faceN = %d\n", poly->faceN);
        fprintf(fp, "      -->\n");
        fprintf(fp, "      <PointData>\n");
        fprintf(fp, "      </PointData>\n");
        fprintf(fp, "      <CellData>\n");
        fprintf(fp, "      </CellData>\n");
        fprintf(fp, "      <Points>\n");
        fprintf(fp, "        <DataArray type=\"%s\" Name=\"points\" NumberOfComponents=\"3\" format=\"ascii\">\n",
                "Float32");
        fprintf(fp, "          ");
        for (int n = 0; n < poly->vertN; ++n) fprintf(fp, "%.6g %.6g %.6g ", poly->v[n][X], poly->v[n][Y], poly->v[n][Z]);
        fprintf(fp, "\n        </DataArray>\n");
        fprintf(fp, "      </Points>\n");
        fprintf(fp, "      <Verts>\n");
        fprintf(fp, "      </Verts>\n");
        fprintf(fp, "      <Polys>\n");
        fprintf(fp, "        <DataArray type=\"%s\" Name=\"connectivity\" format=\"ascii\">\n", "Int32");
        fprintf(fp, "          ");
        for (int n = 0; n < poly->faceN; ++n) fprintf(fp, "%d %d %d ", poly->f[n][0], poly->f[n][1], poly->f[n][2]);
        fprintf(fp, "\n        </DataArray>\n");
        fprintf(fp, "        <DataArray type=\"%s\" Name=\"offsets\" format=\"ascii\">\n", "Int32");
        fprintf(fp, "          ");
        for (int n = 0; n < poly->faceN; ++n) fprintf(fp, "%d ", 3 * (n + 1));
        fprintf(fp, "\n        </DataArray>\n");
        fprintf(fp, "      </Polys>\n");
        fprintf(fp, "    </Piece>\n");
    }
    fprintf(fp, "  </PolyData>\n");
    fprintf(fp, "</VTKFile>\n");
    fprintf(fp, "<!--\n");
    PolyState(fp, pm, pn, geo);
    fprintf(fp, "-->\n");
    fclose(fp);
    return 0;
}

/* WriteStateData (data_stream.c:115-125). */
static int StateLog(const Time *time)
{
    FILE *fp = OpenFile("artracfd.log", "w");
    if (!fp) return 1;
    fprintf(fp, "#------------------------------------------------------------------------------\n");
    fprintf(fp, "%d  # completion checkpoint\n", (time->now == time->end) || (time->stepC == time->stepN));
    fprintf(fp, "%d  # data checkpoint\n", time->dataC);
    fprintf(fp, "#------------------------------------------------------------------------------\n");
    fclose(fp);
    return 0;
}

/* ------------------------------------------------------------ EnSight Gold */
/* The EnSight writer's files (ensight_writer.c:34-452): C-binary geometry
 * and variable files with FP32 values and 80-byte string records, and the
 * text case files.  A string record is the writer's 80-byte buffer after a
 * strncpy (zero-padded) or an snprintf (which leaves the bytes past its
 * terminator as they were), so the buffer persists through a file as the
 * reference's does. */
typedef struct {
    char str[80];
    FILE *fp;
    int failed;
} EnFile;

static void EnCopy(EnFile *f, const char *text)
{
    strncpy(f->str, text, sizeof f->str);
    f->failed |= 1 != fwrite(f->str, sizeof f->str, 1, f->fp);
}

static void EnPrint(EnFile *f, const char *fmt, int a, int b, int c, int nargs)
{
    if (1 == nargs) snprintf(f->str, sizeof f->str, fmt, a);
    else snprintf(f->str, sizeof f->str, fmt, a, b, c);
    f->failed |= 1 != fwrite(f->str, sizeof f->str, 1, f->fp);
}

static void EnInt(EnFile *f, int v)
{
    f->failed |= 1 != fwrite(&v, sizeof v, 1, f->fp);
}

static void EnReal(EnFile *f, double v)
{
    const float x = (float)v;
    f->failed |= 1 != fwrite(&x, sizeof x, 1, f->fp);
}

static int EnOpen(EnFile *f, const char *name)
{
    memset(f->str, 0, sizeof f->str);
    f->failed = 0;
    f->fp = OpenFile(name, "wb");
    return f->fp ? 0 : 1;
}

static int EnClose(EnFile *f, const char *name)
{
    const int bad = (0 != fclose(f->fp)) || f->failed;
    return bad ? artra_set_error(1, "writing %s failed", name) : 0;
}

static void EnHeader(EnFile *f)
{
    EnCopy(f, "C Binary");
    EnCopy(f, "Ensight Geometry File");
    EnCopy(f, "Written by ArtraCFD");
    EnCopy(f, "node id off");
    EnCopy(f, "element id off");
}

/* The transient case file (at the first output) and the per-output case
 * file of one data set (ensight_writer.c:61-134). */
static int EnCaseFiles(const Time *time, const char *rname, const char *bname, const char *gtag, const char *vtag,
                       const char *const *sca, int scaN, const char *const *vec, int vecN)
{
    char name[160];
    FILE *fp;
    if (0 == time->stepC) {
        snprintf(name, sizeof name, "%s.case", rname);
        if (!(fp = OpenFile(name, "w"))) return 1;
        fprintf(fp, "FORMAT\n");
        fprintf(fp, "type: ensight gold\n");
        fprintf(fp, "\n");
        fprintf(fp, "GEOMETRY\n");
        fprintf(fp, "model: %s%s.geo\n", rname, gtag);
        fprintf(fp, "\n");
        fprintf(fp, "VARIABLE\n");
        for (int n = 0; n < scaN; ++n) fprintf(fp, "scalar per node:  1  %3s  %s%s.%s\n", sca[n], rname, vtag, sca[n]);
        for (int n = 0; n < vecN; ++n) fprintf(fp, "vector per node:  1  %3s  %s%s.%s\n", vec[n], rname, vtag, vec[n]);
        fprintf(fp, "\n");
        fprintf(fp, "TIME\n");
        fprintf(fp, "time set: 1\n");
        fprintf(fp, "number of steps:          0          \n");
        fprintf(fp, "filename start number:    0\n");
        fprintf(fp, "filename increment:       1\n");
        fprintf(fp, "time values:  ");
        fclose(fp);
    }
    snprintf(name, sizeof name, "%s.case", bname);
    if (!(fp = OpenFile(name, "w"))) return 1;
    fprintf(fp, "FORMAT\n");
    fprintf(fp, "type: ensight gold\n");
    fprintf(fp, "\n");
    fprintf(fp, "GEOMETRY\n");
    if ('\0' == *gtag) fprintf(fp, "model: %s.geo\n", rname);
    else fprintf(fp, "model: %s.geo\n", bname);
    fprintf(fp, "\n");
    fprintf(fp, "VARIABLE\n");
    fprintf(fp, "constant per case:  Time  %.6g\n", time->now);
    fprintf(fp, "constant per case:  Step  %d\n", time->stepC);
    for (int n = 0; n < scaN; ++n) fprintf(fp, "scalar per node:     %3s  %s.%s\n", sca[n], bname, sca[n]);
    for (int n = 0; n < vecN; ++n) fprintf(fp, "vector per node:     %3s  %s.%s\n", vec[n], bname, vec[n]);
    fprintf(fp, "\n");
    fclose(fp);
    snprintf(name, sizeof name, "%s.case", rname);
    if (!(fp = OpenFile(name, "r+"))) return 1;
    {
        char str[256];
        while (fgets(str, sizeof str, fp)) {
            artra_parse_line(str);
            if (0 == strcmp(str, "time set: 1")) {
                fseek(fp, 0, SEEK_CUR);
                break;
            }
        }
    }
    fprintf(fp, "number of steps:          %d", time->dataC + 1);
    fseek(fp, 0, SEEK_END);
    if (0 == time->dataC % 5) fprintf(fp, "\n");
    fprintf(fp, "%.6g ", time->now);
    fclose(fp);
    return 0;
}

/* WriteStructuredDataEnsight (ensight_writer.c:34-246): the output range as
 * one structured block; the geometry file at the first output only. */
static int EnStructured(ArtraHipSolver *s, const Time *time)
{
    static const char *sca[7] = {"rho", "u", "v", "w", "p", "T", "did"};
    static const char *vec[1] = {"Vel"};
    const Partition *part = &s->space->part;
    char bname[96], name[128];
    snprintf(bname, sizeof bname, "%s%05d", "field", time->dataC);
    EnFile f;
    if (0 == time->stepC) {
        if (EnCaseFiles(time, "field", bname, "", "*****", sca, 7, vec, 1)) return 1;
        snprintf(name, sizeof name, "%s.geo", "field");
        if (EnOpen(&f, name)) return 1;
        EnHeader(&f);
        EnCopy(&f, "part");
        EnInt(&f, 1);
        EnPrint(&f, "part %d", PIO, 0, 0, 1);
        EnCopy(&f, "block");
        int ne[DIMS];
        for (int d = 0; d < DIMS; ++d) ne[d] = part->ns[PIO][d][MAX] - part->ns[PIO][d][MIN];
        for (int d = 0; d < DIMS; ++d) EnInt(&f, ne[d]);
        for (int d = 0; d < DIMS; ++d) {
            for (int k = part->ns[PIO][Z][MIN]; k < part->ns[PIO][Z][MAX]; ++k)
                for (int j = part->ns[PIO][Y][MIN]; j < part->ns[PIO][Y][MAX]; ++j)
                    for (int i = part->ns[PIO][X][MIN]; i < part->ns[PIO][X][MAX]; ++i) {
                        const int ijk[DIMS] = {i, j, k};
                        EnReal(&f, artra_map_point(ijk[d], part->domain[d][MIN], part->d[d], part->ng[d]));
                    }
        }
        if (EnClose(&f, name)) return 1;
    } else if (EnCaseFiles(time, "field", bname, "", "*****", sca, 7, vec, 1)) {
        return 1;
    }
    if (artra_dev_download(s->dev, TO, s->staging)) return artra_dev_failed();
    for (int q = 0; q < 8; ++q) {
        snprintf(name, sizeof name, "%s.%s", bname, (q < 7) ? sca[q] : vec[0]);
        if (EnOpen(&f, name)) return 1;
        EnCopy(&f, (q < 7) ? "scalar variable" : "vector variable");
        EnCopy(&f, "part");
        EnInt(&f, 1);
        EnCopy(&f, "block");
        for (int n = 0; n < ((q < 7) ? 1 : 3); ++n) {
            for (int k = part->ns[PIO][Z][MIN]; k < part->ns[PIO][Z][MAX]; ++k)
                for (int j = part->ns[PIO][Y][MIN]; j < part->ns[PIO][Y][MAX]; ++j)
                    for (int i = part->ns[PIO][X][MIN]; i < part->ns[PIO][X][MAX]; ++i) {
                        double v[3];
                        NodeValues(s, s->staging, k, j, i, (q < 7) ? q : PV_VEL, v);
                        EnReal(&f, (q < 7) ? v[0] : v[n]);
                    }
        }
        if (EnClose(&f, name)) return 1;
    }
    return 0;
}

static int EnPolyState(const char *bname, int pm, int pn, const Geometry *geo)
{
    char name[128];
    snprintf(name, sizeof name, "%s.state", bname);
    FILE *fp = OpenFile(name, "w");
    if (!fp) return 1;
    PolyState(fp, pm, pn, geo);
    fclose(fp);
    return 0;
}

/* WritePolyDataEnsight for the analytical spheres (ensight_writer.c:258-344,
 * 396-442). */
static int EnSpheres(const Time *time, const Geometry *geo)
{
    static const char *sca[2] = {"r", "did"};
    static const char *vec[1] = {"Vel"};
    char bname[96], name[128];
    snprintf(bname, sizeof bname, "%s%05d", "geo_sph", time->dataC);
    if (EnCaseFiles(time, "geo_sph", bname, "*****", "*****", sca, 2, vec, 1)) return 1;
    const int pm = 0, pn = geo->sphN;
    EnFile f;
    snprintf(name, sizeof name, "%s.geo", bname);
    if (EnOpen(&f, name)) return 1;
    EnHeader(&f);
    EnCopy(&f, "part");
    EnInt(&f, 1);
    EnPrint(&f, "%d", pn - pm, 0, 0, 1);
    EnCopy(&f, "coordinates");
    EnInt(&f, pn - pm);
    for (int d = 0; d < DIMS; ++d)
        for (int n = pm; n < pn; ++n) EnReal(&f, geo->poly[n].O[d]);
    EnCopy(&f, "point");
    EnInt(&f, pn - pm);
    for (int n = pm, m = 1; n < pn; ++n, ++m) EnInt(&f, m);
    if (EnClose(&f, name)) return 1;
    for (int q = 0; q < 3; ++q) {
        snprintf(name, sizeof name, "%s.%s", bname, (q < 2) ? sca[q] : vec[0]);
        if (EnOpen(&f, name)) return 1;
        EnCopy(&f, (q < 2) ? "scalar variable" : "vector variable");
        EnCopy(&f, "part");
        EnInt(&f, 1);
        EnCopy(&f, "coordinates");
        if (q < 2) {
            for (int n = pm; n < pn; ++n) EnReal(&f, (0 == q) ? geo->poly[n].r : (Real)(n + 1));
        } else {
            for (int d = 0; d < DIMS; ++d)
                for (int n = pm; n < pn; ++n) EnReal(&f, geo->poly[n].V[TO][d]);
        }
        if (EnClose(&f, name)) return 1;
    }
    return EnPolyState(bname, pm, pn, geo);
}

/* WritePolyDataEnsight for the triangulated bodies (ensight_writer.c:345-395). */
static int EnPolygons(const Time *time, const Geometry *geo)
{
    char bname[96], name[128];
    snprintf(bname, sizeof bname, "%s%05d", "geo_stl", time->dataC);
    if (EnCaseFiles(time, "geo_stl", bname, "*****", "*****", NULL, 0, NULL, 0)) return 1;
    const int pm = geo->sphN, pn = geo->totN;
    EnFile f;
    snprintf(name, sizeof name, "%s.geo", bname);
    if (EnOpen(&f, name)) return 1;
    EnHeader(&f);
    for (int p = pm, pnum = 1; p < pn; ++p, ++pnum) {
        const Polyhedron *poly = geo->poly + p;
        EnCopy(&f, "part");
        EnInt(&f, pnum);
        EnPrint(&f, "%d %d %d", poly->vertN, poly->edgeN, poly->faceN, 3);
        EnCopy(&f, "coordinates");
        EnInt(&f, poly->vertN);
        for (int d = 0; d < DIMS; ++d)
            for (int n = 0; n < poly->vertN; ++n) EnReal(&f, poly->v[n][d]);
        EnCopy(&f, "tria3");
        EnInt(&f, poly->faceN);
        for (int n = 0; n < poly->faceN; ++n)
            for (int q = 0; q < POLYN; ++q) EnInt(&f, poly->f[n][q] + 1);
    }
    if (EnClose(&f, name)) return 1;
    return EnPolyState(bname, pm, pn, geo);
}

int artra_hip_write_space_data(ArtraHipSolver *s, const Time *time)
{
    const Geometry *geo = &s->space->geo;
    if (0 == time->dataStreamer) {
        if (PvStructured(s, time)) return 1;
        if (0 != geo->totN) {
            if (0 != geo->sphN && PvSpheres(time, geo)) return 1;
            if (0 != geo->stlN && PvPolygons(time, geo)) return 1;
        }
    } else if (1 == time->dataStreamer) {
        if (EnStructured(s, time)) return 1;
        if (0 != geo->totN) {
            if (0 != geo->sphN && EnSpheres(time, geo)) return 1;
            if (0 != geo->stlN && EnPolygons(time, geo)) return 1;
        }
    } else {
        return artra_set_error(2, "data streamer %d: 0 (ParaView) or 1 (EnSight)", time->dataStreamer);
    }
    return StateLog(time);
}

/* ----------------------------------------------------------- the solver loop */
static int WriteDataKind(ArtraHipSolver *s, const Time *time, int n)
{
    return (PROSD == n) ? artra_hip_write_space_data(s, time) : artra_hip_write_probe_data(s, time, n);
}

/* Solve (solve.c:38-46) after the case's initialisation: the initial
 * outputs of InitializeComputeDomain (initialization.c:51-55) unless the
 * case restarts, then EvolveSolution (solve.c:48-107) with the time steps on
 * the device and the outputs at the loop's cadence.  The device march stops
 * at the first step where an output falls due, so every output is written
 * from exactly the state the reference writes it from. */
/* EvolveSolution with moving bodies: the host takes every step
 * (artra_moving_step: the solid half-steps around the fluid step) and writes
 * the outputs that fall due after it (solve.c:65-105). */
static int SolveMoving(ArtraHipSolver *s, Time *time, const Real dtData[NPROBE], Real rcData[NPROBE], Real tmInt)
{
    Real rcInt = 0.0;
    while (time->now < time->end && time->stepC < time->stepN) {
        Real dt = 0.0;
        if (artra_moving_step(s, time, tmInt, &rcInt, &dt)) return 1;
        for (int n = 0; n < NPROBE; ++n) {
            rcData[n] = rcData[n] + dt;
            if (rcData[n] >= dtData[n] || time->now == time->end || time->stepC == time->stepN) {
                if (PROFC == n && artra_hip_surface_force(s)) return 1;
                if (PROSD == n) ++time->dataC;
                if (WriteDataKind(s, time, n)) return 1;
                rcData[n] = 0.0;
            }
        }
    }
    s->rcInt = rcInt;
    return artra_hip_synchronize(s);
}

int artra_hip_solve(ArtraHipSolver *s, Time *time)
{
    if (0 == time->restart) {
        if (WriteDataKind(s, time, PROPT) || WriteDataKind(s, time, PROFC) || WriteDataKind(s, time, PROSD)) return 1;
    }
    if (0.0 >= time->end - time->now) return 0;
    Real dtData[NPROBE], rcData[NPROBE];
    for (int n = 0; n < NPROBE; ++n) {
        dtData[n] = time->end / (Real)(time->dataW[n]);
        rcData[n] = 0.0;
    }
    const Real tmInt = (INT_MAX == time->dataW[PROSD]) ? time->end : dtData[PROSD];
    if (0 != s->model->psi) return SolveMoving(s, time, dtData, rcData, tmInt);
    Real rcInt = 0.0, lastDt = 0.0;
    double floor[DIMS];
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
    double rec[ARTRA_DEV_MARCH_CHUNK];
    while (time->now < time->end && time->stepC < time->stepN) {
        /* steps up to the next output on the last step's dt (the march
         * stops by itself at the output; this only bounds the steps that
         * would be enqueued past it) */
        int k = ARTRA_DEV_MARCH_CHUNK;
        if (time->stepN - time->stepC < k) k = time->stepN - time->stepC;
        if (lastDt > 0.0) {
            double ahead = (time->end - time->now) / lastDt;
            for (int n = 0; n < NPROBE; ++n) {
                const double to = (dtData[n] - rcData[n]) / lastDt;
                if (to < ahead) ahead = to;
            }
            if (ahead + 1.0 < (double)k) k = (ahead >= 0.0) ? (int)ahead + 1 : 1;
        } else {
            k = 1;
        }
        double clock[ARTRA_DEV_MARCH_CLOCK] = {time->now, time->end, rcInt, tmInt, (double)time->stepC,
                                               (double)time->stepN, 0.0};
        for (int n = 0; n < NPROBE; ++n) {
            clock[7 + n] = rcData[n];
            clock[12 + n] = dtData[n];
        }
        int adv = 0;
        if (artra_dev_march_steps(s->dev, clock, floor, time->numCFL, k, rec, &adv)) return artra_dev_failed();
        time->now = clock[0];
        rcInt = clock[2];
        time->stepC += adv;
        for (int n = 0; n < NPROBE; ++n) rcData[n] = clock[7 + n];
        if (adv > 0) lastDt = rec[adv - 1];
        if (0.0 != clock[6]) {
            /* the outputs due after this step (solve.c:91-104) */
            for (int n = 0; n < NPROBE; ++n) {
                if (rcData[n] >= dtData[n] || time->now == time->end || time->stepC == time->stepN) {
                    if (PROFC == n && artra_hip_surface_force(s)) return 1;
                    if (PROSD == n) ++time->dataC;
                    if (WriteDataKind(s, time, n)) return 1;
                    rcData[n] = 0.0;
                }
            }
        } else if (0 == adv) {
            break;
        }
    }
    s->rcInt = rcInt;
    return artra_hip_synchronize(s);
}

/* ------------------------------------------------------------- FP64 restart */
/* Restart file: "ARTRARS2", int32 n[3], ng[3], stepC, dataC, double now,
 * rcInt, then level TO of every node (ghosts included) as FP64
 * component-major [5][N], then int32 moving (the phase interaction is on).
 * A moving case adds its bodies and the node classification: int32 totN;
 * per body int32 state, faceN, edgeN, vertN, the FP64 state from r to box
 * as Polyhedron lays it out (centre, inertia, velocities, accelerations,
 * forces, bounding box, ...), and for a triangulated body its vertices and
 * face, edge and vertex normals; then int32 did, fid, lid, gst [4][N].
 * Unlike the reference's restart from its "%.6g" text (paraview_reader.c,
 * data_stream.c:154-188), a run continued from it is the uninterrupted run
 * bit for bit. */
static const char kRestartMagic[8] = {'A', 'R', 'T', 'R', 'A', 'R', 'S', '2'};

/* The FP64 part of a body's state: every member from r to box. */
#define BODY_STATE_BYTES (offsetof(Polyhedron, f) - offsetof(Polyhedron, r))

static int BodyIo(Polyhedron *poly, FILE *fp, int write)
{
    int cnt[4] = {poly->state, poly->faceN, poly->edgeN, poly->vertN};
    if (write) {
        if (fwrite(cnt, sizeof cnt, 1, fp) != 1 || fwrite(&poly->r, BODY_STATE_BYTES, 1, fp) != 1) return 1;
    } else {
        if (fread(cnt, sizeof cnt, 1, fp) != 1) return 1;
        if (cnt[1] != poly->faceN || cnt[2] != poly->edgeN || cnt[3] != poly->vertN) return 1;
        if (fread(&poly->r, BODY_STATE_BYTES, 1, fp) != 1) return 1;
        poly->state = cnt[0];
    }
    if (0 >= poly->faceN) return 0;
    Real (*arrays[4])[DIMS] = {poly->v, poly->Nf, poly->Ne, poly->Nv};
    const int rows[4] = {poly->vertN, poly->faceN, poly->edgeN, poly->vertN};
    for (int a = 0; a < 4; ++a) {
        const size_t n = (size_t)rows[a] * DIMS;
        if ((write ? fwrite(arrays[a], sizeof(Real), n, fp) : fread(arrays[a], sizeof(Real), n, fp)) != n) return 1;
    }
    return 0;
}

static int FlagsIo(Space *space, long N, FILE *fp, int write)
{
    int *buf = malloc((size_t)N * sizeof(int));
    if (!buf) return 1;
    int bad = 0;
    for (int f = 0; f < 4 && !bad; ++f) {
        if (write) {
            for (long q = 0; q < N; ++q) {
                const Node *nd = space->node + q;
                buf[q] = (0 == f) ? nd->did : (1 == f) ? nd->fid : (2 == f) ? nd->lid : nd->gst;
            }
            bad = fwrite(buf, sizeof(int), (size_t)N, fp) != (size_t)N;
        } else {
            bad = fread(buf, sizeof(int), (size_t)N, fp) != (size_t)N;
            for (long q = 0; q < N && !bad; ++q) {
                Node *nd = space->node + q;
                if (0 == f) nd->did = buf[q];
                else if (1 == f) nd->fid = buf[q];
                else if (2 == f) nd->lid = buf[q];
                else nd->gst = buf[q];
            }
        }
    }
    free(buf);
    return bad;
}

int artra_hip_write_restart(ArtraHipSolver *s, const Time *time, const char *path)
{
    const Partition *part = &s->space->part;
    Geometry *geo = &s->space->geo;
    if (artra_dev_download(s->dev, TO, s->staging)) return artra_dev_failed();
    FILE *fp = OpenFile(path, "wb");
    if (!fp) return 1;
    const int hdr[8] = {part->n[X], part->n[Y], part->n[Z], part->ng[X], part->ng[Y], part->ng[Z], time->stepC,
                        time->dataC};
    const double clk[2] = {time->now, s->rcInt};
    const int moving = (0 != s->model->psi);
    int ok = fwrite(kRestartMagic, 1, 8, fp) == 8 && fwrite(hdr, sizeof hdr, 1, fp) == 1 &&
             fwrite(clk, sizeof clk, 1, fp) == 1 &&
             fwrite(s->staging, sizeof(double), (size_t)5 * s->N, fp) == (size_t)5 * s->N &&
             fwrite(&moving, sizeof moving, 1, fp) == 1;
    if (ok && moving) {
        ok = fwrite(&geo->totN, sizeof geo->totN, 1, fp) == 1;
        for (int n = 0; ok && n < geo->totN; ++n) ok = !BodyIo(geo->poly + n, fp, 1);
        ok = ok && !FlagsIo(s->space, s->N, fp, 1);
    }
    ok = (0 == fclose(fp)) && ok;
    return ok ? 0 : artra_set_error(1, "writing the restart file %s failed", path);
}

int artra_hip_read_restart(ArtraHipSolver *s, Time *time, const char *path)
{
    const Partition *part = &s->space->part;
    Geometry *geo = &s->space->geo;
    FILE *fp = OpenFile(path, "rb");
    if (!fp) return 1;
    char magic[8];
    int hdr[8], moving = 0, totN = 0;
    double clk[2];
    int ok = fread(magic, 1, 8, fp) == 8 && 0 == memcmp(magic, kRestartMagic, 8) && fread(hdr, sizeof hdr, 1, fp) == 1 &&
             fread(clk, sizeof clk, 1, fp) == 1;
    for (int d = 0; ok && d < DIMS; ++d) ok = hdr[d] == part->n[d] && hdr[DIMS + d] == part->ng[d];
    ok = ok && fread(s->staging, sizeof(double), (size_t)5 * s->N, fp) == (size_t)5 * s->N &&
         fread(&moving, sizeof moving, 1, fp) == 1 && moving == (0 != s->model->psi);
    if (ok && moving) {
        ok = fread(&totN, sizeof totN, 1, fp) == 1 && totN == geo->totN;
        for (int n = 0; ok && n < geo->totN; ++n) ok = !BodyIo(geo->poly + n, fp, 0);
        ok = ok && !FlagsIo(s->space, s->N, fp, 0);
    }
    fclose(fp);
    if (!ok) return artra_set_error(1, "%s is not a restart file of this case", path);
    if (artra_dev_upload(s->dev, TO, s->staging)) return artra_dev_failed();
    Node *node = s->space->node;
    for (long q = 0; q < s->N; ++q) {
        for (int c = 0; c < DIMU; ++c) node[q].U[TO][c] = s->staging[c * s->N + q];
    }
    if (moving) {
        artra_scope_free(&s->scope);
        if (artra_scope_init(s->space, &s->scope) || artra_geo_device_init(s) || artra_hip_refresh_geometry(s, NULL)) {
            return 1;
        }
    }
    time->stepC = hdr[6];
    time->dataC = hdr[7];
    time->now = clk[0];
    s->rcInt = clk[1];
    return 0;
}
