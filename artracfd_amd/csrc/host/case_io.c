/*
 * case_io.c - error reporting, small numeric helpers, case-file and STL
 * readers of the host side.
 *
 * The file formats are the solver's: artracfd.case with its twelve
 * mandatory sections and optional regional initialisers
 * (case_loader.c:40-232), the geometry count of artracfd.geo
 * (case_loader.c:233-253), boundary entries (case_loader.c:254-285) and
 * binary STL (stl.c:39-79).  Where the reference stops the process through
 * ShowError (commons.c:104-111), these functions return a non-zero code and
 * leave the message in artra_error_text().
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "artra_host.h"
#include "geomcore.h"

/* ------------------------------------------------------------------ errors */
static _Thread_local char errorText[512];

int artra_set_error(int code, const char *fmt, ...)
{
    va_list args;
    va_start(args, fmt);
    vsnprintf(errorText, sizeof errorText, fmt, args);
    va_end(args);
    return code;
}

const char *artra_error_text(void)
{
    return errorText;
}

/* --------------------------------------------------------------- helpers */
void artra_parse_line(char *line)
{
    /* keep text before '#', drop CR/LF, trim, collapse blank runs to one space */
    char *w = line;
    int pendingBlank = 0;
    for (const char *r = line; *r && *r != '#'; ++r) {
        const char ch = *r;
        if ('\r' == ch || '\n' == ch) continue;
        if (' ' == ch || '\t' == ch) {
            pendingBlank = (w != line);
            continue;
        }
        if (pendingBlank) *w++ = ' ';
        pendingBlank = 0;
        *w++ = ch;
    }
    *w = '\0';
}

int artra_index(int k, int j, int i, int ny, int nx)
{
    return (k * ny + j) * nx + i;
}

/* cfd_commons.c:340-351: nearest node of a coordinate, clamp, node coordinate */
int artra_map_node(Real s, Real smin, Real dds, int ng)
{
    return (int)((s - smin) * dds + 0.5) + ng;
}

int artra_confine(int n, int nmin, int nmax)
{
    const int lo = (nmin > n) ? nmin : n;
    return (nmax - 1 < lo) ? nmax - 1 : lo;
}

Real artra_map_point(int n, Real smin, Real ds, int ng)
{
    return smin + (n - ng) * ds;
}

int artra_in_box(int k, int j, int i, const int box[DIMS][LIMIT])
{
    return (box[Z][MIN] <= k) && (box[Z][MAX] > k) && (box[Y][MIN] <= j) && (box[Y][MAX] > j) &&
           (box[X][MIN] <= i) && (box[X][MAX] > i);
}

Real artra_dot(const Real a[DIMS], const Real b[DIMS])
{
    return a[X] * b[X] + a[Y] * b[Y] + a[Z] * b[Z];
}

Real artra_norm(const Real a[DIMS])
{
    return sqrt(artra_dot(a, a));
}

Real artra_dist2(const Real a[DIMS], const Real b[DIMS])
{
    const Real d[DIMS] = {b[X] - a[X], b[Y] - a[Y], b[Z] - a[Z]};
    return artra_dot(d, d);
}

void artra_cross(const Real a[DIMS], const Real b[DIMS], Real out[DIMS])
{
    out[X] = a[Y] * b[Z] - a[Z] * b[Y];
    out[Y] = a[Z] * b[X] - a[X] * b[Z];
    out[Z] = a[X] * b[Y] - a[Y] * b[X];
}

void artra_normalize(int n, Real by, Real *v)
{
    for (int q = 0; q < n; ++q) v[q] = v[q] / by;
}

/* Tangent pair of a unit normal (cfd_commons.c:431-458): the first tangent
 * zeroes the normal's smallest component (ties resolved toward z, then y). */
void artra_orthogonal_space(const Real N[DIMS], Real Ta[DIMS], Real Tb[DIMS])
{
    GcOrthogonalSpace(N, Ta, Tb);
}

/* Primitive (rho, u, v, w, p, T) from conservative (cfd_commons.c:296-305). */
void artra_map_primitive(Real gamma, Real gasR, const Real U[DIMU], Real Uo[DIMUo])
{
    Uo[0] = U[0];
    Uo[1] = U[1] / U[0];
    Uo[2] = U[2] / U[0];
    Uo[3] = U[3] / U[0];
    Uo[4] = (U[4] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]) * (gamma - 1.0);
    Uo[5] = Uo[4] / (Uo[0] * gasR);
}

/* Conservative from primitive (cfd_commons.c:314-322). */
void artra_map_conservative(Real gamma, const Real Uo[DIMUo], Real U[DIMU])
{
    U[0] = Uo[0];
    U[1] = Uo[0] * Uo[1];
    U[2] = Uo[0] * Uo[2];
    U[3] = Uo[0] * Uo[3];
    U[4] = 0.5 * Uo[0] * (Uo[1] * Uo[1] + Uo[2] * Uo[2] + Uo[3] * Uo[3]) + Uo[4] / (gamma - 1.0);
}

/* ------------------------------------------------------------ case loader */
typedef struct {
    FILE *fp;
    const char *fname;
    int line;
} Reader;

static int NextLine(Reader *rd, char *buf, int size)
{
    if (NULL == fgets(buf, size, rd->fp)) {
        buf[0] = '\0';
        return artra_set_error(1, "%s: unexpected end of file after line %d", rd->fname, rd->line);
    }
    ++rd->line;
    return 0;
}

/* One line holding exactly `want` comma-separated values of format fmt. */
static int ReadValues(Reader *rd, int want, const char *fmt, ...)
{
    char buf[STR];
    if (NextLine(rd, buf, sizeof buf)) return 1;
    va_list args;
    va_start(args, fmt);
    const int got = vsscanf(buf, fmt, args);
    va_end(args);
    if (got != want) {
        return artra_set_error(1, "%s:%d: expected %d value(s) of '%s', read '%s'", rd->fname, rd->line, want,
                               fmt, buf);
    }
    return 0;
}

static int ReadExpressions(Reader *rd, char dst[VARIC][VARSTR])
{
    char buf[STR];
    for (int m = 0; m < VARIC; ++m) {
        if (NextLine(rd, buf, sizeof buf)) return 1;
        artra_parse_line(buf);
        if (strlen(buf) >= VARSTR) {
            return artra_set_error(1, "%s:%d: expression longer than %d characters", rd->fname, rd->line,
                                   VARSTR - 1);
        }
        strncpy(dst[m], buf, VARSTR);
    }
    return 0;
}

static int ReadBoundary(Reader *rd, Partition *part, int p)
{
    char buf[STR];
    if (NextLine(rd, buf, sizeof buf)) return 1;
    artra_parse_line(buf);
    if (0 == strcmp(buf, "inflow")) {
        part->typeBC[p] = INFLOW;
        for (int m = 0; m < VARBC - 1; ++m) {
            if (ReadValues(rd, 1, "%lg", &part->varBC[p][m])) return 1;
        }
        return 0;
    }
    if (0 == strcmp(buf, "outflow")) {
        part->typeBC[p] = OUTFLOW;
        return 0;
    }
    if (0 == strcmp(buf, "slip wall")) {
        part->typeBC[p] = SLIPWALL;
        return ReadValues(rd, 1, "%lg", &part->varBC[p][VARBC - 1]);
    }
    if (0 == strcmp(buf, "noslip wall")) {
        part->typeBC[p] = NOSLIPWALL;
        return ReadValues(rd, 1, "%lg", &part->varBC[p][VARBC - 1]);
    }
    if (0 == strcmp(buf, "periodic")) {
        part->typeBC[p] = PERIODIC;
        return 0;
    }
    return artra_set_error(1, "%s:%d: unidentified boundary type '%s' for face %d", rd->fname, rd->line, buf, p);
}

static int ReadTriple(Reader *rd, Real *a, Real *b, Real *c)
{
    return ReadValues(rd, 3, "%lg, %lg, %lg", a, b, c);
}

static int ReadRegion(Reader *rd, Partition *part, int type, int twoPoints, int radius)
{
    if (part->nIC >= NIC) {
        return artra_set_error(1, "%s:%d: more than %d initialisers", rd->fname, rd->line, NIC);
    }
    const int n = part->nIC;
    part->typeIC[n] = type;
    if (ReadTriple(rd, part->posIC[n] + 0, part->posIC[n] + 1, part->posIC[n] + 2)) return 1;
    if (twoPoints && ReadTriple(rd, part->posIC[n] + 3, part->posIC[n] + 4, part->posIC[n] + 5)) return 1;
    if (radius && ReadValues(rd, 1, "%lg", part->posIC[n] + 6)) return 1;
    if (ReadExpressions(rd, part->varIC[n])) return 1;
    ++part->nIC;
    return 0;
}

static int PathJoin(char *out, size_t size, const char *dir, const char *name)
{
    if ((size_t)snprintf(out, size, "%s/%s", dir, name) >= size) {
        return artra_set_error(1, "path too long: %s/%s", dir, name);
    }
    return 0;
}

static int CheckCase(const Time *time, const Space *space, const Model *model)
{
    const Partition *part = &space->part;
    for (int s = 0; s < DIMS; ++s) {
        if (!(part->domain[s][MAX] - part->domain[s][MIN] > 0.0)) {
            return artra_set_error(1, "domain region should have max > min");
        }
        if (part->m[s] < 1) return artra_set_error(1, "mesh number should be positive");
    }
    if ((time->restart < 0) || !(time->end > 0.0) || !(time->numCFL > 0.0)) {
        return artra_set_error(1, "values in time section should not be negative");
    }
    if (time->restart != 0) {
        return artra_set_error(2, "restart from a field file is not supported by this library");
    }
    if ((model->tScheme < 0) || (model->tScheme > 1) || (model->sScheme < 0) || (model->sScheme > 1) ||
        (model->multidim < 0) || (model->multidim > 1) || (model->jacobMean < 0) || (model->jacobMean > 1) ||
        (model->fluxSplit < 0) || (model->fluxSplit > 1) || (model->psi < 0)) {
        return artra_set_error(1, "values in numerical section out of range");
    }
    if (model->mid < 0) return artra_set_error(1, "material type should not be negative");
    if (!(model->refL > 0.0) || !(model->refRho > 0.0) || !(model->refV > 0.0) || !(model->refT > 0.0)) {
        return artra_set_error(1, "reference values should be positive");
    }
    return 0;
}

int artra_load_case(const char *dir, Time *time, Space *space, Model *model)
{
    Partition *part = &space->part;
    part->typeBC = calloc(NBC, sizeof(*part->typeBC));
    part->N = calloc(NBC, sizeof(*part->N));
    part->varBC = calloc(NBC, sizeof(*part->varBC));
    part->typeIC = calloc(NIC, sizeof(*part->typeIC));
    part->posIC = calloc(NIC, sizeof(*part->posIC));
    part->varIC = calloc(NIC, sizeof(*part->varIC));
    if (!part->typeBC || !part->N || !part->varBC || !part->typeIC || !part->posIC || !part->varIC) {
        return artra_set_error(1, "out of host memory");
    }
    part->proc[X] = part->proc[Y] = part->proc[Z] = 1;
    part->procN = 1;
    char path[4096];
    if (PathJoin(path, sizeof path, dir, "artracfd.case")) return 1;
    Reader rd = {fopen(path, "r"), "artracfd.case", 0};
    if (NULL == rd.fp) return artra_set_error(1, "failed to open file: %s", path);
    char buf[STR];
    int nentry = 0;
    int rc = 0;
    static const char *const faces[6] = {"west", "east", "south", "north", "front", "back"};
    while (0 == rc && NULL != fgets(buf, sizeof buf, rd.fp)) {
        ++rd.line;
        artra_parse_line(buf);
        if (0 == strcmp(buf, "space begin")) {
            ++nentry;
            rc = ReadTriple(&rd, &part->domain[X][MIN], &part->domain[Y][MIN], &part->domain[Z][MIN]);
            if (!rc) rc = ReadTriple(&rd, &part->domain[X][MAX], &part->domain[Y][MAX], &part->domain[Z][MAX]);
            if (!rc) rc = ReadValues(&rd, 3, "%d, %d, %d", &part->m[X], &part->m[Y], &part->m[Z]);
            continue;
        }
        if (0 == strcmp(buf, "time begin")) {
            ++nentry;
            rc = ReadValues(&rd, 1, "%d", &time->restart);
            if (!rc) rc = ReadValues(&rd, 1, "%lg", &time->end);
            if (!rc) rc = ReadValues(&rd, 1, "%lg", &time->numCFL);
            if (!rc) rc = ReadValues(&rd, 1, "%d", &time->stepN);
            if (!rc) rc = ReadValues(&rd, 1, "%d", &time->dataW[PROSD]);
            if (!rc) rc = ReadValues(&rd, 1, "%d", &time->dataStreamer);
            continue;
        }
        if (0 == strcmp(buf, "numerical begin")) {
            ++nentry;
            int *dst[7] = {&model->tScheme, &model->sScheme, &model->multidim, &model->jacobMean,
                           &model->fluxSplit, &model->psi, &model->ibmLayer};
            for (int q = 0; q < 7 && !rc; ++q) rc = ReadValues(&rd, 1, "%d", dst[q]);
            continue;
        }
        if (0 == strcmp(buf, "material begin")) {
            ++nentry;
            rc = ReadValues(&rd, 1, "%d", &model->mid);
            if (!rc) rc = ReadValues(&rd, 1, "%lg", &model->refMu);
            if (!rc) rc = ReadValues(&rd, 1, "%d", &model->gState);
            if (!rc) rc = ReadTriple(&rd, &model->g[X], &model->g[Y], &model->g[Z]);
            continue;
        }
        if (0 == strcmp(buf, "reference begin")) {
            ++nentry;
            Real *dst[4] = {&model->refL, &model->refRho, &model->refV, &model->refT};
            for (int q = 0; q < 4 && !rc; ++q) rc = ReadValues(&rd, 1, "%lg", dst[q]);
            continue;
        }
        if (0 == strcmp(buf, "initialization begin")) {
            ++nentry;
            part->nIC = 0; /* the global initialiser always comes first */
            part->typeIC[0] = ICGLOBAL;
            rc = ReadExpressions(&rd, part->varIC[0]);
            part->nIC = 1;
            continue;
        }
        int matched = 0;
        for (int f = 0; f < 6; ++f) {
            char head[64];
            snprintf(head, sizeof head, "%s boundary begin", faces[f]);
            if (0 == strcmp(buf, head)) {
                ++nentry;
                rc = ReadBoundary(&rd, part, PWB + f);
                matched = 1;
                break;
            }
        }
        if (matched) continue;
        if (0 == strcmp(buf, "plane initialization begin")) {
            rc = ReadRegion(&rd, part, ICPLANE, 1, 0);
        } else if (0 == strcmp(buf, "sphere initialization begin")) {
            rc = ReadRegion(&rd, part, ICSPHERE, 0, 1);
        } else if (0 == strcmp(buf, "box initialization begin")) {
            rc = ReadRegion(&rd, part, ICBOX, 1, 0);
        } else if (0 == strcmp(buf, "cylinder initialization begin")) {
            rc = ReadRegion(&rd, part, ICCYLINDER, 1, 1);
        } else if (0 == strcmp(buf, "probe count begin")) {
            for (int q = 0; q < 4 && !rc; ++q) rc = ReadValues(&rd, 1, "%d", &time->dataN[q]);
            if (!rc && time->dataN[PROPT] > 0) time->pp = calloc(time->dataN[PROPT], sizeof(*time->pp));
            if (!rc && time->dataN[PROLN] > 0) time->lp = calloc(time->dataN[PROLN], sizeof(*time->lp));
        } else if (0 == strcmp(buf, "probe control begin")) {
            for (int q = 0; q < 4 && !rc; ++q) rc = ReadValues(&rd, 1, "%d", &time->dataW[q]);
        } else if (0 == strcmp(buf, "point probe begin")) {
            for (int n = 0; n < time->dataN[PROPT] && !rc; ++n) {
                rc = ReadTriple(&rd, time->pp[n] + 0, time->pp[n] + 1, time->pp[n] + 2);
            }
        } else if (0 == strcmp(buf, "line probe begin")) {
            for (int n = 0; n < time->dataN[PROLN] && !rc; ++n) {
                rc = ReadTriple(&rd, time->lp[n] + 0, time->lp[n] + 1, time->lp[n] + 2);
                if (!rc) rc = ReadTriple(&rd, time->lp[n] + 3, time->lp[n] + 4, time->lp[n] + 5);
                if (!rc) rc = ReadValues(&rd, 1, "%lg", time->lp[n] + 6);
            }
        }
    }
    fclose(rd.fp);
    if (rc) return rc;
    if (12 != nentry) {
        return artra_set_error(1, "missing or repeated sections: artracfd.case, entry: %d", nentry);
    }
    /* geometry count (case_loader.c:233-253) */
    if (PathJoin(path, sizeof path, dir, "artracfd.geo")) return 1;
    Reader geo = {fopen(path, "r"), "artracfd.geo", 0};
    if (NULL == geo.fp) return artra_set_error(1, "failed to open file: %s", path);
    nentry = 0;
    while (NULL != fgets(buf, sizeof buf, geo.fp)) {
        ++geo.line;
        artra_parse_line(buf);
        if (0 == strcmp(buf, "count begin")) {
            ++nentry;
            rc = ReadValues(&geo, 1, "%d", &space->geo.sphN);
            if (!rc) rc = ReadValues(&geo, 1, "%d", &space->geo.stlN);
            break;
        }
    }
    fclose(geo.fp);
    if (rc) return rc;
    if (1 != nentry) return artra_set_error(1, "missing or repeated sections: artracfd.geo, entry: %d", nentry);
    return CheckCase(time, space, model);
}

/* ---------------------------------------------------------------- STL */
int artra_read_stl(const char *fname, Polyhedron *poly)
{
    FILE *fp = fopen(fname, "rb");
    if (NULL == fp) return artra_set_error(1, "failed to open file: %s", fname);
    char header[80];
    uint32_t count = 0;
    if (1 != fread(header, sizeof header, 1, fp) || 1 != fread(&count, sizeof count, 1, fp)) {
        fclose(fp);
        return artra_set_error(1, "%s: truncated STL header", fname);
    }
    poly->faceN = (int)count;
    poly->facet = calloc(count ? count : 1, sizeof(*poly->facet));
    if (NULL == poly->facet) {
        fclose(fp);
        return artra_set_error(1, "out of host memory");
    }
    for (uint32_t n = 0; n < count; ++n) {
        float v[12];
        uint16_t attribute;
        if (1 != fread(v, sizeof v, 1, fp) || 1 != fread(&attribute, sizeof attribute, 1, fp)) {
            fclose(fp);
            return artra_set_error(1, "%s: truncated STL facet %u", fname, n);
        }
        for (int s = 0; s < DIMS; ++s) {
            poly->facet[n].N[s] = v[s];
            poly->facet[n].v0[s] = v[3 + s];
            poly->facet[n].v1[s] = v[6 + s];
            poly->facet[n].v2[s] = v[9 + s];
        }
    }
    fclose(fp);
    return 0;
}
