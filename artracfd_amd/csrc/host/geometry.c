/*
 * geometry.c - body geometry on the host: triangulated-surface tables,
 * mass properties, closest-point queries and the node classification that
 * marks bodies and their ghost layers on the grid.
 *
 *   artra_convert_polyhedron            computational_geometry.c:37-136
 *   artra_transform_polyhedron          computational_geometry.c:162-261
 *   artra_compute_geometry_parameters   computational_geometry.c:262-448
 *   artra_point_in_polyhedron           computational_geometry.c:463-506
 *   artra_point_triangle_distance       computational_geometry.c:511-688
 *   closest point and pseudo-normal     computational_geometry.c:689-783
 *   artra_compute_geometric_field       immersed_boundary.c:65-264
 *
 * Bodies are stationary in every supported case, so all of this runs once,
 * before the time loop; the device only consumes the resulting tables.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <limits.h>
#include <string.h>

#include "artra_host.h"
#include "geomcore.h"

static const Real kPi = 3.14159265358979323846;

/* --------------------------------------------------- surface tables */
static int FindOrAddVertex(Polyhedron *poly, const Real v[DIMS])
{
    for (int n = 0; n < poly->vertN; ++n) {
        if (v[X] == poly->v[n][X] && v[Y] == poly->v[n][Y] && v[Z] == poly->v[n][Z]) return n;
    }
    for (int s = 0; s < DIMS; ++s) poly->v[poly->vertN][s] = v[s];
    return poly->vertN++;
}

/* Edges are keyed (larger vertex, smaller vertex); slot 2 holds the first
 * face that uses the edge and slot 3 the second one. */
static void AddEdgeOf(Polyhedron *poly, int a, int b, int face)
{
    const int hi = (a > b) ? a : b;
    const int lo = (a > b) ? b : a;
    for (int n = 0; n < poly->edgeN; ++n) {
        if (hi == poly->e[n][0] && lo == poly->e[n][1]) {
            poly->e[n][3] = face;
            return;
        }
    }
    poly->e[poly->edgeN][0] = hi;
    poly->e[poly->edgeN][1] = lo;
    poly->e[poly->edgeN][2] = face;
    poly->edgeN++;
}

static int EdgeKeyCompare(const void *pa, const void *pb)
{
    return GcEdgeKeyCompare((const int *)pa, (const int *)pb);
}

/* The arrays of a triangulated body as the shared queries take them. */
ArtraPolyView artra_poly_view(const Polyhedron *poly)
{
    ArtraPolyView pv;
    pv.faceN = poly->faceN;
    pv.edgeN = poly->edgeN;
    pv.vertN = poly->vertN;
    pv.f = (const int (*)[POLYN])poly->f;
    pv.e = (const int (*)[EVF])poly->e;
    pv.v = (const Real (*)[DIMS])poly->v;
    pv.Nf = (const Real (*)[DIMS])poly->Nf;
    pv.Ne = (const Real (*)[DIMS])poly->Ne;
    pv.Nv = (const Real (*)[DIMS])poly->Nv;
    return pv;
}

int artra_convert_polyhedron(Polyhedron *poly)
{
    const int fN = poly->faceN;
    poly->f = calloc(fN, sizeof(*poly->f));
    poly->Nf = calloc(fN, sizeof(*poly->Nf));
    poly->e = calloc(3 * fN, sizeof(*poly->e));
    poly->Ne = calloc(3 * fN, sizeof(*poly->Ne));
    poly->v = calloc(3 * fN, sizeof(*poly->v));
    poly->Nv = calloc(3 * fN, sizeof(*poly->Nv));
    if (!poly->f || !poly->Nf || !poly->e || !poly->Ne || !poly->v || !poly->Nv) {
        return artra_set_error(1, "out of host memory for polyhedron tables");
    }
    poly->vertN = 0;
    poly->edgeN = 0;
    for (int n = 0; n < fN; ++n) {
        poly->f[n][0] = FindOrAddVertex(poly, poly->facet[n].v0);
        poly->f[n][1] = FindOrAddVertex(poly, poly->facet[n].v1);
        poly->f[n][2] = FindOrAddVertex(poly, poly->facet[n].v2);
        AddEdgeOf(poly, poly->f[n][0], poly->f[n][1], n);
        AddEdgeOf(poly, poly->f[n][1], poly->f[n][2], n);
        AddEdgeOf(poly, poly->f[n][2], poly->f[n][0], n);
    }
    /* edge keys are unique, so any correct sort gives the reference's order */
    qsort(poly->e, poly->edgeN, sizeof(*poly->e), EdgeKeyCompare);
    free(poly->facet);
    poly->facet = NULL;
    return 0;
}



/* ---------------------------------------------------- transformation */
static void TransformPoints(const Real O[DIMS], const Real scale[DIMS], const Real rot[DIMS][DIMS],
                            const Real offset[DIMS], Real box[DIMS][LIMIT], int count, Real (*v)[DIMS])
{
    for (int n = 0; n < count; ++n) {
        Real t[DIMS];
        for (int s = 0; s < DIMS; ++s) t[s] = (v[n][s] - O[s]) * scale[s];
        for (int s = 0; s < DIMS; ++s) v[n][s] = artra_dot(rot[s], t);
        for (int s = 0; s < DIMS; ++s) {
            v[n][s] = v[n][s] + offset[s] + O[s];
            box[s][MIN] = (box[s][MIN] < v[n][s]) ? box[s][MIN] : v[n][s];
            box[s][MAX] = (box[s][MAX] > v[n][s]) ? box[s][MAX] : v[n][s];
        }
    }
}

static void RotateVectors(const Real rot[DIMS][DIMS], int count, Real (*N)[DIMS])
{
    for (int n = 0; n < count; ++n) {
        const Real t[DIMS] = {N[n][X], N[n][Y], N[n][Z]};
        for (int s = 0; s < DIMS; ++s) N[n][s] = artra_dot(rot[s], t);
    }
}

static Real InertiaAbout(const Real a[DIMS], const Real I[DIMS][DIMS])
{
    return I[X][X] * a[X] * a[X] + I[Y][Y] * a[Y] * a[Y] + I[Z][Z] * a[Z] * a[Z] +
           2.0 * I[X][Y] * a[X] * a[Y] + 2.0 * I[Y][Z] * a[Y] * a[Z] + 2.0 * I[Z][X] * a[Z] * a[X];
}

void artra_transform_polyhedron(const Real O[DIMS], const Real scale[DIMS], const Real angle[DIMS],
                                const Real offset[DIMS], Polyhedron *poly)
{
    const Real Sx = sin(angle[X]), Sy = sin(angle[Y]), Sz = sin(angle[Z]);
    const Real Cx = cos(angle[X]), Cy = cos(angle[Y]), Cz = cos(angle[Z]);
    /* rotation about x, then y, then z */
    const Real rot[DIMS][DIMS] = {
        {Cy * Cz, -Cx * Sz + Sx * Sy * Cz, Sx * Sz + Cx * Sy * Cz},
        {Cy * Sz, Cx * Cz + Sx * Sy * Sz, -Sx * Cz + Cx * Sy * Sz},
        {-Sy, Sx * Cy, Cx * Cy}};
    const Real inv[DIMS][DIMS] = {{rot[0][0], rot[1][0], rot[2][0]},
                                  {rot[0][1], rot[1][1], rot[2][1]},
                                  {rot[0][2], rot[1][2], rot[2][2]}};
    const Real h = 1.0 / sqrt(2.0);
    const Real axes[6][DIMS] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {h, h, 0}, {0, h, h}, {h, 0, h}};
    for (int s = 0; s < DIMS; ++s) {
        poly->box[s][MIN] = FLT_MAX;
        poly->box[s][MAX] = FLT_MIN;
    }
    TransformPoints(O, scale, rot, offset, poly->box, poly->vertN, poly->v);
    RotateVectors(rot, poly->faceN, poly->Nf);
    RotateVectors(rot, poly->edgeN, poly->Ne);
    RotateVectors(rot, poly->vertN, poly->Nv);
    Real I[6];
    for (int n = 0; n < 6; ++n) {
        const Real a[DIMS] = {artra_dot(inv[X], axes[n]), artra_dot(inv[Y], axes[n]), artra_dot(inv[Z], axes[n])};
        I[n] = InertiaAbout(a, (const Real (*)[DIMS])poly->I);
    }
    poly->I[X][X] = I[0];  poly->I[X][Y] = -I[3]; poly->I[X][Z] = -I[5];
    poly->I[Y][X] = -I[3]; poly->I[Y][Y] = I[1];  poly->I[Y][Z] = -I[4];
    poly->I[Z][X] = -I[5]; poly->I[Z][Y] = -I[4]; poly->I[Z][Z] = I[2];
    Real Oc[1][DIMS] = {{poly->O[X], poly->O[Y], poly->O[Z]}};
    TransformPoints(O, scale, rot, offset, poly->box, 1, Oc);
    for (int s = 0; s < DIMS; ++s) poly->O[s] = Oc[0][s];
}

/* ------------------------------------------------------ mass properties */
static void SphereParameters(int collapse, Polyhedron *poly)
{
    for (int s = 0; s < DIMS; ++s) {
        poly->box[s][MIN] = poly->O[s] - poly->r;
        poly->box[s][MAX] = poly->O[s] + poly->r;
    }
    Real k;
    if (COLLAPSEN == collapse) {
        poly->area = 4.0 * kPi * poly->r * poly->r;
        poly->volume = (4.0 / 3.0) * kPi * poly->r * poly->r * poly->r;
        k = 0.4;
    } else {
        poly->area = 2.0 * kPi * poly->r;
        poly->volume = kPi * poly->r * poly->r;
        k = 0.5;
    }
    k = k * poly->r * poly->r * poly->volume;
    memset(poly->I, 0, sizeof poly->I);
    poly->I[X][X] = poly->I[Y][Y] = poly->I[Z][Z] = k;
}

/* Surface-integral mass properties (Eberly, "Polyhedral mass properties")
 * and angle-weighted vertex pseudo-normals (Baerentzen & Aanaes 2005). */
static void PolyhedronParameters(int collapse, Polyhedron *poly)
{
    Real v0[DIMS], v1[DIMS], v2[DIMS], e01[DIMS], e02[DIMS], Nf[DIMS], ang[DIMS];
    Real O[DIMS] = {0.0}, area = 0.0, volume = 0.0, I[6] = {0.0};
    Real f[DIMS][DIMS], g[DIMS][DIMS], box[LIMIT][DIMS];
    for (int s = 0; s < DIMS; ++s) {
        box[MIN][s] = FLT_MAX;
        box[MAX][s] = FLT_MIN;
    }
    memset(poly->Nv, 0, poly->vertN * sizeof(*poly->Nv));
    for (int n = 0; n < poly->vertN; ++n) {
        for (int s = 0; s < DIMS; ++s) {
            box[MIN][s] = (box[MIN][s] < poly->v[n][s]) ? box[MIN][s] : poly->v[n][s];
            box[MAX][s] = (box[MAX][s] > poly->v[n][s]) ? box[MAX][s] : poly->v[n][s];
        }
    }
    const ArtraPolyView pv = artra_poly_view(poly);
    for (int n = 0; n < poly->faceN; ++n) {
        GcTriangleOf(n, &pv, v0, v1, v2, e01, e02);
        artra_cross(e01, e02, Nf);
        for (int s = 0; s < DIMS; ++s) {
            const Real t0 = v0[s] + v1[s];
            f[0][s] = t0 + v2[s];
            const Real t1 = v0[s] * v0[s];
            const Real t2 = t1 + v1[s] * t0;
            f[1][s] = t2 + v2[s] * f[0][s];
            f[2][s] = v0[s] * t1 + v1[s] * t2 + v2[s] * f[1][s];
            g[0][s] = f[1][s] + v0[s] * (f[0][s] + v0[s]);
            g[1][s] = f[1][s] + v1[s] * (f[0][s] + v1[s]);
            g[2][s] = f[1][s] + v2[s] * (f[0][s] + v2[s]);
        }
        area = area + artra_norm(Nf);
        volume = volume + Nf[X] * f[0][X];
        for (int s = 0; s < DIMS; ++s) O[s] = O[s] + Nf[s] * f[1][s];
        I[0] = I[0] + Nf[X] * f[2][X];
        I[1] = I[1] + Nf[Y] * f[2][Y];
        I[2] = I[2] + Nf[Z] * f[2][Z];
        I[3] = I[3] + Nf[X] * (v0[Y] * g[0][X] + v1[Y] * g[1][X] + v2[Y] * g[2][X]);
        I[4] = I[4] + Nf[Y] * (v0[Z] * g[0][Y] + v1[Z] * g[1][Y] + v2[Z] * g[2][Y]);
        I[5] = I[5] + Nf[Z] * (v0[X] * g[0][Z] + v1[X] * g[1][Z] + v2[X] * g[2][Z]);
        artra_normalize(DIMS, artra_norm(Nf), Nf);
        /* interior angles from the law of cosines */
        const Real e12[DIMS] = {v2[X] - v1[X], v2[Y] - v1[Y], v2[Z] - v1[Z]};
        const Real l[DIMS] = {artra_dot(e01, e01), artra_dot(e02, e02), artra_dot(e12, e12)};
        ang[0] = acos((l[0] + l[1] - l[2]) / (2.0 * sqrt(l[0] * l[1])));
        ang[1] = acos((l[0] + l[2] - l[1]) / (2.0 * sqrt(l[0] * l[2])));
        ang[2] = kPi - ang[0] - ang[1];
        for (int v = 0; v < POLYN; ++v) {
            for (int s = 0; s < DIMS; ++s) {
                poly->Nv[poly->f[n][v]][s] = poly->Nv[poly->f[n][v]][s] + ang[v] * Nf[s];
            }
        }
        for (int s = 0; s < DIMS; ++s) poly->Nf[n][s] = Nf[s];
    }
    area = area * (1.0 / 2.0);
    volume = volume * (1.0 / 6.0);
    for (int s = 0; s < DIMS; ++s) O[s] = O[s] * (1.0 / 24.0);
    for (int q = 0; q < 3; ++q) I[q] = I[q] * (1.0 / 60.0);
    for (int q = 3; q < 6; ++q) I[q] = I[q] * (1.0 / 120.0);
    for (int s = 0; s < DIMS; ++s) O[s] = O[s] / volume;
    poly->area = (COLLAPSEN == collapse) ? area : area - 2.0 * volume;
    poly->volume = volume;
    for (int s = 0; s < DIMS; ++s) poly->O[s] = O[s];
    poly->I[X][X] = I[1] + I[2] - volume * (O[Y] * O[Y] + O[Z] * O[Z]);
    poly->I[X][Y] = -I[3] + volume * O[X] * O[Y];
    poly->I[X][Z] = -I[5] + volume * O[Z] * O[X];
    poly->I[Y][X] = poly->I[X][Y];
    poly->I[Y][Y] = I[0] + I[2] - volume * (O[Z] * O[Z] + O[X] * O[X]);
    poly->I[Y][Z] = -I[4] + volume * O[Y] * O[Z];
    poly->I[Z][X] = poly->I[X][Z];
    poly->I[Z][Y] = poly->I[Y][Z];
    poly->I[Z][Z] = I[0] + I[1] - volume * (O[X] * O[X] + O[Y] * O[Y]);
    for (int s = 0; s < DIMS; ++s) {
        poly->box[s][MIN] = box[MIN][s];
        poly->box[s][MAX] = box[MAX][s];
    }
    poly->r = sqrt(artra_dist2(box[MIN], box[MAX]));
    for (int n = 0; n < poly->vertN; ++n) artra_normalize(DIMS, artra_norm(poly->Nv[n]), poly->Nv[n]);
    for (int n = 0; n < poly->edgeN; ++n) {
        for (int s = 0; s < DIMS; ++s) poly->Ne[n][s] = poly->Nf[poly->e[n][2]][s] + poly->Nf[poly->e[n][3]][s];
        artra_normalize(DIMS, artra_norm(poly->Ne[n]), poly->Ne[n]);
    }
}

void artra_compute_geometry_parameters(int collapse, Geometry *geo)
{
    for (int n = 0; n < geo->sphN; ++n) SphereParameters(collapse, geo->poly + n);
    for (int n = geo->sphN; n < geo->totN; ++n) PolyhedronParameters(collapse, geo->poly + n);
}

/* ------------------------------------------------- closest point queries */
/*
 * Squared distance from p to the triangle v0 + s*e01 + t*e02 and the
 * barycentric parameters (1-s-t, s, t) of the closest point, by the region
 * classification of Eberly, "Distance between point and triangle in 3D".
 */
Real artra_point_triangle_distance(const Real p[DIMS], const Real v0[DIMS], const Real e01[DIMS],
                                   const Real e02[DIMS], Real para[DIMS])
{
    return GcPointTriangleDistance(p, v0, e01, e02, para);
}


/* ------------------------------------------------ closest-facet search */
/*
 * The closest facet is the first facet, in facet order, at the minimum of
 * the computed squared distance (computational_geometry.c:486-494).  With
 * bounds, facets that cannot reach that minimum are skipped: a facet whose
 * box lies farther than a distance already reached (plus a slack of 1e-12
 * times the squared extent around the point, far above the rounding of
 * either quantity) can neither lower the minimum nor tie it, so the
 * remaining facets, still visited in facet order with the same strict
 * comparison, give the same facet as the full scan.  Blocks of consecutive
 * facets carry their union box, so most facets are ruled out by block.
 */
static Real BoxDistance2(const Real p[DIMS], const Real b[2 * DIMS])
{
    Real d2 = 0.0;
    for (int s = 0; s < DIMS; ++s) {
        Real g = 0.0;
        if (p[s] < b[s]) g = b[s] - p[s];
        else if (p[s] > b[DIMS + s]) g = p[s] - b[DIMS + s];
        d2 = d2 + g * g;
    }
    return d2;
}

static void GrowBox(Real b[2 * DIMS], const Real v[DIMS])
{
    for (int s = 0; s < DIMS; ++s) {
        if (v[s] < b[s]) b[s] = v[s];
        if (v[s] > b[DIMS + s]) b[DIMS + s] = v[s];
    }
}

int artra_facet_bounds(const Polyhedron *poly, ArtraFacetBounds *fb)
{
    memset(fb, 0, sizeof *fb);
    if (0 >= poly->faceN) return 0;
    fb->blocks = (poly->faceN + ARTRA_FACET_BLOCK - 1) / ARTRA_FACET_BLOCK;
    fb->facet = malloc((size_t)poly->faceN * sizeof(*fb->facet));
    fb->block = malloc((size_t)fb->blocks * sizeof(*fb->block));
    if (!fb->facet || !fb->block) {
        artra_free_facet_bounds(fb);
        return artra_set_error(1, "out of host memory for facet bounds");
    }
    for (int s = 0; s < DIMS; ++s) {
        fb->all[s] = FLT_MAX;
        fb->all[DIMS + s] = -FLT_MAX;
    }
    for (int b = 0; b < fb->blocks; ++b) {
        Real *bb = fb->block[b];
        for (int s = 0; s < DIMS; ++s) {
            bb[s] = FLT_MAX;
            bb[DIMS + s] = -FLT_MAX;
        }
        const int end = (b + 1) * ARTRA_FACET_BLOCK < poly->faceN ? (b + 1) * ARTRA_FACET_BLOCK : poly->faceN;
        for (int n = b * ARTRA_FACET_BLOCK; n < end; ++n) {
            Real *fx = fb->facet[n];
            for (int s = 0; s < DIMS; ++s) {
                fx[s] = FLT_MAX;
                fx[DIMS + s] = -FLT_MAX;
            }
            for (int q = 0; q < 3; ++q) {
                GrowBox(fx, poly->v[poly->f[n][q]]);
                GrowBox(bb, poly->v[poly->f[n][q]]);
                GrowBox(fb->all, poly->v[poly->f[n][q]]);
            }
        }
    }
    return 0;
}

void artra_free_facet_bounds(ArtraFacetBounds *fb)
{
    free(fb->facet);
    free(fb->block);
    memset(fb, 0, sizeof *fb);
}

static Real FacetDistance2(const Real p[DIMS], int n, const Polyhedron *poly)
{
    const ArtraPolyView pv = artra_poly_view(poly);
    return GcFacetDistance2(p, n, &pv);
}

static int ClosestFacet(const Real p[DIMS], const Polyhedron *poly, const ArtraFacetBounds *fb)
{
    Real best = FLT_MAX;
    int cid = 0;
    if (NULL == fb || NULL == fb->facet) {
        for (int n = 0; n < poly->faceN; ++n) {
            const Real d2 = FacetDistance2(p, n, poly);
            if (best > d2) { /* the first facet at the minimum wins */
                best = d2;
                cid = n;
            }
        }
        return cid;
    }
    /* a distance that some facet reaches: that of the facet whose box is
     * nearest within the block whose box is nearest */
    int near = 0;
    Real nearLb = FLT_MAX;
    for (int b = 0; b < fb->blocks; ++b) {
        const Real lb = BoxDistance2(p, fb->block[b]);
        if (lb < nearLb) {
            nearLb = lb;
            near = b;
        }
    }
    const int nearEnd = (near + 1) * ARTRA_FACET_BLOCK < poly->faceN ? (near + 1) * ARTRA_FACET_BLOCK : poly->faceN;
    int probe = near * ARTRA_FACET_BLOCK;
    nearLb = FLT_MAX;
    for (int n = near * ARTRA_FACET_BLOCK; n < nearEnd; ++n) {
        const Real lb = BoxDistance2(p, fb->facet[n]);
        if (lb < nearLb) {
            nearLb = lb;
            probe = n;
        }
    }
    const Real reach = FacetDistance2(p, probe, poly);
    /* slack: 1e-12 of the squared distance to the far corner of all facets */
    Real extent = 0.0;
    for (int s = 0; s < DIMS; ++s) {
        const Real g = fabs(p[s] - fb->all[s]) > fabs(p[s] - fb->all[DIMS + s]) ? fabs(p[s] - fb->all[s])
                                                                                 : fabs(p[s] - fb->all[DIMS + s]);
        extent = extent + g * g;
    }
    const Real bound = reach + 1.0e-12 * extent;
    for (int b = 0; b < fb->blocks; ++b) {
        if (BoxDistance2(p, fb->block[b]) > bound) continue;
        const int end = (b + 1) * ARTRA_FACET_BLOCK < poly->faceN ? (b + 1) * ARTRA_FACET_BLOCK : poly->faceN;
        for (int n = b * ARTRA_FACET_BLOCK; n < end; ++n) {
            if (BoxDistance2(p, fb->facet[n]) > bound) continue;
            const Real d2 = FacetDistance2(p, n, poly);
            if (best > d2) {
                best = d2;
                cid = n;
            }
        }
    }
    return cid;
}

int artra_point_in_polyhedron(const Real p[DIMS], const Polyhedron *poly, const ArtraFacetBounds *bounds, int *fid)
{
    const int cid = ClosestFacet(p, poly, bounds);
    *fid = cid;
    const ArtraPolyView pv = artra_poly_view(poly);
    return GcInsideOfFacet(p, cid, &pv); /* inside or on the surface */
}

/* Boundary point pi, image point pm = 2 pi - p and normal N of a ghost node. */
void artra_compute_geometric_data(const Real p[DIMS], int fid, const Polyhedron *poly, Real pi[DIMS],
                                  Real pm[DIMS], Real N[DIMS])
{
    ArtraPolyView pv;
    if (0 < poly->faceN) {
        pv = artra_poly_view(poly);
    } else {
        memset(&pv, 0, sizeof pv);
    }
    GcGeometricData(p, fid, &pv, poly->O, poly->r, pi, pm, N);
}

/* ------------------------------------------------- node classification */
/* Layer (1..gl) of the first search-path neighbour that satisfies the test,
 * or 0.  sameAs != 0 looks for a node whose did equals `did`; otherwise for
 * one whose did differs from it. */
static int InterState(const Partition *part, const Node *node, int k, int j, int i, int did, int sameAs)
{
    for (int n = 0; n < part->pathSep[0]; ++n) {
        const int kh = k + part->path[n][Z], jh = j + part->path[n][Y], ih = i + part->path[n][X];
        if (!artra_in_box(kh, jh, ih, part->ns[PIN])) continue;
        const int other = node[artra_index(kh, jh, ih, part->n[Y], part->n[X])].did;
        if ((sameAs && other == did) || (!sameAs && other != did)) {
            for (int r = 1; r <= part->gl; ++r) {
                if (part->pathSep[r] > n) return r;
            }
        }
    }
    return 0;
}

int artra_compute_geometric_field(Space *space, const Model *model, ArtraGhostTable *joins,
                                  const int (*region)[LIMIT])
{
    (void)model;
    const Partition *part = &space->part;
    Node *node = space->node;
    const Geometry *geo = &space->geo;
    /* the node range of the two whole-field passes */
    int lo[DIMS], hi[DIMS];
    for (int s = 0; s < DIMS; ++s) {
        lo[s] = part->ns[PIN][s][MIN];
        hi[s] = part->ns[PIN][s][MAX];
        if (region) {
            if (region[s][MIN] > lo[s]) lo[s] = region[s][MIN];
            if (region[s][MAX] < hi[s]) hi[s] = region[s][MAX];
        }
    }
    /* keep the previous domain field in gst (immersed_boundary.c:72-113) */
    for (int k = lo[Z]; k < hi[Z]; ++k)
        for (int j = lo[Y]; j < hi[Y]; ++j)
            for (int i = lo[X]; i < hi[X]; ++i) {
                Node *nd = node + artra_index(k, j, i, part->n[Y], part->n[X]);
                const int gid = nd->did;
                nd->gst = nd->did;
                if (0 == gid) {
                    nd->fid = 0;
                    continue;
                }
                if (1 == geo->poly[gid - 1].state) continue;
                if (0 < nd->lid) nd->did = 0;
            }
    /* nodes inside or on each body (immersed_boundary.c:128-182) */
    for (int n = 0; n < geo->totN; ++n) {
        const Polyhedron *poly = geo->poly + n;
        if (1 == poly->state) continue;
        ArtraFacetBounds bounds;
        if (artra_facet_bounds(poly, &bounds)) return 1;
        int box[DIMS][LIMIT];
        for (int s = 0; s < DIMS; ++s) {
            const int lo = part->ns[PIN][s][MIN], hi = part->ns[PIN][s][MAX];
            box[s][MIN] = artra_confine(artra_map_node(poly->box[s][MIN], part->domain[s][MIN], part->dd[s],
                                                       part->ng[s]), lo, hi);
            box[s][MAX] = artra_confine(artra_map_node(poly->box[s][MAX], part->domain[s][MIN], part->dd[s],
                                                       part->ng[s]), lo, hi) + 1;
        }
        /* the nodes of one body's box are independent of each other */
        const int rowsY = box[Y][MAX] - box[Y][MIN];
        const int rows = (box[Z][MAX] - box[Z][MIN]) * (rowsY > 0 ? rowsY : 0);
#pragma omp parallel for schedule(dynamic, 4) if (rows > 16)
        for (int row = 0; row < rows; ++row) {
            const int k = box[Z][MIN] + row / rowsY, j = box[Y][MIN] + row % rowsY;
            for (int i = box[X][MIN]; i < box[X][MAX]; ++i) {
                Node *nd = node + artra_index(k, j, i, part->n[Y], part->n[X]);
                if (0 != nd->did) continue;
                const Real p[DIMS] = {artra_map_point(i, part->domain[X][MIN], part->d[X], part->ng[X]),
                                      artra_map_point(j, part->domain[Y][MIN], part->d[Y], part->ng[Y]),
                                      artra_map_point(k, part->domain[Z][MIN], part->d[Z], part->ng[Z])};
                if (0 >= poly->faceN) {
                    if (poly->r * poly->r >= artra_dist2(poly->O, p)) {
                        nd->did = n + 1;
                        nd->fid = 0;
                    }
                } else {
                    int fid = 0;
                    if (artra_point_in_polyhedron(p, poly, &bounds, &fid)) {
                        nd->did = n + 1;
                        nd->fid = fid;
                    }
                }
            }
        }
        artra_free_facet_bounds(&bounds);
    }
    /* interfacial and ghost layers (immersed_boundary.c:183-225).  A fluid
     * node that a moving body has uncovered (gst, still the previous domain
     * field, differs from did = 0) is reconstructed from the nodes that were
     * fluid before and are fluid now; that is the set the reference's single
     * pass finds whatever its order (a node the pass has reached holds gst 0
     * and, if uncovered, the mark fid < 0; one it has not reached still holds
     * its previous domain in gst), so the list is built first, on the flags
     * as they stand, and the interfacial states, which read did only, follow
     * in parallel */
    for (int k = lo[Z]; k < hi[Z]; ++k)
        for (int j = lo[Y]; j < hi[Y]; ++j)
            for (int i = lo[X]; i < hi[X]; ++i) {
                const Node *nd = node + artra_index(k, j, i, part->n[Y], part->n[X]);
                if (nd->gst != nd->did && 0 == nd->did) {
                    if (NULL == joins) return artra_set_error(1, "a body uncovered node (%d, %d, %d)", k, j, i);
                    if (artra_add_join(space, k, j, i, joins)) return 1;
                }
            }
    for (int q = 0; joins && q < joins->nghost; ++q) node[joins->ghosts[q].node].fid = NONE;
    const int spanY = hi[Y] - lo[Y];
    const int rows = (hi[Z] - lo[Z]) * (spanY > 0 ? spanY : 0);
#pragma omp parallel for schedule(static) if (rows > 64)
    for (int row = 0; row < rows; ++row) {
        const int k = lo[Z] + row / spanY, j = lo[Y] + row % spanY;
        for (int i = lo[X]; i < hi[X]; ++i) {
            Node *nd = node + artra_index(k, j, i, part->n[Y], part->n[X]);
            nd->lid = 0;
            nd->gst = 0;
            if (0 == nd->did) continue;
            nd->lid = InterState(part, node, k, j, i, nd->did, 0);
            if (0 < nd->lid) nd->gst = InterState(part, node, k, j, i, 0, 1);
        }
    }
    return 0;
}

/* ------------------------------------------------ classification scope */
static void EmptyBox(int box[DIMS][LIMIT])
{
    for (int s = 0; s < DIMS; ++s) {
        box[s][MIN] = INT_MAX;
        box[s][MAX] = INT_MIN;
    }
}

static void AddNode(int box[DIMS][LIMIT], int k, int j, int i)
{
    const int at[DIMS] = {i, j, k};
    for (int s = 0; s < DIMS; ++s) {
        if (at[s] < box[s][MIN]) box[s][MIN] = at[s];
        if (at[s] + 1 > box[s][MAX]) box[s][MAX] = at[s] + 1;
    }
}

static void AddBox(int box[DIMS][LIMIT], const int add[DIMS][LIMIT])
{
    for (int s = 0; s < DIMS; ++s) {
        if (add[s][MIN] >= add[s][MAX]) return;
    }
    for (int s = 0; s < DIMS; ++s) {
        if (add[s][MIN] < box[s][MIN]) box[s][MIN] = add[s][MIN];
        if (add[s][MAX] > box[s][MAX]) box[s][MAX] = add[s][MAX];
    }
}

int artra_scope_init(const Space *space, ArtraGeometryScope *sc)
{
    const Partition *part = &space->part;
    const Geometry *geo = &space->geo;
    memset(sc, 0, sizeof *sc);
    sc->bodies = geo->totN;
    sc->extent = malloc((size_t)(geo->totN > 0 ? geo->totN : 1) * sizeof(*sc->extent));
    if (NULL == sc->extent) return artra_set_error(1, "out of host memory");
    for (int n = 0; n < geo->totN; ++n) EmptyBox(sc->extent[n]);
    EmptyBox(sc->join);
    for (int k = part->ns[PIN][Z][MIN]; k < part->ns[PIN][Z][MAX]; ++k)
        for (int j = part->ns[PIN][Y][MIN]; j < part->ns[PIN][Y][MAX]; ++j)
            for (int i = part->ns[PIN][X][MIN]; i < part->ns[PIN][X][MAX]; ++i) {
                const Node *nd = space->node + artra_index(k, j, i, part->n[Y], part->n[X]);
                if (0 < nd->did && nd->did <= geo->totN) AddNode(sc->extent[nd->did - 1], k, j, i);
                if (0 == nd->did && 0 > nd->fid) AddNode(sc->join, k, j, i);
            }
    return 0;
}

void artra_scope_free(ArtraGeometryScope *sc)
{
    free(sc->extent);
    memset(sc, 0, sizeof *sc);
}

void artra_scope_region(const Space *space, const ArtraGeometryScope *sc, int region[DIMS][LIMIT])
{
    const Partition *part = &space->part;
    const Geometry *geo = &space->geo;
    EmptyBox(region);
    for (int n = 0; n < geo->totN; ++n) {
        const Polyhedron *poly = geo->poly + n;
        if (1 == poly->state) continue;
        int box[DIMS][LIMIT];
        for (int s = 0; s < DIMS; ++s) {
            const int lo = part->ns[PIN][s][MIN], hi = part->ns[PIN][s][MAX];
            box[s][MIN] = artra_confine(artra_map_node(poly->box[s][MIN], part->domain[s][MIN], part->dd[s],
                                                       part->ng[s]), lo, hi);
            box[s][MAX] = artra_confine(artra_map_node(poly->box[s][MAX], part->domain[s][MIN], part->dd[s],
                                                       part->ng[s]), lo, hi) + 1;
        }
        AddBox(region, box);
        AddBox(region, (const int (*)[LIMIT])sc->extent[n]);
    }
    int empty = 0;
    for (int s = 0; s < DIMS; ++s) empty |= (region[s][MIN] >= region[s][MAX]);
    for (int s = 0; s < DIMS && !empty; ++s) {
        region[s][MIN] -= part->gl;
        region[s][MAX] += part->gl;
    }
    AddBox(region, (const int (*)[LIMIT])sc->join);
}

void artra_scope_update(const Space *space, ArtraGeometryScope *sc, const ArtraGhostTable *joins,
                        const int region[DIMS][LIMIT])
{
    const Partition *part = &space->part;
    const Geometry *geo = &space->geo;
    int lo[DIMS], hi[DIMS];
    for (int s = 0; s < DIMS; ++s) {
        lo[s] = (region[s][MIN] > part->ns[PIN][s][MIN]) ? region[s][MIN] : part->ns[PIN][s][MIN];
        hi[s] = (region[s][MAX] < part->ns[PIN][s][MAX]) ? region[s][MAX] : part->ns[PIN][s][MAX];
    }
    for (int n = 0; n < geo->totN; ++n) {
        if (1 != geo->poly[n].state) EmptyBox(sc->extent[n]);
    }
    for (int k = lo[Z]; k < hi[Z]; ++k)
        for (int j = lo[Y]; j < hi[Y]; ++j)
            for (int i = lo[X]; i < hi[X]; ++i) {
                const int did = space->node[artra_index(k, j, i, part->n[Y], part->n[X])].did;
                if (0 < did && did <= geo->totN && 1 != geo->poly[did - 1].state) AddNode(sc->extent[did - 1], k, j, i);
            }
    EmptyBox(sc->join);
    for (int q = 0; joins && q < joins->nghost; ++q) {
        const int idx = joins->ghosts[q].node;
        const int i = idx % part->n[X], j = (idx / part->n[X]) % part->n[Y], k = idx / (part->n[X] * part->n[Y]);
        AddNode(sc->join, k, j, i);
    }
}

int artra_apply_classification(Space *space, const int region[DIMS][LIMIT], const int *packed, long count,
                               ArtraGhostTable *joins)
{
    const Partition *part = &space->part;
    Node *node = space->node;
    int lo[DIMS], hi[DIMS];
    long span = 1;
    for (int s = 0; s < DIMS; ++s) {
        lo[s] = (region[s][MIN] > part->ns[PIN][s][MIN]) ? region[s][MIN] : part->ns[PIN][s][MIN];
        hi[s] = (region[s][MAX] < part->ns[PIN][s][MAX]) ? region[s][MAX] : part->ns[PIN][s][MAX];
        span *= (hi[s] > lo[s]) ? hi[s] - lo[s] : 0;
    }
    if (span != count) return artra_set_error(1, "classification of %ld nodes for a region of %ld", count, span);
    /* the flags as the reference's interfacial pass finds them: the new
     * domain field, gst still holding the previous one */
    long t = 0;
    for (int k = lo[Z]; k < hi[Z]; ++k)
        for (int j = lo[Y]; j < hi[Y]; ++j)
            for (int i = lo[X]; i < hi[X]; ++i, ++t) {
                Node *nd = node + artra_index(k, j, i, part->n[Y], part->n[X]);
                nd->did = packed[t];
                nd->fid = packed[count + t];
                nd->gst = packed[4 * count + t];
            }
    for (int k = lo[Z]; k < hi[Z]; ++k)
        for (int j = lo[Y]; j < hi[Y]; ++j)
            for (int i = lo[X]; i < hi[X]; ++i) {
                const Node *nd = node + artra_index(k, j, i, part->n[Y], part->n[X]);
                if (nd->gst != nd->did && 0 == nd->did && artra_add_join(space, k, j, i, joins)) return 1;
            }
    for (int q = 0; q < joins->nghost; ++q) node[joins->ghosts[q].node].fid = NONE;
    t = 0;
    for (int k = lo[Z]; k < hi[Z]; ++k)
        for (int j = lo[Y]; j < hi[Y]; ++j)
            for (int i = lo[X]; i < hi[X]; ++i, ++t) {
                Node *nd = node + artra_index(k, j, i, part->n[Y], part->n[X]);
                nd->lid = packed[2 * count + t];
                nd->gst = packed[3 * count + t];
            }
    return 0;
}
