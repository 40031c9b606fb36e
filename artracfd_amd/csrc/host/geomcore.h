/*
 * geomcore.h - closest-point queries on a triangulated body, shared by the
 * host classification (geometry.c) and the device classification kernels
 * (artra_device.hip).  One source for both sides, both compiled with
 * -ffp-contract=off, so a node's closest facet, boundary point and
 * inside/outside decision are the same bits on either (the region logic of
 * computational_geometry.c:449-688, Eberly's published method).
 */
#ifndef ARTRA_GEOMCORE_H_
#define ARTRA_GEOMCORE_H_

#include <math.h>

#include "artra_types.h"

#ifdef __HIPCC__
#define ARTRA_HD __host__ __device__ static inline
#else
#define ARTRA_HD static inline
#endif

/* The arrays of a triangulated body in its current position. */
typedef struct {
    int faceN, edgeN, vertN;
    const int (*f)[POLYN];
    const int (*e)[EVF];     /* sorted by (larger vertex, smaller vertex) */
    const Real (*v)[DIMS];
    const Real (*Nf)[DIMS];
    const Real (*Ne)[DIMS];
    const Real (*Nv)[DIMS];
} ArtraPolyView;

ARTRA_HD Real GcDot(const Real a[DIMS], const Real b[DIMS])
{
    return a[X] * b[X] + a[Y] * b[Y] + a[Z] * b[Z];
}

/*
 * Squared distance from p to the triangle v0 + s*e01 + t*e02 and the
 * barycentric parameters (1-s-t, s, t) of the closest point, by the region
 * classification of Eberly, "Distance between point and triangle in 3D".
 */
ARTRA_HD Real GcPointTriangleDistance(const Real p[DIMS], const Real v0[DIMS], const Real e01[DIMS],
                                      const Real e02[DIMS], Real para[DIMS])
{
    const Real D[DIMS] = {v0[X] - p[X], v0[Y] - p[Y], v0[Z] - p[Z]};
    const Real a = GcDot(e01, e01), b = GcDot(e01, e02), c = GcDot(e02, e02);
    const Real d = GcDot(e01, D), e = GcDot(e02, D), f = GcDot(D, D);
    const Real det = a * c - b * b;
    Real s = b * e - c * d;
    Real t = b * d - a * e;
    Real dist;
    /* the closest point on an edge parameterised from its start */
    #define ON_E01(sv) do { s = (sv); dist = d * s + f; } while (0)
    #define ON_E02(tv) do { t = (tv); dist = e * t + f; } while (0)
    #define INTERIOR() (s * (a * s + b * t + 2.0 * d) + t * (b * s + c * t + 2.0 * e) + f)
    if (s + t <= det) {
        if (s < 0.0) {
            if (t < 0.0) {                                    /* region 4 */
                if (d < 0.0) {
                    t = 0.0;
                    if (-d >= a) { s = 1.0; dist = a + 2.0 * d + f; } else ON_E01(-d / a);
                } else {
                    s = 0.0;
                    if (e >= 0.0) { t = 0.0; dist = f; }
                    else if (-e >= c) { t = 1.0; dist = c + 2.0 * e + f; }
                    else ON_E02(-e / c);
                }
            } else {                                          /* region 3 */
                s = 0.0;
                if (e >= 0.0) { t = 0.0; dist = f; }
                else if (-e >= c) { t = 1.0; dist = c + 2.0 * e + f; }
                else ON_E02(-e / c);
            }
        } else if (t < 0.0) {                                 /* region 5 */
            t = 0.0;
            if (d >= 0.0) { s = 0.0; dist = f; }
            else if (-d >= a) { s = 1.0; dist = a + 2.0 * d + f; }
            else ON_E01(-d / a);
        } else {                                              /* region 0 */
            s = s / det;
            t = t / det;
            dist = INTERIOR();
        }
    } else if (s < 0.0) {                                     /* region 2 */
        const Real t0 = b + d, t1 = c + e;
        if (t1 > t0) {
            const Real num = t1 - t0, den = a - 2.0 * b + c;
            if (num >= den) { s = 1.0; t = 0.0; dist = a + 2.0 * d + f; }
            else { s = num / den; t = 1.0 - s; dist = INTERIOR(); }
        } else {
            s = 0.0;
            if (t1 <= 0.0) { t = 1.0; dist = c + 2.0 * e + f; }
            else if (e >= 0.0) { t = 0.0; dist = f; }
            else ON_E02(-e / c);
        }
    } else if (t < 0.0) {                                     /* region 6 */
        const Real t0 = b + e, t1 = a + d;
        if (t1 > t0) {
            const Real num = t1 - t0, den = a - 2.0 * b + c;
            if (num >= den) { t = 1.0; s = 0.0; dist = c + 2.0 * e + f; }
            else { t = num / den; s = 1.0 - t; dist = INTERIOR(); }
        } else {
            t = 0.0;
            if (t1 <= 0.0) { s = 1.0; dist = a + 2.0 * d + f; }
            else if (d >= 0.0) { s = 0.0; dist = f; }
            else ON_E01(-d / a);
        }
    } else {                                                  /* region 1 */
        const Real num = c + e - b - d;
        if (num <= 0.0) { s = 0.0; t = 1.0; dist = c + 2.0 * e + f; }
        else {
            const Real den = a - 2.0 * b + c;
            if (num >= den) { s = 1.0; t = 0.0; dist = a + 2.0 * d + f; }
            else { s = num / den; t = 1.0 - s; dist = INTERIOR(); }
        }
    }
    #undef ON_E01
    #undef ON_E02
    #undef INTERIOR
    para[0] = 1.0 - s - t;
    para[1] = s;
    para[2] = t;
    return (dist < 0.0) ? 0.0 : dist;
}

ARTRA_HD void GcTriangleOf(int fid, const ArtraPolyView *pv, Real v0[DIMS], Real v1[DIMS], Real v2[DIMS],
                           Real e01[DIMS], Real e02[DIMS])
{
    for (int s = 0; s < DIMS; ++s) {
        v0[s] = pv->v[pv->f[fid][0]][s];
        v1[s] = pv->v[pv->f[fid][1]][s];
        v2[s] = pv->v[pv->f[fid][2]][s];
        e01[s] = v1[s] - v0[s];
        e02[s] = v2[s] - v0[s];
    }
}

/* Edge keys (larger vertex, smaller vertex) in lexicographic order. */
ARTRA_HD int GcEdgeKeyCompare(const int *a, const int *b)
{
    if (a[0] != b[0]) return (a[0] < b[0]) ? -1 : 1;
    if (a[1] != b[1]) return (a[1] < b[1]) ? -1 : 1;
    return 0;
}

ARTRA_HD int GcFindEdge(const ArtraPolyView *pv, int a, int b)
{
    const int key[2] = {(a > b) ? a : b, (a > b) ? b : a};
    int lo = 0, hi = pv->edgeN - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        const int c = GcEdgeKeyCompare(key, pv->e[mid]);
        if (0 == c) return mid;
        if (c > 0) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* Closest point on facet fid and the matching face, edge or vertex
 * pseudo-normal, chosen by exact tests on the barycentric parameters. */
ARTRA_HD void GcClosestOnFacet(const Real p[DIMS], int fid, const ArtraPolyView *pv, Real pi[DIMS], Real N[DIMS])
{
    Real v0[DIMS], v1[DIMS], v2[DIMS], e01[DIMS], e02[DIMS], para[DIMS];
    const int vid[3] = {pv->f[fid][0], pv->f[fid][1], pv->f[fid][2]};
    GcTriangleOf(fid, pv, v0, v1, v2, e01, e02);
    GcPointTriangleDistance(p, v0, e01, e02, para);
    const Real *normal;
    int edge;
    if (0.0 == para[1]) {
        if (0.0 == para[2]) {
            for (int s = 0; s < DIMS; ++s) pi[s] = v0[s];
            normal = pv->Nv[vid[0]];
        } else if (1.0 == para[2]) {
            for (int s = 0; s < DIMS; ++s) pi[s] = v2[s];
            normal = pv->Nv[vid[2]];
        } else {
            for (int s = 0; s < DIMS; ++s) pi[s] = v0[s] + para[2] * e02[s];
            edge = GcFindEdge(pv, vid[0], vid[2]);
            normal = (edge >= 0) ? pv->Ne[edge] : pv->Nf[fid];
        }
    } else if (1.0 == para[1]) {
        for (int s = 0; s < DIMS; ++s) pi[s] = v1[s];
        normal = pv->Nv[vid[1]];
    } else if (0.0 == para[2]) {
        for (int s = 0; s < DIMS; ++s) pi[s] = v0[s] + para[1] * e01[s];
        edge = GcFindEdge(pv, vid[0], vid[1]);
        normal = (edge >= 0) ? pv->Ne[edge] : pv->Nf[fid];
    } else if (0.0 == para[0]) {
        for (int s = 0; s < DIMS; ++s) pi[s] = v0[s] + para[1] * e01[s] + para[2] * e02[s];
        edge = GcFindEdge(pv, vid[1], vid[2]);
        normal = (edge >= 0) ? pv->Ne[edge] : pv->Nf[fid];
    } else {
        for (int s = 0; s < DIMS; ++s) pi[s] = v0[s] + para[1] * e01[s] + para[2] * e02[s];
        normal = pv->Nf[fid];
    }
    for (int s = 0; s < DIMS; ++s) N[s] = normal[s];
}

ARTRA_HD Real GcFacetDistance2(const Real p[DIMS], int n, const ArtraPolyView *pv)
{
    Real v0[DIMS], v1[DIMS], v2[DIMS], e01[DIMS], e02[DIMS], para[DIMS];
    GcTriangleOf(n, pv, v0, v1, v2, e01, e02);
    return GcPointTriangleDistance(p, v0, e01, e02, para);
}

/* Inside or on the body, by the pseudo-normal at the closest point of the
 * closest facet cid (computational_geometry.c:495-505). */
ARTRA_HD int GcInsideOfFacet(const Real p[DIMS], int cid, const ArtraPolyView *pv)
{
    Real pi[DIMS], N[DIMS];
    GcClosestOnFacet(p, cid, pv, pi, N);
    const Real r[DIMS] = {p[X] - pi[X], p[Y] - pi[Y], p[Z] - pi[Z]};
    return !(0.0 < GcDot(r, N));
}

ARTRA_HD void GcCross(const Real a[DIMS], const Real b[DIMS], Real out[DIMS])
{
    out[X] = a[Y] * b[Z] - a[Z] * b[Y];
    out[Y] = a[Z] * b[X] - a[X] * b[Z];
    out[Z] = a[X] * b[Y] - a[Y] * b[X];
}

/* Node of a coordinate (cfd_commons.c:352-355). */
ARTRA_HD int GcMapNode(Real s, Real smin, Real dds, int ng)
{
    return (int)((s - smin) * dds + 0.5) + ng;
}

/* Tangent pair of a unit normal (cfd_commons.c:431-458): the first tangent
 * zeroes the normal's smallest component (ties resolved toward z, then y). */
ARTRA_HD void GcOrthogonalSpace(const Real N[DIMS], Real Ta[DIMS], Real Tb[DIMS])
{
    int small = Z;
    if (fabs(N[small]) > fabs(N[Y])) small = Y;
    if (fabs(N[small]) > fabs(N[X])) small = X;
    if (X == small) {
        Ta[X] = 0.0; Ta[Y] = -N[Z]; Ta[Z] = N[Y];
    } else if (Y == small) {
        Ta[X] = N[Z]; Ta[Y] = 0.0; Ta[Z] = -N[X];
    } else {
        Ta[X] = -N[Y]; Ta[Y] = N[X]; Ta[Z] = 0.0;
    }
    const Real norm = sqrt(GcDot(Ta, Ta));
    for (int s = 0; s < DIMS; ++s) Ta[s] = Ta[s] / norm;
    GcCross(Ta, N, Tb);
}

/* Boundary point pi, image point pm = 2 pi - p and normal N of a ghost node
 * (computational_geometry.c:689-783): a sphere (faceN <= 0: centre O,
 * radius r) or facet fid of a triangulated body. */
ARTRA_HD void GcGeometricData(const Real p[DIMS], int fid, const ArtraPolyView *pv, const Real O[DIMS], Real r,
                              Real pi[DIMS], Real pm[DIMS], Real N[DIMS])
{
    if (0 >= pv->faceN) {
        for (int s = 0; s < DIMS; ++s) N[s] = p[s] - O[s];
        Real dist = sqrt(GcDot(N, N));
        for (int s = 0; s < DIMS; ++s) N[s] = N[s] / dist;
        dist = r - dist;
        for (int s = 0; s < DIMS; ++s) pi[s] = p[s] + dist * N[s];
    } else {
        GcClosestOnFacet(p, fid, pv, pi, N);
    }
    for (int s = 0; s < DIMS; ++s) pm[s] = pi[s] + pi[s] - p[s];
}

#endif /* ARTRA_GEOMCORE_H_ */
