/*
 * scope_check.c - CPU check of the moving-body classification shortcuts.
 *
 * Prepares a case twice and moves its non-stationary bodies identically in
 * both, `moves` times by (dx, dy, dz) and a rotation az about z (spheres:
 * centre and box; triangulated bodies: the polyhedron transform).  After
 * each move copy A is classified over the whole interior (region NULL) and
 * copy B only over the scope's region; the node flags of every node and the
 * reconstruction lists (nodes, stencil nodes, weights) must be identical.
 * Also checks the bounded closest-facet search against the full scan on
 * every interior node and on points placed exactly on vertices, edges and
 * facets, and that the ghost table built on every thread equals the one
 * built on one.  Prints "ok <joined nodes>" or the first difference and
 * exits 1.
 *
 *   scope_check <case dir> <moves> <dx> <dy> <dz> <az>
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

#include "artra_host.h"

static void Move(Space *space, const Real off[DIMS], const Real ang[DIMS])
{
    static const Real unit[DIMS] = {1.0, 1.0, 1.0};
    for (int n = 0; n < space->geo.totN; ++n) {
        Polyhedron *poly = space->geo.poly + n;
        if (1 == poly->state) continue;
        if (0 >= poly->faceN) {
            for (int s = 0; s < DIMS; ++s) {
                poly->O[s] = poly->O[s] + off[s];
                poly->box[s][MIN] = poly->O[s] - poly->r;
                poly->box[s][MAX] = poly->O[s] + poly->r;
            }
        } else {
            artra_transform_polyhedron(poly->O, unit, ang, off, poly);
        }
    }
}

static int SameJoins(const ArtraGhostTable *a, const ArtraGhostTable *b)
{
    if (a->nghost != b->nghost || a->nstencil != b->nstencil) return 0;
    for (int q = 0; q < a->nghost; ++q) {
        if (a->ghosts[q].node != b->ghosts[q].node || a->ghosts[q].first != b->ghosts[q].first ||
            a->ghosts[q].count != b->ghosts[q].count) return 0;
    }
    return 0 == memcmp(a->stencil_node, b->stencil_node, a->nstencil * sizeof(int)) &&
           0 == memcmp(a->stencil_w, b->stencil_w, a->nstencil * sizeof(Real));
}

/* The ghost table built with every thread equals the one built with one. */
static int SameGhostTables(const Space *space, const Model *model)
{
    ArtraGhostTable a, b;
    if (artra_build_ghost_table(space, model, &a)) return 0;
    const int threads = omp_get_max_threads();
    omp_set_num_threads(1);
    const int rc = artra_build_ghost_table(space, model, &b);
    omp_set_num_threads(threads);
    if (rc) return 0;
    int same = a.nghost == b.nghost && a.nstencil == b.nstencil && a.layers == b.layers &&
               0 == memcmp(a.layer_start, b.layer_start, (a.layers + 1) * sizeof(int)) &&
               0 == memcmp(a.ghosts, b.ghosts, a.nghost * sizeof(ArtraGhost)) &&
               0 == memcmp(a.stencil_node, b.stencil_node, a.nstencil * sizeof(int)) &&
               0 == memcmp(a.stencil_w, b.stencil_w, a.nstencil * sizeof(Real));
    same = same && a.nghost > 0;
    artra_free_ghost_table(&a);
    artra_free_ghost_table(&b);
    return same;
}

static int FacetSearch(const Space *space)
{
    const Partition *part = &space->part;
    for (int n = 0; n < space->geo.totN; ++n) {
        const Polyhedron *poly = space->geo.poly + n;
        if (0 >= poly->faceN) continue;
        ArtraFacetBounds fb;
        if (artra_facet_bounds(poly, &fb)) return 1;
        long tested = 0;
        for (int k = part->ns[PIN][Z][MIN]; k < part->ns[PIN][Z][MAX]; ++k)
            for (int j = part->ns[PIN][Y][MIN]; j < part->ns[PIN][Y][MAX]; ++j)
                for (int i = part->ns[PIN][X][MIN]; i < part->ns[PIN][X][MAX]; ++i) {
                    const Real p[DIMS] = {artra_map_point(i, part->domain[X][MIN], part->d[X], part->ng[X]),
                                          artra_map_point(j, part->domain[Y][MIN], part->d[Y], part->ng[Y]),
                                          artra_map_point(k, part->domain[Z][MIN], part->d[Z], part->ng[Z])};
                    int fa = -1, fbid = -1;
                    const int ia = artra_point_in_polyhedron(p, poly, NULL, &fa);
                    const int ib = artra_point_in_polyhedron(p, poly, &fb, &fbid);
                    if (ia != ib || fa != fbid) {
                        printf("facet search differs at node (%d, %d, %d): %d/%d vs %d/%d\n", k, j, i, ia, fa, ib, fbid);
                        return 1;
                    }
                    ++tested;
                }
        /* points exactly on vertices, on edge midpoints and on facet centroids */
        for (int f = 0; f < poly->faceN; f += 7) {
            const Real *v[3] = {poly->v[poly->f[f][0]], poly->v[poly->f[f][1]], poly->v[poly->f[f][2]]};
            Real pts[3][DIMS];
            for (int s = 0; s < DIMS; ++s) {
                pts[0][s] = v[0][s];
                pts[1][s] = 0.5 * (v[0][s] + v[1][s]);
                pts[2][s] = (v[0][s] + v[1][s] + v[2][s]) / 3.0;
            }
            for (int q = 0; q < 3; ++q) {
                int fa = -1, fbid = -1;
                const int ia = artra_point_in_polyhedron(pts[q], poly, NULL, &fa);
                const int ib = artra_point_in_polyhedron(pts[q], poly, &fb, &fbid);
                if (ia != ib || fa != fbid) {
                    printf("facet search differs on facet %d point %d: %d/%d vs %d/%d\n", f, q, ia, fa, ib, fbid);
                    return 1;
                }
            }
        }
        artra_free_facet_bounds(&fb);
        if (0 == tested) return 1;
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 7) {
        fprintf(stderr, "usage: scope_check <case dir> <moves> <dx> <dy> <dz> <az>\n");
        return 2;
    }
    const int moves = atoi(argv[2]);
    const Real off[DIMS] = {atof(argv[3]), atof(argv[4]), atof(argv[5])};
    const Real ang[DIMS] = {0.0, 0.0, atof(argv[6])};
    Time ta, tb;
    Space sa, sb;
    Model ma, mb;
    memset(&ta, 0, sizeof ta); memset(&sa, 0, sizeof sa); memset(&ma, 0, sizeof ma);
    memset(&tb, 0, sizeof tb); memset(&sb, 0, sizeof sb); memset(&mb, 0, sizeof mb);
    if (artra_prepare_case(argv[1], &ta, &sa, &ma) || artra_prepare_case(argv[1], &tb, &sb, &mb)) {
        fprintf(stderr, "%s\n", artra_error_text());
        return 2;
    }
    if (FacetSearch(&sa)) return 1;
    ArtraGeometryScope scope;
    if (artra_scope_init(&sb, &scope)) return 2;
    const long N = (long)sa.part.n[X] * sa.part.n[Y] * sa.part.n[Z];
    long joined = 0;
    for (int m = 0; m < moves; ++m) {
        Move(&sa, off, ang);
        Move(&sb, off, ang);
        ArtraGhostTable ja, jb;
        memset(&ja, 0, sizeof ja);
        memset(&jb, 0, sizeof jb);
        int region[DIMS][LIMIT];
        artra_scope_region(&sb, &scope, region);
        if (artra_compute_geometric_field(&sa, &ma, &ja, NULL) ||
            artra_compute_geometric_field(&sb, &mb, &jb, (const int (*)[LIMIT])region)) {
            fprintf(stderr, "%s\n", artra_error_text());
            return 2;
        }
        artra_scope_update(&sb, &scope, &jb, (const int (*)[LIMIT])region);
        for (long q = 0; q < N; ++q) {
            const Node *a = sa.node + q, *b = sb.node + q;
            if (a->did != b->did || a->fid != b->fid || a->lid != b->lid || a->gst != b->gst) {
                printf("move %d: node %ld flags %d %d %d %d vs %d %d %d %d\n", m, q, a->did, a->fid, a->lid, a->gst,
                       b->did, b->fid, b->lid, b->gst);
                return 1;
            }
        }
        if (!SameJoins(&ja, &jb)) {
            printf("move %d: reconstruction lists differ (%d vs %d nodes)\n", m, ja.nghost, jb.nghost);
            return 1;
        }
        if (!SameGhostTables(&sb, &mb)) {
            printf("move %d: ghost tables differ between %d threads and one\n", m, omp_get_max_threads());
            return 1;
        }
        joined += ja.nghost;
        artra_free_ghost_table(&ja);
        artra_free_ghost_table(&jb);
    }
    artra_scope_free(&scope);
    printf("ok %ld\n", joined);
    return 0;
}
