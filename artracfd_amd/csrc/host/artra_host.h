/*
 * artra_host.h - internal interface of the host side of libartra_hip.
 *
 * The host side prepares a case exactly as the reference solver does before
 * its time loop (case files -> Space/Model/Time, node flags, initial field,
 * geometry), builds the static tables the device kernels consume (boundary
 * boxes, immersed-boundary ghost stencils), and drives the device time loop.
 * Everything here is plain C; the device side is reached through the
 * functions declared in artra_device.h.
 */
#ifndef ARTRA_HOST_H_
#define ARTRA_HOST_H_

#include <stddef.h>
#include <stdint.h>
#include "artra_types.h"
#include "geomcore.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
/* Record a formatted error message; returns the error code for chaining. */
int artra_set_error(int code, const char *fmt, ...);
const char *artra_error_text(void);

/* ------------------------------------------------------------- utilities */
/* Normalise a case-file line in place (case grammar of commons.c:26-84):
 * strip the newline and everything after '#', collapse blanks. */
void artra_parse_line(char *line);
int artra_index(int k, int j, int i, int ny, int nx);
int artra_map_node(Real s, Real smin, Real dds, int ng);
int artra_confine(int n, int nmin, int nmax);
Real artra_map_point(int n, Real smin, Real ds, int ng);
int artra_in_box(int k, int j, int i, const int box[DIMS][LIMIT]);
Real artra_dot(const Real a[DIMS], const Real b[DIMS]);
Real artra_norm(const Real a[DIMS]);
Real artra_dist2(const Real a[DIMS], const Real b[DIMS]);
void artra_cross(const Real a[DIMS], const Real b[DIMS], Real out[DIMS]);
void artra_normalize(int n, Real by, Real *v);
void artra_orthogonal_space(const Real N[DIMS], Real Ta[DIMS], Real Tb[DIMS]);
void artra_map_primitive(Real gamma, Real gasR, const Real U[DIMU], Real Uo[DIMUo]);
void artra_map_conservative(Real gamma, const Real Uo[DIMUo], Real U[DIMU]);

/* ------------------------------------------------------------ calculator */
/* A compiled initial-condition expression (grammar of calculator.c:126-718). */
typedef struct ArtraExpr ArtraExpr;
int artra_expr_compile(const char *text, ArtraExpr **out);
/* Evaluate with variables t, x, y, z; returns 0 on success.  On a domain
 * error (ln of a non-positive number, division by zero, ...) the result is
 * left at *ans unchanged and 1 is returned, as the reference does. */
int artra_expr_eval(const ArtraExpr *e, Real t, Real x, Real y, Real z, Real *ans);
void artra_expr_free(ArtraExpr *e);
/* One-shot evaluation (used by tests): returns 0 on success. */
int artra_calc(const char *text, Real t, Real x, Real y, Real z, Real *ans);

/* -------------------------------------------------------------- case set-up */
/* Read artracfd.case and artracfd.geo in dir (case_loader.c:32-253). */
int artra_load_case(const char *dir, Time *time, Space *space, Model *model);
/* Ghost counts, collapse flags, nondimensionalisation (cfd_parameters.c:28-186). */
int artra_compute_parameters(Time *time, Space *space, Model *model);
/* Index ranges, boundary normals, search path (domain_partition.c:19-150). */
void artra_partition_domain(Space *space);
/* Node and geometry storage (preprocess.c:47-59). */
int artra_allocate(Space *space, Model *model);
/* Initial field and geometry (initialization.c:38-162) up to, not including,
 * the first boundary treatment. */
int artra_initialize_domain(const char *dir, Time *time, Space *space, Model *model);
/* Whole preparation: load, parameters, partition, allocate, initialise. */
int artra_prepare_case(const char *dir, Time *time, Space *space, Model *model);
/* Release everything artra_prepare_case allocated. */
void artra_release_case(Space *space, Model *model, Time *time);

/* Immersed-boundary stencil table (defined with the device-side tables). */
typedef struct ArtraGhostTable ArtraGhostTable;

/* --------------------------------------------------------------- geometry */
int artra_read_stl(const char *fname, Polyhedron *poly);
int artra_convert_polyhedron(Polyhedron *poly);
void artra_transform_polyhedron(const Real O[DIMS], const Real scale[DIMS], const Real angle[DIMS],
                                const Real offset[DIMS], Polyhedron *poly);
void artra_compute_geometry_parameters(int collapse, Geometry *geo);
Real artra_point_triangle_distance(const Real p[DIMS], const Real v0[DIMS], const Real e01[DIMS],
                                   const Real e02[DIMS], Real para[DIMS]);
/* Bounds that speed up the closest-facet search of a triangulated body in
 * its current position: the box of every facet, of every block of
 * ARTRA_FACET_BLOCK consecutive facets and of all facets. */
#define ARTRA_FACET_BLOCK 32
typedef struct {
    int blocks;
    Real (*facet)[2 * DIMS];   /* min xyz, max xyz */
    Real (*block)[2 * DIMS];
    Real all[2 * DIMS];
} ArtraFacetBounds;
int artra_facet_bounds(const Polyhedron *poly, ArtraFacetBounds *bounds);
void artra_free_facet_bounds(ArtraFacetBounds *bounds);
/* The arrays of a triangulated body as the shared closest-point queries
 * (geomcore.h) take them. */
ArtraPolyView artra_poly_view(const Polyhedron *poly);
/* Inside-or-on test of a triangulated body with the closest facet in *fid
 * (computational_geometry.c:463-506); bounds (may be NULL: every facet is
 * measured) only skip facets that cannot be the closest. */
int artra_point_in_polyhedron(const Real p[DIMS], const Polyhedron *poly, const ArtraFacetBounds *bounds, int *fid);
void artra_compute_geometric_data(const Real p[DIMS], int fid, const Polyhedron *poly, Real pi[DIMS],
                                  Real pm[DIMS], Real N[DIMS]);
/* Node classification did/fid/lid/gst (immersed_boundary.c:65-264).  A
 * fluid node that was inside a body before the call is marked fid = -1 and,
 * when joins is not NULL, listed there with its reconstruction stencil
 * (artra_add_join); the caller evaluates the list.  Returns 0 or an error. */
int artra_compute_geometric_field(Space *space, const Model *model, ArtraGhostTable *joins,
                                  const int (*region)[LIMIT]);

/*
 * Classification after a move, restricted to where it can change anything.
 * The two passes of ComputeGeometricField over every interior node only
 * change nodes that a moving body covers before (reset of its interfacial
 * nodes) or after the move (domain field), the interfacial state of nodes
 * within gl layers of those, and the reconstruction mark of the nodes the
 * previous classification reconstructed; every other node keeps its flags.
 * A scope tracks, per body, the node box of its nodes (extent) and the box of
 * the last reconstructed nodes; artra_scope_region gives, after the bodies
 * moved, the box that the passes must cover (region argument above), and
 * artra_scope_update refreshes the scope from the new classification (the
 * moving bodies' nodes all lie inside that region).
 */
/* The host's copy of a classification made on the device over region:
 * packed[5][count] = did, fid, lid, gst and the previous domain field of
 * the region's nodes in k-j-i order (artra_dev_geo_classify).  The nodes a
 * body uncovered are listed in joins (artra_add_join, on the flags as the
 * reference's interfacial pass finds them) and marked fid = -1. */
int artra_apply_classification(Space *space, const int region[DIMS][LIMIT], const int *packed, long count,
                               ArtraGhostTable *joins);

typedef struct {
    int bodies;
    int (*extent)[DIMS][LIMIT];   /* [bodies]: node box [min, max) per axis; empty when min >= max */
    int join[DIMS][LIMIT];
} ArtraGeometryScope;
int artra_scope_init(const Space *space, ArtraGeometryScope *scope);
void artra_scope_free(ArtraGeometryScope *scope);
void artra_scope_region(const Space *space, const ArtraGeometryScope *scope, int region[DIMS][LIMIT]);
void artra_scope_update(const Space *space, ArtraGeometryScope *scope, const ArtraGhostTable *joins,
                        const int region[DIMS][LIMIT]);
void artra_identify_geometry_state(Geometry *geo);

/* ------------------------------------------------- device-side tables */
/* One external-boundary box of TreatBoundary (boundary_treatment.c:51-63). */
typedef struct {
    int face;            /* PWB..PBB */
    int layer;           /* r */
    int box[DIMS][LIMIT];
} ArtraBcBox;

/* Boxes in the reference's order, empty boxes dropped.  Returns the count. */
int artra_bc_boxes(const Space *space, ArtraBcBox *out, int cap);

/* Launch groups: consecutive boxes that touch disjoint node sets and may run
 * in one kernel.  group_start has ngroups+1 entries. */
int artra_bc_groups(const Space *space, const ArtraBcBox *boxes, int nbox, int *group_start, int cap);

/*
 * Resolved form of the external-boundary fill.  Replaying the box order on
 * node labels shows, for outflow/periodic/inflow boundaries, that every
 * node a fill writes ends up as a copy of one interior node or as the inflow
 * state, independently of the order; the whole fill is then one gather.
 * Wall reflections (or an order-dependent result) leave resolved = 0 and the
 * grouped box kernels are used instead.
 */
typedef struct {
    int resolved;
    long count;
    int *dst;            /* padded-grid index written */
    int *src;            /* >= 0: interior source node; -3 - p: inflow state of face p */
    long axis_count[DIMS];
    int *axis_dst[DIMS]; /* the subset outside the interior range along one axis only */
    int *axis_src[DIMS];
} ArtraBcGather;

int artra_bc_resolve(const Space *space, const ArtraBcBox *boxes, int nbox, ArtraBcGather *out);
/* The same replay for the boxes of the faces normal to axis s alone (the fill
 * between the stages of a sweep along s), kept to the nodes outside the
 * interior range along s only, the ghosts that sweep reads: resolved when
 * each of those is a copy of a node the fill does not write or an inflow
 * state (dst/src/count hold them; the axis tables are unused). */
int artra_bc_resolve_axis(const Space *space, const ArtraBcBox *boxes, int nbox, int s, ArtraBcGather *out);
void artra_free_bc_gather(ArtraBcGather *g);

/*
 * Immersed-boundary ghost table.  For a stationary body the stencil of every
 * ghost node is fixed, so the host precomputes, per ghost node and in the
 * reference's exact summation order, the stencil node indices and weights of
 * InverseDistanceWeighting (immersed_boundary.c:413-473) and the geometry of
 * ReconstructFlow (immersed_boundary.c:357-412).
 */
typedef struct {
    int node;            /* padded-grid index of the ghost node */
    int kind;            /* 0: image-point reconstruction, 1: IDW of layer r-1 */
    int first;           /* first stencil entry */
    int count;           /* number of stencil entries */
    int noslip;          /* cf > 0 */
    int isothermal;      /* wall T >= 0 */
    Real wallT;
    Real wO;             /* weight of the boundary point (correction step) */
    Real Vs[DIMS];       /* body velocity at the boundary point */
    Real N[DIMS];
    Real Ta[DIMS];
    Real Tb[DIMS];
} ArtraGhost;

struct ArtraGhostTable {
    int layers;          /* gl */
    int *layer_start;    /* layers+1 offsets into ghosts (ghosts grouped by layer) */
    ArtraGhost *ghosts;
    int nghost;
    int *stencil_node;
    Real *stencil_w;
    int nstencil;
    int ghost_cap;       /* allocated entries (tables grown one node at a time) */
    int stencil_cap;
};

int artra_build_ghost_table(const Space *space, const Model *model, ArtraGhostTable *table);
void artra_free_ghost_table(ArtraGhostTable *table);

/*
 * Nodes that a moving body has just uncovered (immersed_boundary.c:193-204):
 * artra_add_join appends node (k, j, i) as an entry of kind 1 (normalized
 * inverse-distance weighting) whose stencil is the search from radius 2 over
 * the fluid nodes that were fluid before the move (gst, still holding the
 * previous domain field, is 0, and fid is not the reconstruction mark),
 * read from the flags before the interfacial pass changes them.
 * artra_apply_joins evaluates the
 * entries on the host node array (level tn); the device evaluates the same
 * table with the immersed-boundary kernel.
 */
int artra_add_join(const Space *space, int k, int j, int i, ArtraGhostTable *joins);
void artra_apply_joins(Space *space, const Model *model, const ArtraGhostTable *joins, int tn);

#ifdef __cplusplus
}
#endif

#endif /* ARTRA_HOST_H_ */
