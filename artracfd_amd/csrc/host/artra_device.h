/*
 * artra_device.h - C interface between the host side and the HIP device
 * side of libartra_hip.  All device memory is owned by an ArtraDevice; the
 * host passes plain arrays (structure-of-arrays fields, flag masks, tables).
 *
 * The host exchanges a time level as component-major arrays [5][N] in the
 * solver's index order idx = (k * ny + j) * nx + i; on the device each
 * level is node-major, U[level][idx][5].
 */
#ifndef ARTRA_DEVICE_H_
#define ARTRA_DEVICE_H_

#include <stdint.h>
#include "artra_host.h"
#include "artra_hip.h"
#include "geomcore.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ArtraDevice ArtraDevice;

typedef struct {
    int n[DIMS];             /* padded node counts */
    int ng[DIMS];
    int m[DIMS];             /* cells (after the at-least-two rule) */
    int pin[DIMS][LIMIT];    /* interior range (ns[PIN]) */
    int collapse;
    int multidim;
    int tScheme;
    int sScheme;
    int jacobMean;
    int fluxSplit;
    double gamma, gasR, cv, refMu, refT;
    int sState;              /* gravity source operator on */
    double g[DIMS];          /* gravity (nondimensional) */
    double d[DIMS], dd[DIMS];
    int typeBC[NBC];
    int normal[NBC][DIMS];
    double varBC[NBC][VARBC];
} ArtraDevParams;

/* One layer-2 ghost node of a body in the surface-force integration
 * (solid_dynamics.c:96-148): its padded-grid index, the body's slot in the
 * body table, the normal at its boundary point, the boundary point relative
 * to the body centre and its distance from the ghost node. */
typedef struct {
    int node;
    int body;
    double N[DIMS];
    double r[DIMS];
    double dist;
} ArtraForceNode;

/* A body of the surface-force integration: its nodes [first, first+count)
 * in the reference's (k, j, i) order, whether the wall shear counts
 * (viscous flow and no-slip body), its velocity and angular velocity. */
typedef struct {
    int first;
    int count;
    int viscous;
    int pad;
    double V[DIMS];
    double W[DIMS];
} ArtraForceBody;

/* kernel timing categories */
enum { ARTRA_T_SWEEP = 0, ARTRA_T_BC = 1, ARTRA_T_IB = 2, ARTRA_T_CFL = 3, ARTRA_T_COUNT = 4 };

int artra_dev_create(const ArtraDevParams *par, const uint8_t *fluid, ArtraDevice **out);
void artra_dev_destroy(ArtraDevice *dev);
/* box groups of the external boundary: mode 0 the full fill, 1 + s the
 * faces of axis s only */
int artra_dev_set_boundary(ArtraDevice *dev, int mode, const ArtraBcBox *boxes, int nbox, const int *group_start,
                           int ngroup);
/* resolved fill tables (used instead of the box groups when resolved) */
int artra_dev_set_gather(ArtraDevice *dev, const ArtraBcGather *gather);
/* use_gather = 0 forces the box groups; partial = 0 fills everything after
 * every stage */
int artra_dev_set_fill_options(ArtraDevice *dev, int use_gather, int partial);
/* the resolved fill between the stages of a sweep along s when the full
 * fill is not resolved (artra_bc_resolve_axis) */
int artra_dev_set_axis_gather(ArtraDevice *dev, int s, const ArtraBcGather *gather);
int artra_dev_set_ghosts(ArtraDevice *dev, const ArtraGhostTable *table);
/* the nodes a moving body has just uncovered: their reconstruction
 * (artra_add_join entries) evaluated on level `level` */
int artra_dev_reconstruct(ArtraDevice *dev, int level, const ArtraGhostTable *joins);
/* new node flags after the bodies moved: fluid[q] != 0 for a fluid node
 * (the whole grid's array); box (may be NULL: all) bounds the nodes copied,
 * [min, max) per axis */
int artra_dev_set_fluid(ArtraDevice *dev, const uint8_t *fluid, const int (*box)[2]);
/* ---- moving bodies: the node classification on the device ----
 * Node coordinates x = smin + (i - ng) d (cfd_commons.c:348-351) and the
 * interfacial search path (path offsets, layer r ending at pathSep[r]). */
typedef struct {
    double smin[DIMS], d[DIMS];
    int ng[DIMS];
    int path[PATHN][DIMS];
    int pathSep[PATHSEP];
    int gl;
} ArtraGeoSetup;

/* A body as the classification takes it: moving (state != 1), the node box
 * its domain field searches ([min, max) per axis), the sphere (faceN <= 0)
 * or the triangulated body's host arrays in their current position. */
typedef struct {
    int moving;
    int box[DIMS][LIMIT];
    double O[DIMS], r;
    ArtraPolyView pv;
    double V[DIMS], W[DIMS];  /* velocities of level TO (ghost wall velocity) */
    double cf, T;             /* friction (> 0 no-slip), wall temperature (< 0 adiabatic) */
} ArtraGeoBody;

/* The immersed-boundary ghost table built on the device from its flags
 * (artra_build_ghost_table's entries; within a layer the order may
 * differ, which the fill does not depend on).  *fallback = 1 when a
 * stencil search needs more than radius 3 (the caller then builds the
 * table on the host). */
int artra_dev_build_ghosts(ArtraDevice *dev, const ArtraGeoBody *bodies, int nbody, double dd[DIMS], double tinyL,
                           int ibmLayer, int *fallback);

/* The surface-force table built on the device from its flags: the nodes,
 * their order and geometry of artra_build_force_table.  Slot b integrates
 * body slotBody[b] over its node box box[b] with the wall shear when
 * viscous[b]; gstN[b] and lidN[b] return its layer-2 ghost and interfacial
 * node counts. */
int artra_dev_build_force_table(ArtraDevice *dev, const ArtraGeoBody *bodies, int nbody, const int *slotBody,
                                const int (*box)[DIMS][LIMIT], const int *viscous, int nb, int *gstN, int *lidN);

/* The flags did, fid, lid, gst ([4][N], host order) onto the device. */
int artra_dev_geo_init(ArtraDevice *dev, const ArtraGeoSetup *setup, const int *flags, int nbody);
/* ComputeGeometricField's three passes (immersed_boundary.c:65-264) over
 * region on the device (the domain field over each moving body's box), the
 * sweeps' fluid flags updated; *packed points to [5][*count] ints, the
 * region's new did, fid, lid, gst and previous domain field in k-j-i order,
 * held by the device until its next classification. */
int artra_dev_geo_classify(ArtraDevice *dev, const ArtraGeoBody *bodies, int nbody, const int region[DIMS][LIMIT],
                           const int **packed, long *count);

/* replay the device march's step from a captured graph (1, the default) or
 * launch its kernels one by one (0) */
int artra_dev_set_graph(ArtraDevice *dev, int on);
/* copy one time level: soa is [5][N] doubles */
int artra_dev_upload(ArtraDevice *dev, int level, const double *soa);
int artra_dev_download(ArtraDevice *dev, int level, double *soa);
/* set every byte of the three levels to 0x01 (the solver's poison pattern) */
int artra_dev_poison(ArtraDevice *dev);
int artra_dev_treat_boundary(ArtraDevice *dev, int level);
/* the immersed-boundary ghost fill alone */
int artra_dev_treat_immersed(ArtraDevice *dev, int level);
/* per-axis maxima of |u_s| + c over interior fluid nodes of level TO,
 * folded with the given floors (body speeds) */
int artra_dev_max_speed(ArtraDevice *dev, const double floor[DIMS], double vmax[DIMS]);
/* one EvolveFluidDynamics step of size dt */
int artra_dev_evolve(ArtraDevice *dev, double dt);
/* up to ARTRA_DEV_MARCH_CHUNK whole time steps with dt computed on the
 * device (CFL maxima folded with the floors, the solver loop's clock,
 * EvolveFluidDynamics), no host round trip between steps.  clock =
 * {now, end, rcInt, tmInt, stepC, stepN, halt, rcData[5], dtData[5]}: the
 * loop's time, output-instant and step counters and its output cadence
 * (solve.c:57-104); the march stops moving the clock after the first step
 * at which an output falls due (halt becomes 1) or the loop ends.  clock
 * is updated, dts[i] receives each step's dt (0 once the clock stopped),
 * *advanced the steps that moved it */
#define ARTRA_DEV_MARCH_CHUNK 64
#define ARTRA_DEV_MARCH_CLOCK 17
int artra_dev_march_steps(ArtraDevice *dev, double clock[ARTRA_DEV_MARCH_CLOCK], const double floor[DIMS],
                          double cfl, int steps, double *dts, int *advanced);
/* the same with the device scratch in a static per-GPU array (no device
 * allocation at all); at most ARTRA_DEV_NOALLOC_STEPS steps per call */
#define ARTRA_DEV_NOALLOC_STEPS ARTRA_NOALLOC_MAX_STEPS
int artra_dev_march_noalloc(ArtraDevice *dev, double clock[ARTRA_DEV_MARCH_CLOCK], const double floor[DIMS],
                            double cfl, int steps, double *dts, int *advanced);
int artra_dev_sync(ArtraDevice *dev);
/* surface force: the table once, then per call the per-body sums of the
 * pressure force, wall shear and torque over its nodes in table order and
 * the running pressure statistics of solid_dynamics.c:120-123, sums[b][12]
 * = Fp[3], Fv[3], Tt[3], fvar[3] */
int artra_dev_set_force_table(ArtraDevice *dev, const ArtraForceNode *nodes, int count, const ArtraForceBody *bodies,
                              int nbody);
int artra_dev_surface_force(ArtraDevice *dev, double *sums);
/* primitive variables (rho, u, v, w, p, T) of level TO at count nodes:
 * prim[count][6] (the probes) */
int artra_dev_gather_primitive(ArtraDevice *dev, const int *nodes, int count, double *prim);
/* kernel timing: when on, every launch is bracketed by events on the
 * solver's stream; totals are in milliseconds. */
int artra_dev_set_timing(ArtraDevice *dev, int on);
int artra_dev_get_timing(ArtraDevice *dev, double ms[ARTRA_T_COUNT], long launches[ARTRA_T_COUNT]);
/* arithmetic of the flux kernels: 1 keeps the reference's operation order
 * (bit-faithful), 0 (default) uses the fused/reciprocal form */
int artra_dev_set_exact(ArtraDevice *dev, int exact);
/* ------------------------------------------------ learned indicator */
/* FP32 parameter vector of the 12-32-32-1 MLP (layout of indicator.hip) */
#define ARTRA_MLP_PARAMS 1633
#define ARTRA_MLP_W1 0
#define ARTRA_MLP_B1 512
#define ARTRA_MLP_W2 544
#define ARTRA_MLP_B2 1568
#define ARTRA_MLP_W3 1600
#define ARTRA_MLP_B3 1632

/* route WENO5 faces through the indicator (params NULL disables) */
int artra_dev_set_indicator(ArtraDevice *dev, const float *params, float threshold);
/* faces computed with the linear flux, faces classified, and faces of lines
 * that overflowed their queue (plain WENO5) since the last reset */
int artra_dev_hybrid_counts(ArtraDevice *dev, unsigned long long counts[ARTRA_HYBRID_TALLIES], int reset);
int artra_dev_set_fused(ArtraDevice *dev, int fused);
int artra_dev_set_indicator_options(ArtraDevice *dev, int every_stage, float smooth_variation);
int artra_dev_fused_axes(ArtraDevice *dev);
/* faces along one axis of the interior lines, and their features/labels on level TO */
long artra_dev_face_count(ArtraDevice *dev, int axis);
int artra_dev_sample_faces(ArtraDevice *dev, int axis, float *feat, float *label);
/* training / gradient / inference on the device (indicator.hip) */
int artra_mlp_train_dev(float *params, const float *feat, const float *label, long n, const ArtraMlpTraining *opt,
                        float *loss_history);
int artra_mlp_gradient_dev(const float *params, const float *feat, const float *label, long n, float tau, float wpos,
                           float *grad, float *loss);
int artra_mlp_infer_dev(const float *params, const float *feat, long n, float *logit);
const char *artra_mlp_error(void);

/* batched cases: many independent cases advanced together, one persistent
 * launch per directional operator over the lines of every case */
typedef struct ArtraBatch ArtraBatch;
/* clock[c] = {now, end, rcInt, tmInt, stepC, stepN} of case c (6 doubles);
 * cfl[c] its CFL number */
int artra_dev_batch_create(ArtraDevice *const *devs, int count, const double *clock, const double *cfl,
                           ArtraBatch **out);
void artra_dev_batch_destroy(ArtraBatch *b);
int artra_dev_batch_set_indicator(ArtraBatch *b, const float *params, float threshold);
int artra_dev_batch_counts(ArtraBatch *b, unsigned long long counts[ARTRA_HYBRID_TALLIES], int reset);
int artra_dev_batch_march(ArtraBatch *b, int steps, int use_graph, double *ms);
int artra_dev_batch_clock(ArtraBatch *b, double *clock);
int artra_dev_batch_set_clock(ArtraBatch *b, const double *clock);
int artra_dev_batch_set_threads(ArtraBatch *b, const int *threads);
int artra_dev_batch_tune(ArtraBatch *b, int reps, int *threads, int *grid);

/* tuning knobs (0 keeps the default) */
int artra_dev_set_tuning(ArtraDevice *dev, int march_segment, int xblock);
/* streaming copy of `bytes` bytes, iters times; returns read+write GB/s */
int artra_dev_copy_probe(long bytes, int iters, double *gbps);
const char *artra_dev_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ARTRA_DEVICE_H_ */
