/*
 * solver.h - the resident solver's host state, shared by the C-ABI entry
 * points (capi.c) and the output side (output.c: surface force, probes,
 * field files, restart, the solver loop with its output cadence).
 */
#ifndef ARTRA_SOLVER_H_
#define ARTRA_SOLVER_H_

#include "artra_device.h"
#include "artra_host.h"

/* A body whose surface force the solver integrates (state <= 0,
 * solid_dynamics.c:86-88): its index in Geometry, the interfacial layer-2
 * node count that calibrates the sum, and its layer-2 ghost nodes. */
typedef struct {
    int body;
    int lidN;
    int gstN;
} ArtraForceBodyInfo;

struct ArtraHipSolver {
    Space *space;
    const Model *model;
    ArtraDevice *dev;
    ArtraGhostTable ghosts;
    long N;
    double *staging;          /* [5][N] host buffer for AoS <-> SoA */
    Real rcInt;               /* time since the last output instant (solve.c:64) */
    int forceBodies;          /* bodies in force[] */
    ArtraForceBodyInfo *force;
    ArtraGeometryScope scope; /* moving bodies: where the next classification must look */
    uint8_t *fluid;           /* [N] host copy of the device's fluid flags */
};

/* The node box of a body's surface-force scan: its bounding box over the
 * interior range. */
void artra_force_body_box(const Partition *part, const Polyhedron *poly, int box[DIMS][LIMIT]);
/* Build the surface-force table of the solver's bodies and hand it to the
 * device (at attach). */
int artra_build_force_table(ArtraHipSolver *s);
/* Error from the device side, as an artra_set_error code. */
int artra_dev_failed(void);
/* The device tables that follow the node classification (fluid flags,
 * immersed-boundary ghosts, surface-force nodes), rebuilt from the host
 * flags after the bodies moved; region (may be NULL: everywhere) bounds the
 * nodes whose flags changed. */
int artra_hip_refresh_geometry(ArtraHipSolver *s, const int (*region)[LIMIT]);
/* The device's copy of the node classification, from the host's flags
 * (at attach and after a restart of a moving case). */
int artra_geo_device_init(ArtraHipSolver *s);
/* EvolveSolidDynamics (solid_dynamics.c:37-48) over dt at time now. */
int artra_solid_step(ArtraHipSolver *s, Real now, Real dt);
/* One pass of the time loop's body with moving bodies (solve.c:65-89): the
 * step count, dt with its rectifications (rcInt, tmInt), the clock, and the
 * solid half-steps around the fluid step.  *dt receives the step taken. */
int artra_moving_step(ArtraHipSolver *s, Time *time, Real tmInt, Real *rcInt, Real *dt);

#endif /* ARTRA_SOLVER_H_ */
