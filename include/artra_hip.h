/*
 * artra_hip.h - C-ABI of libartra_hip, the MI355X (gfx950) implementation of
 * ArtraCFD's fluid-update hot path.
 *
 * Three groups of entry points:
 *
 *  1. Drop-in symbols with the reference's own signatures.  A solver built
 *     from the reference sources links against these instead of
 *     fluid_dynamics.c / boundary_treatment.c and runs unchanged; each call
 *     moves the node array to the device and back, so it is correct but
 *     bound by host<->device copies.
 *
 *  2. The resident API (artra_hip_*): attach once, keep the field in HBM,
 *     step with the device CFL reduction, copy back only for output.  Every
 *     function returns 0 on success and a non-zero code otherwise; the
 *     message is available from artra_hip_error().
 *
 *  3. Case preparation (artra_case_*): reads the solver's case directory
 *     (artracfd.case, artracfd.geo, STL) into Space/Model/Time exactly as the
 *     reference's Preprocess and initialisation do, for callers that do not
 *     link the reference at all.
 *
 * All structures are those of artra_types.h (layout-identical to commons.h).
 */
#ifndef ARTRA_HIP_H_
#define ARTRA_HIP_H_

#include "artra_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------- 1. drop-in */

/* Advance space->node[].U[TO] by one time step dt: Strang-split (or
 * unsplit) Runge-Kutta stages with the boundary and immersed-boundary fill
 * after every stage.  Replaces fluid_dynamics.c:55-116
 * (fluid_dynamics.h:32).  Errors are fatal (message on stderr, exit(1)), as
 * in the reference. */
void EvolveFluidDynamics(const Real dt, Space *space, const Model *model);

/* Fill the domain-boundary and immersed-boundary ghost nodes of level tn.
 * Replaces boundary_treatment.c:28-65 (boundary_treatment.h:32). */
void TreatBoundary(const int tn, Space *space, const Model *model);

/* ----------------------------------------------------- 2. resident API */

typedef struct ArtraHipSolver ArtraHipSolver;

/* Upload level TO of space (all nodes) and build the boundary tables.  The
 * solver keeps pointers to space and model; both must outlive it.  Returns
 * 2 for a case this library does not support (message says which). */
int artra_hip_attach(Space *space, const Model *model, ArtraHipSolver **out);
int artra_hip_detach(ArtraHipSolver *solver);

/* TreatBoundary on the device for time level tn (0, 1 or 2). */
int artra_hip_treat_boundary(ArtraHipSolver *solver, int tn);

/* The CFL time step of solve.c:108-153: numCFL * min_s(d_s / Vmax_s). */
int artra_hip_compute_dt(ArtraHipSolver *solver, const Time *time, Real *dt);

/* EvolveFluidDynamics on the device-resident field (fluid_dynamics.c:55-116). */
int artra_hip_evolve(ArtraHipSolver *solver, Real dt);

/* The time loop of EvolveSolution (solve.c:48-107) without output: advance
 * until time->now reaches time->end, time->stepC reaches time->stepN or
 * max_steps steps were taken.  dts (may be NULL) receives each step's dt. */
int artra_hip_march(ArtraHipSolver *solver, Time *time, int max_steps, int *steps_done, Real *dts);

/* Copy a time level between the device and the node array of space. */
int artra_hip_download(ArtraHipSolver *solver, int tn);
int artra_hip_upload(ArtraHipSolver *solver, int tn);

/* Copy a time level between the device and a component-major array
 * field[5][nx*ny*nz] (no node array involved). */
int artra_hip_read_field(ArtraHipSolver *solver, int tn, Real *field);
int artra_hip_write_field(ArtraHipSolver *solver, int tn, const Real *field);

/* Wait for all queued device work. */
int artra_hip_synchronize(ArtraHipSolver *solver);

/* Kernel timing with HIP events on the solver's stream.  Categories:
 * 0 directional sweeps, 1 domain boundary, 2 immersed boundary, 3 CFL.
 * Turning timing on resets the totals. */
int artra_hip_set_timing(ArtraHipSolver *solver, int on);
int artra_hip_get_timing(ArtraHipSolver *solver, double ms[4], long launches[4]);

/* Flux arithmetic.  exact = 1 evaluates every expression in the reference's
 * operation order without contraction (results bit-identical to the CPU
 * for inviscid cases); exact = 0, the default, uses fused multiply-adds,
 * shared reciprocals, the structured characteristic projection and
 * single-division WENO weights (rounding-level differences, far inside the
 * 1e-10 relative-L2 parity contract).  The environment variable
 * ARTRA_EXACT=1 selects the exact form at attach time. */
int artra_hip_set_arithmetic(ArtraHipSolver *solver, int exact);

/* Boundary-fill strategy.  use_gather = 1 (default) replaces the ordered box
 * sequence by one gather wherever replaying the order on the host proves
 * every ghost a copy of an interior node or an inflow state (outflow,
 * periodic and inflow faces; walls keep the ordered boxes).  partial = 1
 * (default) fills only the sweep axis's ghosts between the stages of an
 * inviscid split sweep, the only ghosts those stages read.  Both leave the
 * result unchanged; 0, 0 restores the literal fill after every stage. */
int artra_hip_set_fill_options(ArtraHipSolver *solver, int use_gather, int partial);

/* Launch-shape knobs: column segment of the y/z sweeps (0 keeps the default). */
int artra_hip_set_tuning(ArtraHipSolver *solver, int march_segment);

/* Device copy bandwidth probe: copies `bytes` bytes `iters` times with a
 * plain streaming kernel and reports read+write GB/s.  Used to calibrate
 * HBM byte counters and to state the achievable copy bandwidth. */
int artra_hip_copy_probe(long bytes, int iters, double *gbps);

/* Message of the last failed call on this thread. */
const char *artra_hip_error(void);

/* ------------------------------------------------ 3. case preparation */

typedef struct ArtraCase ArtraCase;

/* Preprocess + initialisation of the case in directory dir, up to (not
 * including) the first boundary treatment. */
int artra_case_open(const char *dir, ArtraCase **out);
void artra_case_close(ArtraCase *c);
Space *artra_case_space(ArtraCase *c);
Model *artra_case_model(ArtraCase *c);
Time *artra_case_time(ArtraCase *c);
/* Grid description: n[3], ng[3], pin[3][2] (interior ranges). */
int artra_case_grid(const ArtraCase *c, int n[3], int ng[3], int pin[6]);
/* Component-major copy of one level of the node array, and its flags
 * flags[4][N] = did, fid, lid, gst. */
int artra_case_get_field(const ArtraCase *c, int tn, Real *field);
int artra_case_set_field(ArtraCase *c, int tn, const Real *field);
int artra_case_get_flags(const ArtraCase *c, int *flags);
/* Evaluate an initial-condition expression (t = 0); returns 0 on success. */
int artra_calc_expression(const char *text, Real x, Real y, Real z, Real *ans);

#ifdef __cplusplus
}
#endif

#endif /* ARTRA_HIP_H_ */
