/*
 * artra_hip.h - C-ABI of libartra_hip, the MI355X (gfx950) implementation of
 * ArtraCFD's fluid-update hot path.
 *
 * Three groups of entry points:
 *
 *  1. Drop-in symbols with the reference's own signatures.  A solver built
 *     from the reference sources links against these instead of
 *     fluid_dynamics.c, boundary_treatment.c and immersed_boundary.c (and the
 *     flux files only they call) and runs unchanged; each device call moves
 *     the node array to the device and back, so it is correct but bound by
 *     host<->device copies.
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

/* Fill the immersed-boundary ghost nodes of level tn only.  Replaces
 * immersed_boundary.c:270-340 (immersed_boundary.h:39), called by the
 * reference's solid dynamics (solid_dynamics.c:47). */
void TreatImmersedBoundary(const int tn, Space *space, const Model *model);

/* Classify every interior node against the bodies: did (inside body k),
 * fid (closest facet), lid and gst (interfacial and ghost layers), on the
 * host, and reconstruct in node[].U[TO] the fluid nodes a moving body has
 * uncovered (inverse-distance weighting over the nodes that were fluid
 * before).  Replaces immersed_boundary.c:65-264 (immersed_boundary.h:32),
 * called at initialisation (initialization.c:47) and after every move of
 * the bodies (solid_dynamics.c:46).  The device solver cached for space
 * rebuilds the tables that derive from the flags. */
void ComputeGeometricField(Space *space, const Model *model);

/* ----------------------------------------------------- 2. resident API */

/* Face tallies of the learned hybrid (artra_hip_hybrid_counts). */
#define ARTRA_HYBRID_TALLIES 4

typedef struct ArtraHipSolver ArtraHipSolver;

/* Upload level TO of space (all nodes) and build the boundary tables.  The
 * solver keeps pointers to space and model; both must outlive it.  Returns
 * 2 for a case this library does not support (message says which). */
int artra_hip_attach(Space *space, const Model *model, ArtraHipSolver **out);
int artra_hip_detach(ArtraHipSolver *solver);

/* TreatBoundary on the device for time level tn (0, 1 or 2). */
int artra_hip_treat_boundary(ArtraHipSolver *solver, int tn);
/* The immersed-boundary part of it alone (TreatImmersedBoundary). */
int artra_hip_treat_immersed_boundary(ArtraHipSolver *solver, int tn);

/* The CFL time step of solve.c:108-153: numCFL * min_s(d_s / Vmax_s). */
int artra_hip_compute_dt(ArtraHipSolver *solver, const Time *time, Real *dt);

/* EvolveFluidDynamics on the device-resident field (fluid_dynamics.c:55-116). */
int artra_hip_evolve(ArtraHipSolver *solver, Real dt);

/* The time loop of EvolveSolution (solve.c:48-107) without output: advance
 * until time->now reaches time->end, time->stepC reaches time->stepN or
 * max_steps steps were taken.  dts (may be NULL) receives each step's dt.
 * With moving bodies (model->psi != 0) every step also takes the solid
 * half-steps of solid_dynamics.c:37-48 before and after the fluid step
 * (surface force on the device; kinematics, collisions unless psi == 1,
 * motion and node classification on the host; the uncovered nodes'
 * reconstruction and the ghost fill on the device), updating the bodies in
 * space->geo and the node flags in space->node. */
int artra_hip_march(ArtraHipSolver *solver, Time *time, int max_steps, int *steps_done, Real *dts);

/* The time loop of artra_hip_march for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc); no counterpart in the reference.  Same steps, same dt
 * sequence and same field as artra_hip_march, all max_steps steps enqueued
 * in one call with no host round trip between them (steps past the end
 * time change nothing).  Its device scratch (CFL speed maxima, clock, the
 * record of the steps) is a static array of each GPU, not an allocation:
 * max_steps is at most ARTRA_NOALLOC_MAX_STEPS, kernel timing must be off,
 * and calls that overlap on one GPU share that array and race.  Moving
 * bodies (psi != 0) are refused: their solid steps rebuild device tables. */
#define ARTRA_NOALLOC_MAX_STEPS 1024
int artra_hip_march_noalloc(ArtraHipSolver *solver, Time *time, int max_steps, int *steps_done, Real *dts);

/* IntegrateSurfaceForce (solid_dynamics.c:50-171) on the device-resident
 * level TO: for every body with state <= 0, the pressure force, the wall
 * shear of a no-slip body in a viscous flow and their torque about the
 * body centre, into space->geo.poly[n].Fp, Fv, Tt.  A reduction over the
 * bodies' layer-2 ghost nodes on the device (in the reference's node
 * order); only the per-body sums cross to the host. */
int artra_hip_surface_force(ArtraHipSolver *solver);

/* WriteData for one probe kind (data_probe.c:22-197; kind PROPT point,
 * PROLN line, PROCV body-conformal curve, PROFC surface force): the same
 * files, in the current directory, as the reference writes, from device
 * gathers of the probed nodes only (the field stays on the device).  The
 * surface-force file holds the values of the last artra_hip_surface_force. */
int artra_hip_write_probe_data(ArtraHipSolver *solver, const Time *time, int kind);

/* WriteData(PROSD) (data_stream.c:76-125): the field file of the output
 * range, the bodies' geometry files and artracfd.log, in the current
 * directory, from one download of level TO, as the reference's writer of
 * the case's data streamer makes them: 0, ParaView (paraview_writer.c:
 * 31-398: the same names, collection files, layout and "%.6g" text,
 * formatted in parallel); 1, EnSight Gold (ensight_writer.c:34-452: the
 * transient and per-output case files, C-binary FP32 geometry and variable
 * files, body state files).  Another streamer is refused with code 2. */
int artra_hip_write_space_data(ArtraHipSolver *solver, const Time *time);

/* Solve (solve.c:38-107) on the device after the case's initialisation: the
 * initial outputs unless the case restarts, then the time loop with its
 * output cadence (probes, surface force, field files) written exactly as
 * the reference writes them, from the device-resident field. */
int artra_hip_solve(ArtraHipSolver *solver, Time *time);

/* FP64 restart: level TO of every node (ghosts included), the clock
 * (now, stepC, dataC) and the output-instant recorder, and with moving
 * bodies their whole state (position, orientation, velocities,
 * accelerations, forces) and the node classification, exactly; a run
 * continued from it is the uninterrupted run bit for bit (the reference
 * restarts from its "%.6g" text, paraview_reader.c, data_stream.c:154-188).
 * Reading checks the grid size and bodies, uploads the field and sets
 * time, bodies, flags and solver. */
int artra_hip_write_restart(ArtraHipSolver *solver, const Time *time, const char *path);
int artra_hip_read_restart(ArtraHipSolver *solver, Time *time, const char *path);

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

/* on = 1 (default): artra_hip_march and artra_hip_solve replay each device
 * step (CFL maxima, dt, the operators and fills) from a captured HIP graph,
 * one launch per step; on = 0 launches the kernels one by one.  Same
 * kernels, same results; kernel timing and the learned indicator always
 * launch one by one.  No counterpart in the reference (a tuning switch). */
int artra_hip_set_graph(ArtraHipSolver *solver, int on);

/* Fused directional sweeps.  fused = 1 (default) runs all RK stages of a
 * split directional operator as one launch per grid line with the line
 * held in LDS, wherever a stage along that axis reads nothing but its own
 * line (inviscid, split, no immersed body, between-stage fill copying along
 * the line, line fits in LDS); the final level TO is identical, the
 * intermediate levels TN and TM are left as scratch.  *axes receives a bit
 * mask of the axes that qualify. */
int artra_hip_set_fused(ArtraHipSolver *solver, int fused);
int artra_hip_fused_axes(ArtraHipSolver *solver, int *axes);

/* Launch-shape knobs: column segment of the y/z sweeps (0 keeps the default). */
int artra_hip_set_tuning(ArtraHipSolver *solver, int march_segment);

/* Device copy bandwidth probe: copies `bytes` bytes `iters` times with a
 * plain streaming kernel and reports read+write GB/s.  Used to calibrate
 * HBM byte counters and to state the achievable copy bandwidth. */
int artra_hip_copy_probe(long bytes, int iters, double *gbps);

/* Message of the last failed call on this thread. */
const char *artra_hip_error(void);

/* ------------------------------------------------------------ batched cases */
/*
 * Many independent cases advanced together.  Each time step is one device
 * CFL reduction over every case, one kernel computing every case's dt (with
 * the time loop's rectification, solve.c:65-78), then the split directional
 * operators Z Y X X Y Z, each as ONE persistent launch over the lines of
 * every case (blocks walk a tile list; a tile is 256 / T lines of one case,
 * T threads per line) followed by every case's boundary fill.  Nothing of a
 * step needs the host, so a step can be captured once in a HIP graph and
 * replayed.  No reference counterpart: an addition of this library.
 *
 * Cases must be three-dimensional, fluid-only, inviscid, split WENO5 with
 * the fast arithmetic and share the time scheme and flux model; they run in
 * lockstep.  A case that reaches its end time or its step limit (stepN)
 * inside the marched steps stops changing (dt = 0) from then on, and its
 * step count stops with it, as the solver loop's does (solve.c:65).
 */
typedef struct ArtraHipBatch ArtraHipBatch;
int artra_hip_batch_create(ArtraHipSolver *const *solvers, Time *const *times, int count, ArtraHipBatch **out);
int artra_hip_batch_destroy(ArtraHipBatch *batch);
/* Autotuner: per grid size the threads per line (tile shape), then the
 * persistent grid size; threads[count] and *grid receive the choice.  The
 * fields and clocks are restored afterwards. */
int artra_hip_batch_tune(ArtraHipBatch *batch, int reps, int *threads, int *grid);
/* Set the tile shape instead of tuning it: threads[c] threads per line of
 * case c (64, 128 or 256, at least its longest interior line). */
int artra_hip_batch_set_threads(ArtraHipBatch *batch, const int *threads);
/* Learned indicator for every case (NULL: plain WENO5). */
int artra_hip_batch_set_indicator(ArtraHipBatch *batch, const float *params, float logit_threshold);
int artra_hip_batch_counts(ArtraHipBatch *batch, unsigned long long counts[ARTRA_HYBRID_TALLIES], int reset);
/* steps lockstep time steps; use_graph replays one captured step; *ms the
 * device time of the steps (may be NULL).  Each case starts from its Time
 * (now, stepC) and solver clock as they stand, and its Time advances. */
int artra_hip_batch_march(ArtraHipBatch *batch, int steps, int use_graph, double *ms);

/* ------------------------------------------- learned troubled-face indicator */
/*
 * A 12-32-32-1 MLP (ReLU, BF16 weights and activations on MFMA, FP32
 * accumulation) reads, per face, the six-node stencils of density and
 * pressure normalised by their face average, and routes the face either to
 * full WENO5 or to the linear fifth-order upwind flux.  FP64 is kept for the
 * flux, the characteristic projection and the update.  No reference
 * counterpart: an addition of this library.
 *
 * Parameters: one FP32 vector of ARTRA_MLP_PARAMS values,
 *   W1[16][32] (rows 12..15 zero), b1[32], W2[32][32], b2[32], w3[32], b3,
 * weights stored [input][output].
 */
#define ARTRA_MLP_PARAMS 1633

typedef struct {
    int epochs;          /* passes over the samples */
    int batch;           /* samples per optimiser step (multiple of 64) */
    float lr;            /* peak Adam step size, cosine-decayed to zero */
    float tau;           /* a face is troubled when its label exceeds tau */
    float wpos;          /* loss weight of troubled samples */
} ArtraMlpTraining;

/* Faces along an axis of the interior lines, and their 12 features and
 * WENO5 weight-deviation labels (max over fields and sides of
 * sum_k |omega_k - C_k|) computed on level TO. */
int artra_hip_face_count(ArtraHipSolver *solver, int axis, long *count);
int artra_hip_sample_faces(ArtraHipSolver *solver, int axis, float *features, float *labels);

/* Use the indicator for the WENO5 sweeps: faces whose logit exceeds
 * logit_threshold get WENO5, the others the linear flux.  params NULL
 * restores plain WENO5.  The indicator runs inside the fused sweeps, so
 * every swept axis must qualify for them (see artra_hip_set_fused). */
int artra_hip_set_indicator(ArtraHipSolver *solver, const float *params, float logit_threshold);
/* When and where the indicator runs.  It classifies the faces at the first
 * RK stage of each directional operator and keeps that set of troubled
 * faces for its later stages (the solution moves less than a cell within
 * one operator); every_stage must be 0 (per-stage classification is not
 * provided: it would hold the network's registers through the flux
 * passes).  Faces whose density and pressure stencils vary by at most
 * smooth_variation (relative to the face average; default 1e-3, 0 turns
 * the gate off) take the linear flux without the network, and a wavefront
 * of such faces skips the network altogether. */
int artra_hip_set_indicator_options(ArtraHipSolver *solver, int every_stage, float smooth_variation);
/* Face tallies per classification (once per directional operator):
 * counts[0] faces computed with the linear flux, counts[1] faces
 * classified, counts[2] faces computed with the full WENO5 flux
 * (counts[0] + counts[2] = counts[1]; the routing is per wavefront of 64
 * faces, which takes WENO5 when any of its faces is troubled), counts[3]
 * faces the indicator flagged troubled (at most counts[2]). */
int artra_hip_hybrid_counts(ArtraHipSolver *solver, unsigned long long counts[ARTRA_HYBRID_TALLIES], int reset);

/* Train in place on the device (Adam, weighted logistic loss of
 * label > tau); loss_history[epoch] receives the last batch loss per epoch
 * (may be NULL).  Deterministic for identical inputs. */
int artra_mlp_train(float *params, const float *features, const float *labels, long n,
                    const ArtraMlpTraining *options, float *loss_history);
/* Gradient of the mean weighted loss over n samples (device kernels). */
int artra_mlp_gradient(const float *params, const float *features, const float *labels, long n, float tau,
                       float wpos, float *grad, float *loss);
/* Logits of n feature rows (device kernels). */
int artra_mlp_logits(const float *params, const float *features, long n, float *logits);

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
/* current time and step count of the case's clock */
int artra_case_clock(const ArtraCase *c, double *now, int *step);
/* Surface force of every body as the last integration left it:
 * forces[n][9] = Fp[3], Fv[3], Tt[3]; returns the number of bodies (and
 * writes nothing when forces is NULL). */
int artra_case_forces(const ArtraCase *c, Real *forces);
/* Motion state of every body: state[n][15] = O[3], V[3], W[3], at[3],
 * ar[3] (velocities and accelerations of level TO); returns the number of
 * bodies (and writes nothing when state is NULL). */
int artra_case_bodies(const ArtraCase *c, Real *state);
/* Evaluate an initial-condition expression (t = 0); returns 0 on success. */
int artra_calc_expression(const char *text, Real x, Real y, Real z, Real *ans);

#ifdef __cplusplus
}
#endif

#endif /* ARTRA_HIP_H_ */
