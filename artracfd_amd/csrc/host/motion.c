/*
 * motion.c - moving bodies around the resident fluid solver.
 *
 * EvolveSolidDynamics (solid_dynamics.c:37-48) for a field that lives on the
 * device.  The surface force is the device reduction of output.c; the
 * rigid-body kinematics, the collisions and the motion of the bodies run on
 * the host over the Geometry shared with the caller.  The node
 * classification runs on the device (its own copy of the flags, where the
 * moving bodies were or are now) and comes back to the host's flags for the
 * host's consumers (collisions, tables, outputs).  What the tables need goes
 * back to the device: the ghost stencils of the immersed-boundary fill, the
 * surface-force nodes, and the reconstruction of the nodes a body has
 * uncovered, which the device evaluates on level TO before the ghost fill.
 * The field never crosses PCIe.
 *
 *   artra_solid_step          solid_dynamics.c:37-48
 *   Kinematics, SolveInertia  solid_dynamics.c:172-212, linear_system.c:24-174
 *   Collisions                solid_dynamics.c:213-370
 *   Motion                    solid_dynamics.c:371-404
 *   ClassifyOnDevice          immersed_boundary.c:65-264 on the device
 *   artra_hip_refresh_geometry  the device tables of a new classification
 *   artra_moving_step         solve.c:65-89 (one step of the time loop)
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "artra_device.h"
#include "artra_hip.h"
#include "artra_host.h"
#include "solver.h"

/*
 * b <- A^-1 b for the 3x3 inertia tensor by LU factorization with implicit
 * row scaling and partial pivoting, then forward and back substitution
 * (linear_system.c:24-174, the Crout/Doolittle scheme of Numerical
 * Recipes).  The operation order is the reference's so that the angular
 * acceleration is the same to the last bit.  A zero row is an error; a zero
 * pivot is replaced by 1e-15 as there.
 */
static int SolveInertia(const Real I[DIMS][DIMS], Real b[DIMS])
{
    Real A[DIMS][DIMS], scale[DIMS];
    int perm[DIMS];
    memcpy(A, I, sizeof A);
    for (int r = 0; r < DIMS; ++r) {
        Real big = 0.0;
        for (int c = 0; c < DIMS; ++c) {
            if (fabs(A[r][c]) > big) big = fabs(A[r][c]);
        }
        if (0.0 == big) return artra_set_error(1, "singular inertia tensor");
        scale[r] = 1.0 / big;
    }
    for (int p = 0; p < DIMS; ++p) {
        Real big = 0.0;
        int pivot = p;
        for (int r = p; r < DIMS; ++r) {
            const Real t = scale[r] * fabs(A[r][p]);
            if (t > big) {
                big = t;
                pivot = r;
            }
        }
        if (pivot != p) {
            for (int c = 0; c < DIMS; ++c) {
                const Real t = A[pivot][c];
                A[pivot][c] = A[p][c];
                A[p][c] = t;
            }
            scale[pivot] = scale[p];
        }
        perm[p] = pivot;
        if (0.0 == A[p][p]) A[p][p] = 1.0e-15;
        for (int r = p + 1; r < DIMS; ++r) {
            A[r][p] = A[r][p] / A[p][p];
            for (int c = p + 1; c < DIMS; ++c) A[r][c] = A[r][c] - A[r][p] * A[p][c];
        }
    }
    /* the row interchanges in factorization order, then L (unit diagonal)
     * and U */
    for (int r = 0; r < DIMS; ++r) {
        const Real t = b[r];
        b[r] = b[perm[r]];
        b[perm[r]] = t;
    }
    for (int r = 0; r < DIMS; ++r) {
        for (int c = 0; c < r; ++c) b[r] = b[r] - A[r][c] * b[c];
    }
    for (int r = DIMS - 1; r >= 0; --r) {
        for (int c = r + 1; c < DIMS; ++c) b[r] = b[r] - A[r][c] * b[c];
        b[r] = b[r] / A[r][r];
    }
    return 0;
}

/* Accelerations from the surface force, the body force and the powered
 * acceleration until time `to`, then the velocities: V[TN] the mean over the
 * step, V[TO] the end value (solid_dynamics.c:172-212). */
static int Kinematics(Space *space, Real now, Real dt)
{
    Geometry *geo = &space->geo;
    for (int n = 0; n < geo->totN; ++n) {
        Polyhedron *poly = geo->poly + n;
        if (1 == poly->state) continue;
        if (now > poly->to) {
            memset(poly->at[TN], 0, sizeof poly->at[TN]);
            memset(poly->ar[TN], 0, sizeof poly->ar[TN]);
            poly->to = FLT_MAX;
        }
        Real alpha[DIMS];
        for (int s = 0; s < DIMS; ++s) alpha[s] = poly->Tt[s] / poly->rho;
        if (SolveInertia((const Real (*)[DIMS])poly->I, alpha)) return 1;
        for (int s = 0; s < DIMS; ++s) {
            poly->at[TO][s] = (poly->Fp[s] + poly->Fv[s]) / (poly->rho * poly->volume) + poly->at[TN][s] + poly->g[s];
            poly->ar[TO][s] = alpha[s] + poly->ar[TN][s];
        }
        for (int s = 0; s < DIMS; ++s) {
            poly->V[TN][s] = 0.5 * (poly->V[TO][s] + poly->V[TO][s] + poly->at[TO][s] * dt);
            poly->W[TN][s] = 0.5 * (poly->W[TO][s] + poly->W[TO][s] + poly->ar[TO][s] * dt);
            poly->V[TO][s] = poly->V[TO][s] + poly->at[TO][s] * dt;
            poly->W[TO][s] = poly->W[TO][s] + poly->ar[TO][s] * dt;
        }
    }
    return 0;
}

/* Node-space search box of a body within the interior node range
 * (solid_dynamics.c:249-253). */
static void SearchBox(const Partition *part, const Polyhedron *poly, int box[DIMS][LIMIT])
{
    for (int s = 0; s < DIMS; ++s) {
        const int lo = part->ns[PIN][s][MIN], hi = part->ns[PIN][s][MAX];
        box[s][MIN] = artra_confine(artra_map_node(poly->box[s][MIN], part->domain[s][MIN], part->dd[s], part->ng[s]),
                                    lo, hi);
        box[s][MAX] = artra_confine(artra_map_node(poly->box[s][MAX], part->domain[s][MIN], part->dd[s], part->ng[s]),
                                    lo, hi) + 1;
    }
}

/* One contact partner of the body being examined: the partner's did, and
 * the line of impact as the sum of the path offsets that reach it
 * (solid_dynamics.c:342-370). */
static void AddContact(Geometry *geo, int did, const int offset[DIMS])
{
    for (int n = 0; n < geo->colN; ++n) {
        Collision *col = geo->col + n;
        if (did == col->gid) {
            for (int s = 0; s < DIMS; ++s) col->N[s] = col->N[s] + offset[s];
            return;
        }
    }
    Collision *col = geo->col + geo->colN;
    col->gid = did;
    for (int s = 0; s < DIMS; ++s) col->N[s] = offset[s];
    ++geo->colN;
}

/*
 * Pairwise collisions (solid_dynamics.c:213-341): a first-layer interfacial
 * node of body p whose face neighbours belong to another body marks a
 * contact; a contact that approaches along the summed line of impact gets
 * the restitution and sliding-friction impulse of the pair.  A body that
 * collides keeps its post-collision velocity as V[TN] too.
 */
static int Collisions(Space *space)
{
    const Partition *part = &space->part;
    const Node *node = space->node;
    Geometry *geo = &space->geo;
    const int tag = INT_MAX / 2;       /* added to the state of a colliding body */
    static const Real restitution[5] = {0.0, 0.25, 0.5, 0.75, 1.0};
    for (int p = 0; p < geo->totN; ++p) {
        Polyhedron *pp = geo->poly + p;
        if (1 == pp->state) continue;
        geo->colN = 0;
        int box[DIMS][LIMIT];
        SearchBox(part, pp, box);
        for (int k = box[Z][MIN]; k < box[Z][MAX]; ++k)
            for (int j = box[Y][MIN]; j < box[Y][MAX]; ++j)
                for (int i = box[X][MIN]; i < box[X][MAX]; ++i) {
                    const Node *nd = node + artra_index(k, j, i, part->n[Y], part->n[X]);
                    if (1 != nd->lid || p + 1 != nd->did) continue;
                    for (int e = 0; e < part->pathSep[1]; ++e) {
                        const int kh = k + part->path[e][Z], jh = j + part->path[e][Y], ih = i + part->path[e][X];
                        if (!artra_in_box(kh, jh, ih, part->ns[PIN])) continue;
                        const int other = node[artra_index(kh, jh, ih, part->n[Y], part->n[X])].did;
                        if (0 == other || p + 1 == other) continue;
                        AddContact(geo, other, part->path[e]);
                    }
                }
        if (0 == geo->colN) continue;
        Real Vo[DIMS], Wo[DIMS];
        memcpy(Vo, pp->V[TO], sizeof Vo);
        memcpy(Wo, pp->W[TO], sizeof Wo);
        memcpy(pp->V[TO], pp->V[TN], sizeof Vo);
        memcpy(pp->W[TO], pp->W[TN], sizeof Wo);
        const Real mp = pp->rho * pp->volume;
        for (int n = 0; n < geo->colN; ++n) {
            const Collision *col = geo->col + n;
            const int last = (geo->colN - 1 == n);
            if (0 == abs(col->N[X]) + abs(col->N[Y]) + abs(col->N[Z])) {
                if (last && tag > pp->state) {
                    memcpy(pp->V[TO], Vo, sizeof Vo);
                    memcpy(pp->W[TO], Wo, sizeof Wo);
                }
                continue;
            }
            Real N[DIMS] = {col->N[X], col->N[Y], col->N[Z]};
            artra_normalize(DIMS, artra_norm(N), N);
            const Polyhedron *pn = geo->poly + col->gid - 1;
            if (pp->mid < 0 || pp->mid > 4 || pn->mid < 0 || pn->mid > 4) {
                return artra_set_error(1, "collision of bodies %d and %d: material id outside 0..4", p, col->gid - 1);
            }
            Real V[DIMS], W[DIMS];
            for (int s = 0; s < DIMS; ++s) {
                V[s] = pp->V[TN][s] - pn->V[TN][s];
                W[s] = pp->W[TN][s] - pn->W[TN][s];
            }
            const Real Vn = artra_dot(V, N);
            if (0.0 >= Vn) {
                if (last && tag > pp->state) {
                    memcpy(pp->V[TO], Vo, sizeof Vo);
                    memcpy(pp->W[TO], Wo, sizeof Wo);
                }
                continue;
            }
            if (tag > pp->state) pp->state = pp->state + tag;
            const Real mn = pn->rho * pn->volume;
            const Real meff = mn / (mp + mn);
            const Real cr = 0.5 * (restitution[pp->mid] + restitution[pn->mid]);
            const Real cf = 0.5 * (pp->cf + pn->cf);
            for (int s = 0; s < DIMS; ++s) {
                pp->V[TO][s] = pp->V[TO][s] - meff * (1.0 + cr) * Vn * N[s] - cf * (V[s] - Vn * N[s]);
                pp->W[TO][s] = pp->W[TO][s] - meff * W[s];
            }
        }
    }
    for (int p = 0; p < geo->totN; ++p) {
        Polyhedron *pp = geo->poly + p;
        if (tag > pp->state) continue;
        pp->state = pp->state - tag;
        memcpy(pp->V[TN], pp->V[TO], sizeof pp->V[TN]);
        memcpy(pp->W[TN], pp->W[TO], sizeof pp->W[TN]);
    }
    return 0;
}

/* Positions over the step with the mean velocities: a sphere moves its
 * centre and box, a triangulated body is rotated and translated about its
 * centroid (solid_dynamics.c:371-404). */
static void Motion(Space *space, Real dt)
{
    Geometry *geo = &space->geo;
    static const Real unit[DIMS] = {1.0, 1.0, 1.0};
    for (int n = 0; n < geo->totN; ++n) {
        Polyhedron *poly = geo->poly + n;
        if (1 == poly->state) continue;
        Real offset[DIMS], angle[DIMS];
        for (int s = 0; s < DIMS; ++s) {
            offset[s] = poly->V[TN][s] * dt;
            angle[s] = poly->W[TN][s] * dt;
        }
        if (0 >= poly->faceN) {
            for (int s = 0; s < DIMS; ++s) poly->O[s] = poly->O[s] + offset[s];
            for (int s = 0; s < DIMS; ++s) {
                poly->box[s][MIN] = poly->O[s] - poly->r;
                poly->box[s][MAX] = poly->O[s] + poly->r;
            }
        } else {
            artra_transform_polyhedron(poly->O, unit, angle, offset, poly);
        }
    }
}

int artra_geo_device_init(ArtraHipSolver *s)
{
    const Space *space = s->space;
    const Partition *part = &space->part;
    ArtraGeoSetup gs;
    memset(&gs, 0, sizeof gs);
    for (int d = 0; d < DIMS; ++d) {
        gs.smin[d] = part->domain[d][MIN];
        gs.d[d] = part->d[d];
        gs.ng[d] = part->ng[d];
    }
    memcpy(gs.path, part->path, sizeof gs.path);
    memcpy(gs.pathSep, part->pathSep, sizeof gs.pathSep);
    gs.gl = part->gl;
    int *flags = malloc((size_t)4 * s->N * sizeof(int));
    if (NULL == flags) return artra_set_error(1, "out of host memory for %ld node flags", s->N);
    for (long q = 0; q < s->N; ++q) {
        const Node *nd = space->node + q;
        flags[q] = nd->did;
        flags[s->N + q] = nd->fid;
        flags[2 * s->N + q] = nd->lid;
        flags[3 * s->N + q] = nd->gst;
    }
    const int rc = artra_dev_geo_init(s->dev, &gs, flags, space->geo.totN);
    free(flags);
    return rc ? artra_dev_failed() : 0;
}

/* The bodies as the device's classification and ghost table take them. */
static ArtraGeoBody *GeoBodies(const Space *space)
{
    const Geometry *geo = &space->geo;
    ArtraGeoBody *bodies = calloc(geo->totN > 0 ? geo->totN : 1, sizeof *bodies);
    if (NULL == bodies) return NULL;
    for (int n = 0; n < geo->totN; ++n) {
        const Polyhedron *poly = geo->poly + n;
        ArtraGeoBody *b = bodies + n;
        b->moving = (1 != poly->state);
        SearchBox(&space->part, poly, b->box);
        for (int d = 0; d < DIMS; ++d) {
            b->O[d] = poly->O[d];
            b->V[d] = poly->V[TO][d];
            b->W[d] = poly->W[TO][d];
        }
        b->r = poly->r;
        b->cf = poly->cf;
        b->T = poly->T;
        if (0 < poly->faceN) b->pv = artra_poly_view(poly);
    }
    return bodies;
}

/* ComputeGeometricField over region on the device, the host's flags
 * updated from it and the uncovered nodes listed in joins. */
static int ClassifyOnDevice(ArtraHipSolver *s, const int region[DIMS][LIMIT], const ArtraGeoBody *bodies,
                            ArtraGhostTable *joins)
{
    Space *space = s->space;
    const int *packed = NULL;
    long count = 0;
    if (artra_dev_geo_classify(s->dev, bodies, space->geo.totN, region, &packed, &count)) return artra_dev_failed();
    return artra_apply_classification(space, region, packed, count, joins);
}

/* The surface-force table built on the device from its flags
 * (artra_build_force_table's nodes, order and geometry); the host keeps
 * each body's counts for the calibration. */
static int ForceTableOnDevice(ArtraHipSolver *s, const ArtraGeoBody *bodies)
{
    const Space *space = s->space;
    const Geometry *geo = &space->geo;
    free(s->force);
    s->force = NULL;
    s->forceBodies = 0;
    int nb = 0;
    for (int n = 0; n < geo->totN; ++n) nb += (0 >= geo->poly[n].state);
    if (0 == nb) return artra_dev_set_force_table(s->dev, NULL, 0, NULL, 0) ? artra_dev_failed() : 0;
    s->force = calloc((size_t)nb, sizeof *s->force);
    int *ints = malloc((size_t)4 * nb * sizeof(int));
    int (*box)[DIMS][LIMIT] = malloc((size_t)nb * sizeof *box);
    if (!s->force || !ints || !box) {
        free(ints);
        free(box);
        return artra_set_error(1, "out of host memory for the surface-force table");
    }
    int *slot = ints, *viscous = ints + nb, *gstN = ints + 2 * nb, *lidN = ints + 3 * nb;
    int b = 0;
    for (int n = 0; n < geo->totN; ++n) {
        const Polyhedron *poly = geo->poly + n;
        if (0 < poly->state) continue;
        slot[b] = n;
        viscous[b] = (0.0 < s->model->refMu) && (0.0 < poly->cf);
        artra_force_body_box(&space->part, poly, box[b]);
        ++b;
    }
    int rc = artra_dev_build_force_table(s->dev, bodies, geo->totN, slot, (const int (*)[DIMS][LIMIT])box, viscous,
                                         nb, gstN, lidN) ? artra_dev_failed() : 0;
    if (!rc) {
        for (b = 0; b < nb; ++b) {
            s->force[b].body = slot[b];
            s->force[b].gstN = gstN[b];
            s->force[b].lidN = lidN[b];
        }
        s->forceBodies = nb;
    }
    free(ints);
    free(box);
    return rc;
}

/* The device tables after a classification on the device: the ghost table
 * built there (on the host when a stencil search reaches beyond radius 3),
 * the surface-force table likewise from the device's flags.  The fluid flags
 * are already current on the device. */
static int TablesOnDevice(ArtraHipSolver *s, const ArtraGeoBody *bodies)
{
    const Space *space = s->space;
    const Partition *part = &space->part;
    double dd[DIMS] = {part->dd[X], part->dd[Y], part->dd[Z]};
    int fallback = 0;
    if (artra_dev_build_ghosts(s->dev, bodies, space->geo.totN, dd, part->tinyL, s->model->ibmLayer, &fallback)) {
        return artra_dev_failed();
    }
    if (fallback) {
        artra_free_ghost_table(&s->ghosts);
        if (artra_build_ghost_table(space, s->model, &s->ghosts)) return 1;
        if (artra_dev_set_ghosts(s->dev, &s->ghosts)) return artra_dev_failed();
    }
    return ForceTableOnDevice(s, bodies);
}

int artra_hip_refresh_geometry(ArtraHipSolver *s, const int (*region)[LIMIT])
{
    const Space *space = s->space;
    const Partition *part = &space->part;
    int box[DIMS][LIMIT];                   /* the changed flags' node box */
    if (NULL == region) {
        for (long q = 0; q < s->N; ++q) s->fluid[q] = (0 == space->node[q].did);
    } else {
        for (int d = 0; d < DIMS; ++d) {
            box[d][MIN] = (region[d][MIN] > part->ns[PIN][d][MIN]) ? region[d][MIN] : part->ns[PIN][d][MIN];
            box[d][MAX] = (region[d][MAX] < part->ns[PIN][d][MAX]) ? region[d][MAX] : part->ns[PIN][d][MAX];
        }
        for (int k = box[Z][MIN]; k < box[Z][MAX]; ++k)
            for (int j = box[Y][MIN]; j < box[Y][MAX]; ++j)
                for (int i = box[X][MIN]; i < box[X][MAX]; ++i) {
                    const int q = artra_index(k, j, i, part->n[Y], part->n[X]);
                    s->fluid[q] = (0 == space->node[q].did);
                }
    }
    if (artra_dev_set_fluid(s->dev, s->fluid, region ? (const int (*)[LIMIT])box : NULL)) return artra_dev_failed();
    artra_free_ghost_table(&s->ghosts);
    if (artra_build_ghost_table(space, s->model, &s->ghosts)) return 1;
    if (artra_dev_set_ghosts(s->dev, &s->ghosts)) return artra_dev_failed();
    return artra_build_force_table(s);
}

int artra_solid_step(ArtraHipSolver *s, Real now, Real dt)
{
    Space *space = s->space;
    const Model *model = s->model;
    if (artra_hip_surface_force(s)) return 1;
    if (Kinematics(space, now, dt)) return 1;
    if (1 != model->psi && Collisions(space)) return 1;
    Motion(space, dt);
    /* the classification only where the bodies were or are now */
    int region[DIMS][LIMIT];
    artra_scope_region(space, &s->scope, region);
    ArtraGhostTable joins;
    memset(&joins, 0, sizeof joins);
    ArtraGeoBody *bodies = GeoBodies(space);
    int rc = bodies ? 0 : artra_set_error(1, "out of host memory");
    if (!rc) rc = ClassifyOnDevice(s, (const int (*)[LIMIT])region, bodies, &joins);
    if (!rc) artra_scope_update(space, &s->scope, &joins, (const int (*)[LIMIT])region);
    if (!rc) rc = TablesOnDevice(s, bodies);
    free(bodies);
    if (!rc && artra_dev_reconstruct(s->dev, TO, &joins)) rc = artra_dev_failed();
    if (!rc && artra_dev_treat_immersed(s->dev, TO)) rc = artra_dev_failed();
    artra_free_ghost_table(&joins);
    return rc;
}

int artra_moving_step(ArtraHipSolver *s, Time *time, Real tmInt, Real *rcInt, Real *dt)
{
    ++time->stepC;
    Real h = 0.0;
    /* the CFL step folds in the body speeds of this instant */
    if (artra_hip_compute_dt(s, time, &h)) return 1;
    if (*rcInt + h > tmInt) {
        h = tmInt - *rcInt;
        *rcInt = 0.0;
    } else {
        *rcInt = *rcInt + h;
    }
    time->now = time->now + h;
    if (time->now > time->end) {
        h = time->end - (time->now - h);
        time->now = time->end;
    }
    if (artra_solid_step(s, time->now, 0.5 * h)) return 1;
    if (artra_hip_evolve(s, h)) return 1;
    if (artra_solid_step(s, time->now, 0.5 * h)) return 1;
    *dt = h;
    return 0;
}
