/*
 * tables.c - static tables consumed by the device boundary kernels.
 *
 * External boundary: TreatBoundary (boundary_treatment.c:28-65) visits, for
 * r = 0..R and faces W E S N F B in that order, a box that extends r layers
 * into the ghost region tangentially; later boxes read values that earlier
 * boxes have just written at edges and corners.  artra_bc_boxes lists the
 * boxes in that order; artra_bc_groups packs consecutive boxes whose node
 * sets (written and read) do not interact, so one kernel launch can serve a
 * whole group with results identical to the sequential order.
 *
 * Immersed boundary: while the node flags stay as they are, every ghost
 * node's interpolation stencil is fixed.  artra_build_ghost_table replays the
 * neighbour search of InverseDistanceWeighting (immersed_boundary.c:413-473)
 * once per node classification (at attach, and after every move of the
 * bodies), recording the stencil nodes and inverse-distance weights in
 * summation order, and the boundary-point data of ReconstructFlow
 * (immersed_boundary.c:357-412).  artra_add_join records, in the same form,
 * the reconstruction of a node that a moving body has just uncovered
 * (immersed_boundary.c:193-204).
 */
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#include "artra_host.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* --------------------------------------------------- external boundary */
int artra_bc_boxes(const Space *space, ArtraBcBox *out, int cap)
{
    const Partition *part = &space->part;
    const int R = (part->ng[X] > part->ng[Y]) ? ((part->ng[X] > part->ng[Z]) ? part->ng[X] : part->ng[Z])
                                              : ((part->ng[Y] > part->ng[Z]) ? part->ng[Y] : part->ng[Z]);
    int count = 0;
    for (int r = 0; r <= R; ++r) {
        for (int p = PWB; p <= PBB; ++p) {
            /* the physical layer of a periodic face is computed, not treated */
            if (0 == r && PERIODIC == part->typeBC[p]) continue;
            ArtraBcBox b;
            b.face = p;
            b.layer = r;
            int empty = 0;
            for (int s = 0; s < DIMS; ++s) {
                const int N = part->N[p][s];
                const int reach = (r < part->ng[s]) ? r : part->ng[s];
                b.box[s][MIN] = part->ns[p][s][MIN] + reach * (N - !N);
                b.box[s][MAX] = part->ns[p][s][MAX] + reach * (N + !N) - (part->ng[s] < r) * (!!N);
                if (b.box[s][MIN] >= b.box[s][MAX]) empty = 1;
            }
            if (empty) continue;
            if (count >= cap) return -1;
            out[count++] = b;
        }
    }
    return count;
}

typedef struct {
    int lo[DIMS];
    int hi[DIMS];
} Region;

static int Overlap(const Region *a, const Region *b)
{
    for (int s = 0; s < DIMS; ++s) {
        if (a->hi[s] <= b->lo[s] || b->hi[s] <= a->lo[s]) return 0;
    }
    return 1;
}

/* Nodes a box writes and the (up to two) shifted regions it reads. */
static int BoxRegions(const Partition *part, const ArtraBcBox *b, Region *write, Region read[2])
{
    const int p = b->face, r = b->layer;
    int shift[2][DIMS] = {{0}};
    int nread = 0;
    const int type = part->typeBC[p];
    if (r > 0) {
        if (SLIPWALL == type || NOSLIPWALL == type) {
            for (int s = 0; s < DIMS; ++s) {
                shift[0][s] = -r * part->N[p][s];
                shift[1][s] = -2 * r * part->N[p][s];
            }
            nread = 2;
        } else if (PERIODIC == type) {
            for (int s = 0; s < DIMS; ++s) shift[0][s] = -part->m[s] * part->N[p][s];
            nread = 1;
        } else {
            for (int s = 0; s < DIMS; ++s) shift[0][s] = -part->N[p][s];
            nread = 1;
        }
    } else if (INFLOW != type) {
        for (int s = 0; s < DIMS; ++s) shift[0][s] = -part->N[p][s];
        nread = 1;
    }
    for (int s = 0; s < DIMS; ++s) {
        write->lo[s] = b->box[s][MIN];
        write->hi[s] = b->box[s][MAX];
        for (int q = 0; q < nread; ++q) {
            read[q].lo[s] = b->box[s][MIN] + shift[q][s];
            read[q].hi[s] = b->box[s][MAX] + shift[q][s];
        }
    }
    return nread;
}

int artra_bc_groups(const Space *space, const ArtraBcBox *boxes, int nbox, int *group_start, int cap)
{
    const Partition *part = &space->part;
    int ngroup = 0;
    int first = 0;
    while (first < nbox) {
        if (ngroup >= cap) return -1;
        group_start[ngroup++] = first;
        int last = first + 1;
        while (last < nbox) {
            Region wNew, rNew[2];
            const int nrNew = BoxRegions(part, boxes + last, &wNew, rNew);
            int clash = 0;
            for (int q = first; q < last && !clash; ++q) {
                Region wOld, rOld[2];
                const int nrOld = BoxRegions(part, boxes + q, &wOld, rOld);
                clash = Overlap(&wNew, &wOld);
                for (int t = 0; t < nrNew && !clash; ++t) clash = Overlap(rNew + t, &wOld);
                for (int t = 0; t < nrOld && !clash; ++t) clash = Overlap(rOld + t, &wNew);
            }
            if (clash) break;
            ++last;
        }
        first = last;
    }
    group_start[ngroup] = nbox;
    return ngroup;
}

/* --------------------------------------------------- immersed boundary */
typedef struct {
    int *node;
    Real *w;
    int n;
    int cap;
} StencilBuf;

static int Push(StencilBuf *sb, int node, Real w)
{
    if (sb->n == sb->cap) {
        const int cap = sb->cap ? 2 * sb->cap : 4096;
        int *nn = realloc(sb->node, cap * sizeof(int));
        Real *nw = realloc(sb->w, cap * sizeof(Real));
        if (!nn || !nw) {
            free(nn ? nn : sb->node);
            free(nw ? nw : sb->w);
            sb->node = NULL;
            sb->w = NULL;
            return artra_set_error(1, "out of host memory for ghost stencils");
        }
        sb->node = nn;
        sb->w = nw;
        sb->cap = cap;
    }
    sb->node[sb->n] = node;
    sb->w[sb->n] = w;
    sb->n++;
    return 0;
}

static Real InverseDistance(Real tiny, Real d2)
{
    if (tiny > d2) d2 = tiny;
    return 1.0 / d2;
}

/* Target set of the stencil search besides did (immersed_boundary.c:446-462):
 * every node of the domain (TYPED), a ghost layer (layer >= 1), or the nodes
 * of the domain that were in it before the bodies moved and have not been
 * reconstructed since (TYPEF: gst, which still holds the previous domain
 * field, equal to did, and fid not the reconstruction mark). */
enum { kTypeDomain = 0, kTypeOriginal = -2 };

/*
 * Stencil search around node c for point p: the box of half-width h grows
 * until it contains at least one node of the target set, scanned k-j-i
 * within the interior range.  Appends the entries and returns the running
 * weight sum in *wsum.
 */
static int SearchStencil(const Space *space, const int c[DIMS], const Real p[DIMS], int h, int did, int layer,
                         StencilBuf *sb, int *count, Real *wsum)
{
    const Partition *part = &space->part;
    const Node *node = space->node;
    const int maxR = part->n[X] + part->n[Y] + part->n[Z];
    *count = 0;
    *wsum = 0.0;
    for (int r = h; 0 == *count; ++r) {
        if (r > maxR) return artra_set_error(1, "immersed boundary: no stencil node found");
        for (int kh = -r; kh <= r; ++kh)
            for (int jh = -r; jh <= r; ++jh)
                for (int ih = -r; ih <= r; ++ih) {
                    const int nh[DIMS] = {c[X] + ih, c[Y] + jh, c[Z] + kh};
                    if (!artra_in_box(nh[Z], nh[Y], nh[X], part->ns[PIN])) continue;
                    const int q = artra_index(nh[Z], nh[Y], nh[X], part->n[Y], part->n[X]);
                    if (did != node[q].did) continue;
                    if (layer > 0 && layer != node[q].gst) continue;
                    if (kTypeOriginal == layer && (did != node[q].gst || 0 > node[q].fid)) continue;
                    const Real ph[DIMS] = {artra_map_point(nh[X], part->domain[X][MIN], part->d[X], part->ng[X]),
                                           artra_map_point(nh[Y], part->domain[Y][MIN], part->d[Y], part->ng[Y]),
                                           artra_map_point(nh[Z], part->domain[Z][MIN], part->d[Z], part->ng[Z])};
                    const Real w = InverseDistance(part->tinyL, artra_dist2(p, ph));
                    if (Push(sb, q, w)) return 1;
                    *wsum = *wsum + w;
                    ++*count;
                }
    }
    return 0;
}

/* A ghost node found by the scan, in table order. */
typedef struct {
    int node, k, j, i, body;
} GhostSite;

/* One ghost's entry and stencil (appended to sb; gh->first is the offset
 * in sb). */
static int BuildGhost(const Space *space, const Model *model, const GhostSite *site, int r, StencilBuf *sb,
                      ArtraGhost *gh)
{
    const Partition *part = &space->part;
    const Polyhedron *poly = space->geo.poly + site->body;
    const Node *nd = space->node + site->node;
    memset(gh, 0, sizeof *gh);
    gh->node = site->node;
    gh->first = sb->n;
    const Real pG[DIMS] = {artra_map_point(site->i, part->domain[X][MIN], part->d[X], part->ng[X]),
                           artra_map_point(site->j, part->domain[Y][MIN], part->d[Y], part->ng[Y]),
                           artra_map_point(site->k, part->domain[Z][MIN], part->d[Z], part->ng[Z])};
    Real wsum = 0.0;
    if (model->ibmLayer >= r) {
        Real pO[DIMS], pI[DIMS], N[DIMS];
        artra_compute_geometric_data(pG, nd->fid, poly, pO, pI, N);
        const int nI[DIMS] = {artra_map_node(pI[X], part->domain[X][MIN], part->dd[X], part->ng[X]),
                              artra_map_node(pI[Y], part->domain[Y][MIN], part->dd[Y], part->ng[Y]),
                              artra_map_node(pI[Z], part->domain[Z][MIN], part->dd[Z], part->ng[Z])};
        if (SearchStencil(space, nI, pI, 2, 0, kTypeDomain, sb, &gh->count, &wsum)) return 1;
        gh->kind = 0;
        gh->wO = InverseDistance(part->tinyL, artra_dist2(pI, pO));
        /* body surface velocity V + W x (pO - O) */
        const Real arm[DIMS] = {pO[X] - poly->O[X], pO[Y] - poly->O[Y], pO[Z] - poly->O[Z]};
        artra_cross(poly->W[TO], arm, gh->Vs);
        for (int s = 0; s < DIMS; ++s) gh->Vs[s] = poly->V[TO][s] + gh->Vs[s];
        for (int s = 0; s < DIMS; ++s) gh->N[s] = N[s];
        artra_orthogonal_space(N, gh->Ta, gh->Tb);
        gh->noslip = (0.0 < poly->cf);
        gh->isothermal = !(0.0 > poly->T);
        gh->wallT = poly->T;
    } else {
        const int nG[DIMS] = {site->i, site->j, site->k};
        if (SearchStencil(space, nG, pG, 1, site->body + 1, r - 1, sb, &gh->count, &wsum)) return 1;
        gh->kind = 1;
    }
    return 0;
}

/*
 * The ghost table: ghosts by layer, then body, then k-j-i within the body's
 * box (the order of immersed_boundary.c:301-337 up to the body/layer loop
 * order, which the fill does not depend on: a layer reads fluid nodes or
 * the previous layer of the same body only).  The scan is serial; the
 * entries and stencils, independent of each other, are built in parallel
 * into per-thread buffers and joined in table order.
 */
int artra_build_ghost_table(const Space *space, const Model *model, ArtraGhostTable *table)
{
    const Partition *part = &space->part;
    const Node *node = space->node;
    const Geometry *geo = &space->geo;
    memset(table, 0, sizeof *table);
    table->layers = part->gl;
    table->layer_start = calloc(part->gl + 1, sizeof(int));
    if (NULL == table->layer_start) return artra_set_error(1, "out of host memory");
    int cap = 0, count = 0;
    GhostSite *sites = NULL;
    int *layer = NULL;
    for (int r = 1; r <= part->gl; ++r) {
        table->layer_start[r - 1] = count;
        for (int n = 0; n < geo->totN; ++n) {
            const Polyhedron *poly = geo->poly + n;
            int box[DIMS][LIMIT];
            for (int s = 0; s < DIMS; ++s) {
                const int lo = part->ns[PIN][s][MIN], hi = part->ns[PIN][s][MAX];
                box[s][MIN] = artra_confine(artra_map_node(poly->box[s][MIN], part->domain[s][MIN], part->dd[s],
                                                           part->ng[s]), lo, hi);
                box[s][MAX] = artra_confine(artra_map_node(poly->box[s][MAX], part->domain[s][MIN], part->dd[s],
                                                           part->ng[s]), lo, hi) + 1;
            }
            for (int k = box[Z][MIN]; k < box[Z][MAX]; ++k)
                for (int j = box[Y][MIN]; j < box[Y][MAX]; ++j)
                    for (int i = box[X][MIN]; i < box[X][MAX]; ++i) {
                        const int idx = artra_index(k, j, i, part->n[Y], part->n[X]);
                        if (r != node[idx].gst || n + 1 != node[idx].did) continue;
                        if (count == cap) {
                            cap = cap ? 2 * cap : 1024;
                            GhostSite *gs = realloc(sites, cap * sizeof *sites);
                            int *gl = realloc(layer, cap * sizeof *layer);
                            if (gs) sites = gs;
                            if (gl) layer = gl;
                            if (!gs || !gl) {
                                free(sites);
                                free(layer);
                                return artra_set_error(1, "out of host memory for ghost table");
                            }
                        }
                        sites[count] = (GhostSite){idx, k, j, i, n};
                        layer[count] = r;
                        ++count;
                    }
        }
    }
    table->layer_start[part->gl] = count;
    table->ghosts = calloc(count > 0 ? count : 1, sizeof(ArtraGhost));
    int threads = 1;
#ifdef _OPENMP
    threads = omp_get_max_threads();
#endif
    StencilBuf *tb = calloc(threads, sizeof *tb);
    int *owner = malloc((count > 0 ? count : 1) * sizeof *owner);
    int failed = (NULL == table->ghosts || NULL == tb || NULL == owner);
    if (!failed) {
#pragma omp parallel if (count > 64)
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
#pragma omp for schedule(dynamic, 16)
            for (int q = 0; q < count; ++q) {
                int stop;
#pragma omp atomic read
                stop = failed;
                if (stop) continue;
                owner[q] = tid;
                if (BuildGhost(space, model, sites + q, layer[q], tb + tid, table->ghosts + q)) {
#pragma omp atomic write
                    failed = 1;
                }
            }
        }
    }
    /* join the stencils in table order: offsets by a running sum, then the
     * segments copied in parallel */
    StencilBuf sb = {0};
    int *from = malloc((count > 0 ? count : 1) * sizeof *from);
    failed = failed || NULL == from;
    long total = 0;
    for (int q = 0; q < count && !failed; ++q) {
        ArtraGhost *gh = table->ghosts + q;
        from[q] = gh->first;
        gh->first = (int)total;
        total += gh->count;
    }
    if (!failed && total > INT_MAX) failed = 1;
    if (!failed && total > 0) {
        sb.node = malloc((size_t)total * sizeof(int));
        sb.w = malloc((size_t)total * sizeof(Real));
        sb.n = sb.cap = (int)total;
        failed = (NULL == sb.node || NULL == sb.w);
    }
    if (!failed) {
#pragma omp parallel for schedule(static) if (count > 64)
        for (int q = 0; q < count; ++q) {
            const ArtraGhost *gh = table->ghosts + q;
            const StencilBuf *src = tb + owner[q];
            memcpy(sb.node + gh->first, src->node + from[q], (size_t)gh->count * sizeof(int));
            memcpy(sb.w + gh->first, src->w + from[q], (size_t)gh->count * sizeof(Real));
        }
    }
    free(from);
    for (int t = 0; tb && t < threads; ++t) {
        free(tb[t].node);
        free(tb[t].w);
    }
    free(tb);
    free(owner);
    free(sites);
    free(layer);
    table->nghost = count;
    table->ghost_cap = count;
    table->stencil_node = sb.node;
    table->stencil_w = sb.w;
    table->nstencil = sb.n;
    table->stencil_cap = sb.cap;
    if (failed) return artra_set_error(1, "immersed boundary: ghost stencil search failed (no stencil node or "
                                          "out of host memory)");
    return 0;
}

int artra_add_join(const Space *space, int k, int j, int i, ArtraGhostTable *joins)
{
    const Partition *part = &space->part;
    if (NULL == joins->layer_start) {
        joins->layers = 1;
        joins->layer_start = calloc(2, sizeof(int));
        if (NULL == joins->layer_start) return artra_set_error(1, "out of host memory");
    }
    if (joins->nghost == joins->ghost_cap) {
        const int cap = joins->ghost_cap ? 2 * joins->ghost_cap : 256;
        ArtraGhost *g = realloc(joins->ghosts, cap * sizeof(ArtraGhost));
        if (NULL == g) return artra_set_error(1, "out of host memory for reconstructed nodes");
        joins->ghosts = g;
        joins->ghost_cap = cap;
    }
    ArtraGhost *gh = joins->ghosts + joins->nghost;
    memset(gh, 0, sizeof *gh);
    gh->node = artra_index(k, j, i, part->n[Y], part->n[X]);
    gh->kind = 1;
    gh->first = joins->nstencil;
    StencilBuf sb = {joins->stencil_node, joins->stencil_w, joins->nstencil, joins->stencil_cap};
    const int c[DIMS] = {i, j, k};
    const Real p[DIMS] = {artra_map_point(i, part->domain[X][MIN], part->d[X], part->ng[X]),
                          artra_map_point(j, part->domain[Y][MIN], part->d[Y], part->ng[Y]),
                          artra_map_point(k, part->domain[Z][MIN], part->d[Z], part->ng[Z])};
    Real wsum = 0.0;
    const int rc = SearchStencil(space, c, p, 2, 0, kTypeOriginal, &sb, &gh->count, &wsum);
    joins->stencil_node = sb.node;
    joins->stencil_w = sb.w;
    joins->nstencil = sb.n;
    joins->stencil_cap = sb.cap;
    if (rc) return 1;
    joins->nghost++;
    joins->layer_start[1] = joins->nghost;
    return 0;
}

void artra_apply_joins(Space *space, const Model *model, const ArtraGhostTable *joins, int tn)
{
    Node *node = space->node;
    for (int q = 0; q < joins->nghost; ++q) {
        const ArtraGhost *gh = joins->ghosts + q;
        Real Uo[DIMUo] = {0.0}, Uoh[DIMUo];
        Real wsum = 0.0;
        for (int e = gh->first; e < gh->first + gh->count; ++e) {
            const Real w = joins->stencil_w[e];
            artra_map_primitive(model->gamma, model->gasR, node[joins->stencil_node[e]].U[tn], Uoh);
            for (int c = 0; c < DIMUo; ++c) Uo[c] = Uo[c] + Uoh[c] * w;
            wsum = wsum + w;
        }
        artra_normalize(DIMUo, wsum, Uo);
        Uo[0] = Uo[4] / (Uo[5] * model->gasR);
        artra_map_conservative(model->gamma, Uo, node[gh->node].U[tn]);
    }
}

void artra_free_ghost_table(ArtraGhostTable *table)
{
    free(table->layer_start);
    free(table->ghosts);
    free(table->stencil_node);
    free(table->stencil_w);
    memset(table, 0, sizeof *table);
}

/* ------------------------------------------------ resolved boundary fill */
/*
 * Labels of the symbolic replay: a node label is
 *   >= 0           "copy of interior node <label>" (an ns[PIN] node, which
 *                  the boundary treatment never writes),
 *   LABEL_CONST-p  "inflow state of face p",
 *   LABEL_STALE    "value left from before this fill",
 *   LABEL_MIXED    "function of several nodes" (wall reflections).
 */
enum { LABEL_STALE = -1, LABEL_MIXED = -2, LABEL_CONST = -3 };

static int LabelOf(const int *label, const Partition *part, int k, int j, int i)
{
    if (artra_in_box(k, j, i, part->ns[PIN])) return artra_index(k, j, i, part->n[Y], part->n[X]);
    return label[artra_index(k, j, i, part->n[Y], part->n[X])];
}

/* The symbolic replay of the boxes in order: label[q] as described above. */
static void ReplayLabels(const Space *space, const ArtraBcBox *boxes, int nbox, int *label)
{
    const Partition *part = &space->part;
    const long N = (long)part->n[X] * part->n[Y] * part->n[Z];
    for (long q = 0; q < N; ++q) label[q] = LABEL_STALE;
    for (int b = 0; b < nbox; ++b) {
        const ArtraBcBox *bx = boxes + b;
        const int p = bx->face, r = bx->layer, type = part->typeBC[p];
        const int *Nn = part->N[p];
        for (int k = bx->box[Z][MIN]; k < bx->box[Z][MAX]; ++k)
            for (int j = bx->box[Y][MIN]; j < bx->box[Y][MAX]; ++j)
                for (int i = bx->box[X][MIN]; i < bx->box[X][MAX]; ++i) {
                    int lab;
                    if (r > 0) {
                        if (SLIPWALL == type || NOSLIPWALL == type) lab = LABEL_MIXED;
                        else if (PERIODIC == type)
                            lab = LabelOf(label, part, k - part->m[Z] * Nn[Z], j - part->m[Y] * Nn[Y],
                                          i - part->m[X] * Nn[X]);
                        else lab = LabelOf(label, part, k - Nn[Z], j - Nn[Y], i - Nn[X]);
                    } else if (INFLOW == type) {
                        lab = LABEL_CONST - p;
                    } else if (OUTFLOW == type) {
                        lab = LabelOf(label, part, k - Nn[Z], j - Nn[Y], i - Nn[X]);
                    } else {
                        lab = LABEL_MIXED;
                    }
                    label[artra_index(k, j, i, part->n[Y], part->n[X])] = lab;
                }
    }
}

/* A node outside the interior range along axis s only. */
static int AxisOnly(const Partition *part, int q, int s)
{
    const int c[DIMS] = {q % part->n[X], (q / part->n[X]) % part->n[Y], q / (part->n[X] * part->n[Y])};
    for (int t = 0; t < DIMS; ++t) {
        if (t == s) continue;
        if (c[t] < part->ns[PIN][t][MIN] || c[t] >= part->ns[PIN][t][MAX]) return 0;
    }
    return !(c[s] >= part->ns[PIN][s][MIN] && c[s] < part->ns[PIN][s][MAX]);
}

int artra_bc_resolve_axis(const Space *space, const ArtraBcBox *boxes, int nbox, int s, ArtraBcGather *out)
{
    const Partition *part = &space->part;
    const long N = (long)part->n[X] * part->n[Y] * part->n[Z];
    memset(out, 0, sizeof *out);
    int *label = malloc(N * sizeof(int));
    if (NULL == label) return artra_set_error(1, "out of host memory");
    ReplayLabels(space, boxes, nbox, label);
    /* every written node outside the interior along s only must resolve */
    long count = 0;
    int ok = 1;
    for (int b = 0; b < nbox && ok; ++b) {
        const ArtraBcBox *bx = boxes + b;
        for (int k = bx->box[Z][MIN]; k < bx->box[Z][MAX] && ok; ++k)
            for (int j = bx->box[Y][MIN]; j < bx->box[Y][MAX] && ok; ++j)
                for (int i = bx->box[X][MIN]; i < bx->box[X][MAX]; ++i) {
                    const int q = artra_index(k, j, i, part->n[Y], part->n[X]);
                    if (AxisOnly(part, q, s) && !(label[q] >= 0 || label[q] <= LABEL_CONST)) {
                        ok = 0;
                        break;
                    }
                }
    }
    for (long q = 0; q < N && ok; ++q) count += AxisOnly(part, (int)q, s) && (label[q] >= 0 || label[q] <= LABEL_CONST);
    if (ok) {
        out->dst = malloc((count ? count : 1) * sizeof(int));
        out->src = malloc((count ? count : 1) * sizeof(int));
        if (!out->dst || !out->src) {
            free(label);
            artra_free_bc_gather(out);
            return artra_set_error(1, "out of host memory");
        }
        long n = 0;
        for (long q = 0; q < N; ++q) {
            if (AxisOnly(part, (int)q, s) && (label[q] >= 0 || label[q] <= LABEL_CONST)) {
                out->dst[n] = (int)q;
                out->src[n] = label[q];
                ++n;
            }
        }
        out->count = n;
        out->resolved = 1;
    }
    free(label);
    return 0;
}

int artra_bc_resolve(const Space *space, const ArtraBcBox *boxes, int nbox, ArtraBcGather *out)
{
    const Partition *part = &space->part;
    const long N = (long)part->n[X] * part->n[Y] * part->n[Z];
    memset(out, 0, sizeof *out);
    int *label = malloc(N * sizeof(int));
    if (NULL == label) return artra_set_error(1, "out of host memory");
    ReplayLabels(space, boxes, nbox, label);
    /* count entries; any written node that does not resolve disables the gather */
    long count = 0;
    int ok = 1;
    for (long q = 0; q < N; ++q) {
        const int lab = label[q];
        if (LABEL_MIXED == lab) ok = 0;
        if (lab >= 0 || lab <= LABEL_CONST) ++count;
    }
    /* a stale label only matters on a node that some box wrote */
    if (ok) {
        for (int b = 0; b < nbox && ok; ++b) {
            const ArtraBcBox *bx = boxes + b;
            for (int k = bx->box[Z][MIN]; k < bx->box[Z][MAX] && ok; ++k)
                for (int j = bx->box[Y][MIN]; j < bx->box[Y][MAX] && ok; ++j)
                    for (int i = bx->box[X][MIN]; i < bx->box[X][MAX]; ++i) {
                        if (LABEL_STALE == label[artra_index(k, j, i, part->n[Y], part->n[X])]) {
                            ok = 0;
                            break;
                        }
                    }
        }
    }
    if (!ok) {
        free(label);
        return 0;
    }
    out->dst = malloc(count * sizeof(int));
    out->src = malloc(count * sizeof(int));
    if (!out->dst || !out->src) {
        free(label);
        artra_free_bc_gather(out);
        return artra_set_error(1, "out of host memory");
    }
    /* full table, then per-axis subsets: nodes outside the interior range
     * along that axis only (the ghosts a sweep along it reads) */
    long n = 0;
    for (long q = 0; q < N; ++q) {
        if (label[q] >= 0 || label[q] <= LABEL_CONST) {
            out->dst[n] = (int)q;
            out->src[n] = label[q];
            ++n;
        }
    }
    out->count = n;
    for (int s = 0; s < DIMS; ++s) {
        long m = 0;
        for (long e = 0; e < n; ++e) {
            const int q = out->dst[e];
            const int c[DIMS] = {q % part->n[X], (q / part->n[X]) % part->n[Y], q / (part->n[X] * part->n[Y])};
            int inside = 1;
            for (int t = 0; t < DIMS; ++t) {
                if (t == s) continue;
                if (c[t] < part->ns[PIN][t][MIN] || c[t] >= part->ns[PIN][t][MAX]) inside = 0;
            }
            if (inside) ++m;
        }
        out->axis_dst[s] = malloc((m ? m : 1) * sizeof(int));
        out->axis_src[s] = malloc((m ? m : 1) * sizeof(int));
        if (!out->axis_dst[s] || !out->axis_src[s]) {
            free(label);
            artra_free_bc_gather(out);
            return artra_set_error(1, "out of host memory");
        }
        long w = 0;
        for (long e = 0; e < n; ++e) {
            const int q = out->dst[e];
            const int c[DIMS] = {q % part->n[X], (q / part->n[X]) % part->n[Y], q / (part->n[X] * part->n[Y])};
            int inside = 1;
            for (int t = 0; t < DIMS; ++t) {
                if (t == s) continue;
                if (c[t] < part->ns[PIN][t][MIN] || c[t] >= part->ns[PIN][t][MAX]) inside = 0;
            }
            if (inside) {
                out->axis_dst[s][w] = q;
                out->axis_src[s][w] = out->src[e];
                ++w;
            }
        }
        out->axis_count[s] = m;
    }
    out->resolved = 1;
    free(label);
    return 0;
}

void artra_free_bc_gather(ArtraBcGather *g)
{
    free(g->dst);
    free(g->src);
    for (int s = 0; s < DIMS; ++s) {
        free(g->axis_dst[s]);
        free(g->axis_src[s]);
    }
    memset(g, 0, sizeof *g);
}
