/*
 * abi_fields.h - the structure members whose offsets make up the drop-in
 * ABI (commons.h:128-268), as an X-macro list: ABI_SIZE(type) for every
 * structure and ABI_FIELD(type, member) for every member.
 *
 * Test infrastructure: the reference harness expands it against the
 * reference's commons.h (oracle/ref_harness.c, "layout" mode) and
 * tests/abi_layout.c against include/artra_types.h; tests/test_host.py
 * compares the two dumps.
 */
#define ABI_LIST                                                                                          \
    ABI_SIZE(Node) ABI_FIELD(Node, did) ABI_FIELD(Node, fid) ABI_FIELD(Node, lid) ABI_FIELD(Node, gst)      \
    ABI_FIELD(Node, U)                                                                                    \
    ABI_SIZE(Partition) ABI_FIELD(Partition, m) ABI_FIELD(Partition, n) ABI_FIELD(Partition, ng)          \
    ABI_FIELD(Partition, gl) ABI_FIELD(Partition, collapse) ABI_FIELD(Partition, d) ABI_FIELD(Partition, dd) \
    ABI_FIELD(Partition, tinyL) ABI_FIELD(Partition, ns) ABI_FIELD(Partition, np) ABI_FIELD(Partition, path) \
    ABI_FIELD(Partition, pathSep) ABI_FIELD(Partition, typeBC) ABI_FIELD(Partition, N)                    \
    ABI_FIELD(Partition, varBC) ABI_FIELD(Partition, nIC) ABI_FIELD(Partition, typeIC)                    \
    ABI_FIELD(Partition, posIC) ABI_FIELD(Partition, varIC) ABI_FIELD(Partition, domain)                  \
    ABI_FIELD(Partition, proc) ABI_FIELD(Partition, procN)                                                \
    ABI_SIZE(Facet) ABI_FIELD(Facet, N) ABI_FIELD(Facet, v0) ABI_FIELD(Facet, v1) ABI_FIELD(Facet, v2)     \
    ABI_SIZE(Collision) ABI_FIELD(Collision, gid) ABI_FIELD(Collision, N)                                 \
    ABI_SIZE(Polyhedron) ABI_FIELD(Polyhedron, faceN) ABI_FIELD(Polyhedron, edgeN)                        \
    ABI_FIELD(Polyhedron, vertN) ABI_FIELD(Polyhedron, state) ABI_FIELD(Polyhedron, mid)                  \
    ABI_FIELD(Polyhedron, r) ABI_FIELD(Polyhedron, O) ABI_FIELD(Polyhedron, I) ABI_FIELD(Polyhedron, V)   \
    ABI_FIELD(Polyhedron, W) ABI_FIELD(Polyhedron, at) ABI_FIELD(Polyhedron, g) ABI_FIELD(Polyhedron, ar) \
    ABI_FIELD(Polyhedron, Fp) ABI_FIELD(Polyhedron, Fv) ABI_FIELD(Polyhedron, Tt) ABI_FIELD(Polyhedron, to) \
    ABI_FIELD(Polyhedron, rho) ABI_FIELD(Polyhedron, T) ABI_FIELD(Polyhedron, cf)                         \
    ABI_FIELD(Polyhedron, area) ABI_FIELD(Polyhedron, volume) ABI_FIELD(Polyhedron, box)                  \
    ABI_FIELD(Polyhedron, f) ABI_FIELD(Polyhedron, Nf) ABI_FIELD(Polyhedron, e) ABI_FIELD(Polyhedron, Ne) \
    ABI_FIELD(Polyhedron, v) ABI_FIELD(Polyhedron, Nv) ABI_FIELD(Polyhedron, facet)                       \
    ABI_SIZE(Geometry) ABI_FIELD(Geometry, totN) ABI_FIELD(Geometry, sphN) ABI_FIELD(Geometry, stlN)      \
    ABI_FIELD(Geometry, colN) ABI_FIELD(Geometry, poly) ABI_FIELD(Geometry, col)                          \
    ABI_SIZE(Material) ABI_FIELD(Material, eos)                                                           \
    ABI_SIZE(Space) ABI_FIELD(Space, node) ABI_FIELD(Space, geo) ABI_FIELD(Space, part)                   \
    ABI_SIZE(Time) ABI_FIELD(Time, restart) ABI_FIELD(Time, stepN) ABI_FIELD(Time, stepC)                 \
    ABI_FIELD(Time, dataN) ABI_FIELD(Time, dataW) ABI_FIELD(Time, dataStreamer) ABI_FIELD(Time, dataC)    \
    ABI_FIELD(Time, end) ABI_FIELD(Time, now) ABI_FIELD(Time, numCFL) ABI_FIELD(Time, pp)                 \
    ABI_FIELD(Time, lp)                                                                                   \
    ABI_SIZE(Model) ABI_FIELD(Model, tScheme) ABI_FIELD(Model, sScheme) ABI_FIELD(Model, sL)              \
    ABI_FIELD(Model, sR) ABI_FIELD(Model, multidim) ABI_FIELD(Model, jacobMean) ABI_FIELD(Model, fluxSplit) \
    ABI_FIELD(Model, psi) ABI_FIELD(Model, ibmLayer) ABI_FIELD(Model, mid) ABI_FIELD(Model, gState)       \
    ABI_FIELD(Model, sState) ABI_FIELD(Model, refMa) ABI_FIELD(Model, refMu) ABI_FIELD(Model, gamma)      \
    ABI_FIELD(Model, gasR) ABI_FIELD(Model, cv) ABI_FIELD(Model, refL) ABI_FIELD(Model, refRho)           \
    ABI_FIELD(Model, refV) ABI_FIELD(Model, refT) ABI_FIELD(Model, g) ABI_FIELD(Model, mat)

/* One line per entry: "size Type N" or "field Type.member offset size". */
#define ABI_SIZE(T) printf("size %s %zu\n", #T, sizeof(T));
#define ABI_FIELD(T, f) printf("field %s.%s %zu %zu\n", #T, #f, offsetof(T, f), sizeof(((T *)0)->f));
