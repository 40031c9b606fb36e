/*
 * artra_types.h - data types shared with ArtraCFD's C call surface.
 *
 * The drop-in entry points take the caller's own Space/Model/Time
 * structures, so every structure below reproduces the member order, member
 * types and array extents of the reference declarations in commons.h:27-273
 * (enum ComConst commons.h:27-116, Node commons.h:128-134, Partition
 * commons.h:136-159, Polyhedron commons.h:174-205, Geometry commons.h:207-214,
 * Space commons.h:224-228, Time commons.h:230-243, Model commons.h:245-268).
 * A caller compiled against the reference headers can therefore pass its
 * pointers straight to this library.  The names of the constants are the
 * reference's own so that code ported from the solver reads the same.
 *
 * When this header is used alongside the reference's commons.h, define
 * ARTRA_TYPES_FROM_REFERENCE before including it and the declarations below
 * are skipped (the two sets are layout-identical).
 */
#ifndef ARTRA_TYPES_H_
#define ARTRA_TYPES_H_

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef ARTRA_TYPES_FROM_REFERENCE

/* Index and size constants of the solver (values fixed by commons.h:27-116). */
enum ArtraConst {
    DIMO = 4, X = 0, Y = 1, Z = 2, DIMS = 3, PHI = 4,
    COLLAPSEN = 0, COLLAPSEX = 1, COLLAPSEY = 2, COLLAPSEZ = 3,
    COLLAPSEXY = 5, COLLAPSEXZ = 7, COLLAPSEYZ = 8, COLLAPSEXYZ = 17,
    DIMT = 3, TO = 0, TN = 1, TM = 2,
    DIMU = 5, DIMUo = 6,
    PATHN = 30, PATHSEP = 4, NONE = -1,
    WENOTHREE = 0, WENOFIVE = 1, OPTSPLIT = 0, OPTBYOPT = 1,
    NPART = 15, PIO = 0, PIN = 0, PWB = 1, PEB = 2, PSB = 3, PNB = 4, PFB = 5,
    PBB = 6, PWG = 7, PEG = 8, PSG = 9, PNG = 10, PFG = 11, PBG = 12, PHY = 13,
    PAL = 14, LIMIT = 2, MIN = 0, MAX = 1,
    NBC = 7, INFLOW = 0, OUTFLOW = 1, SLIPWALL = 2, NOSLIPWALL = 3, PERIODIC = 4,
    VARBC = 6,
    NIC = 10, ICGLOBAL = 0, ICPLANE = 1, ICSPHERE = 2, ICBOX = 3, ICCYLINDER = 4,
    POSIC = 7, VARIC = 5,
    DIMTK = 2, POLYN = 3, EVF = 4,
    NPROBE = 5, PROPT = 0, PROLN = 1, PROCV = 2, PROFC = 3, PROSD = 4, POSLN = 7,
    STR = 200, VARSTR = 100
};

typedef double Real;
typedef char String[STR];
typedef int IntVec[DIMS];
typedef Real RealVec[DIMS];

/* One grid node: flags plus the conservative state at three time levels. */
typedef struct {
    int did;                /* 0 fluid, k inside body k, -1 exterior */
    int fid;                /* closest facet of the owning body */
    int lid;                /* interfacial layer */
    int gst;                /* ghost layer */
    Real U[DIMT][DIMU];     /* (rho, rho u, rho v, rho w, rho eT) per level */
} Node;

typedef struct {
    IntVec m;
    IntVec n;
    IntVec ng;
    int gl;
    int collapse;
    RealVec d;
    RealVec dd;
    Real tinyL;
    int ns[NPART][DIMS][LIMIT];
    int np[DIMS][DIMS][LIMIT];
    int path[PATHN][DIMS];
    int pathSep[PATHSEP];
    int *typeBC;
    int (*N)[DIMS];
    Real (*varBC)[VARBC];
    int nIC;
    int *typeIC;
    Real (*posIC)[POSIC];
    char (*varIC)[VARIC][VARSTR];
    Real domain[DIMS][LIMIT];
    IntVec proc;
    int procN;
} Partition;

typedef struct {
    RealVec N;
    RealVec v0;
    RealVec v1;
    RealVec v2;
} Facet;

typedef struct {
    int gid;
    IntVec N;
} Collision;

typedef struct {
    int faceN;              /* <= 0 for an analytical sphere */
    int edgeN;
    int vertN;
    int state;              /* 1 stationary, 2 force-free, 0 moving */
    int mid;
    Real r;
    RealVec O;
    Real I[DIMS][DIMS];
    Real V[DIMTK][DIMS];
    Real W[DIMTK][DIMS];
    Real at[DIMTK][DIMS];
    RealVec g;
    Real ar[DIMTK][DIMS];
    RealVec Fp;
    RealVec Fv;
    RealVec Tt;
    Real to;
    Real rho;
    Real T;
    Real cf;
    Real area;
    Real volume;
    Real box[DIMS][LIMIT];
    int (*f)[POLYN];
    Real (*Nf)[DIMS];
    int (*e)[EVF];
    Real (*Ne)[DIMS];
    Real (*v)[DIMS];
    Real (*Nv)[DIMS];
    Facet *facet;
} Polyhedron;

typedef struct {
    int totN;
    int sphN;
    int stlN;
    int colN;
    Polyhedron *poly;
    Collision *col;
} Geometry;

typedef struct {
    Real eos;
} Material;

typedef struct {
    Node *node;
    Geometry geo;
    Partition part;
} Space;

typedef struct {
    int restart;
    int stepN;
    int stepC;
    int dataN[NPROBE];
    int dataW[NPROBE];
    int dataStreamer;
    int dataC;
    Real end;
    Real now;
    Real numCFL;
    Real (*pp)[DIMS];
    Real (*lp)[POSLN];
} Time;

typedef struct {
    int tScheme;
    int sScheme;
    int sL;
    int sR;
    int multidim;
    int jacobMean;
    int fluxSplit;
    int psi;
    int ibmLayer;
    int mid;
    int gState;
    int sState;
    Real refMa;
    Real refMu;
    Real gamma;
    Real gasR;
    Real cv;
    Real refL;
    Real refRho;
    Real refV;
    Real refT;
    RealVec g;
    Material *mat;
} Model;

#endif /* ARTRA_TYPES_FROM_REFERENCE */

#ifdef __cplusplus
}
#endif

#endif /* ARTRA_TYPES_H_ */
