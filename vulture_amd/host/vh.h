/*
 * vh.h - host-side mesh parser and solver set-up for the MI355X FDTD core.
 *
 * This library reads Vulture mesh files and builds, for one solve, every
 * array and list the time-stepping loop needs: grid extents and per-line edge
 * lengths, update coefficients, UPML profiles and slab limits, Mur faces,
 * Debye blocks, SIBC surfaces, plane waves, sources and observers.  All state
 * lives in one vh_solve object, so many solves can coexist in one process
 * (the reference keeps the same state in module globals, one solve per
 * process).
 *
 * Arithmetic follows the reference's types: `real` is the working precision
 * (double for the product, float when the library is compiled with
 * -DVH_REAL=float to reproduce the reference float build bit for bit), Debye
 * and SIBC coefficients are double/double complex as in the reference.
 *
 * Array layout is the reference's: A[i][j][k] with k fastest, allocation
 * extents numCells[3] = gobox extent + 2 ghost layers (grid.c:357-387), flat
 * index ((i*NY)+j)*NZ+k.  The device library transposes to x-fastest.
 */
#ifndef VH_H
#define VH_H

#include <stdio.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <complex.h>

#ifndef VH_REAL
#define VH_REAL double
#endif
typedef VH_REAL real;

#ifdef __cplusplus
extern "C" {
#endif

/* Faces, axes and field components: same order and values as the reference
 * (fdtd_types.h:58-112). */
enum { XLO = 0, XHI, YLO, YHI, ZLO, ZHI };
enum { XDIR = 0, YDIR, ZDIR, CA_UNDEFINED };
enum { EX = 0, EY, EZ, HX, HY, HZ };
enum { BB_INVALID = 0, BB_POINT, BB_LINE, BB_SURFACE, BB_VOLUME };

typedef unsigned char FaceMask;
#define FACE_MASK_ERROR 0xFF
#define FACE_MASK_ALL 0x3F

/* Grid types (grid.h:46-54). */
enum { GT_CUBIC = 0, GT_UNIFORM, GT_NONUNIFORM };

/* Medium types (medium.h:49-57). */
enum { MT_FREE_SPACE = 0, MT_PEC = 1, MT_SIMPLE = 2, MT_DEBYE = 3, MT_UNDEFINED };

/* Boundary types (boundary.h:51-63). */
enum { BT_PMC = 0, BT_PML = 1, BT_PEC = 2, BT_FREE_SPACE = 3, BT_PERIODIC = 4,
       BT_MUR = 5, BT_SIBC = 6, BT_UNDEFINED };

/* Waveform types (waveform.h:43-57). */
enum { WT_GAUSSIAN_PULSE = 0, WT_NARROW_GAUSSIAN_PULSE, WT_DIFFERENTIATED_GAUSSIAN_PULSE,
       WT_RICKER_WAVELET, WT_MODULATED_GAUSSIAN_PULSE, WT_COMPACT_PULSE,
       WT_DIFFERENTIATED_COMPACT_PULSE, WT_MODULATED_COMPACT_PULSE, WT_RAMPED_SINUSOID,
       WT_EXTERNAL, WT_UNDEFINED };

/* Source types (source.h:41-59). */
enum { ST_EFIELD = 0, ST_HFIELD, ST_ELEC_CURR_DENSITY, ST_MAGN_CURR_DENSITY,
       ST_ELEC_SURF_CURR_DENSITY, ST_MAGN_SURF_CURR_DENSITY, ST_ELEC_CURRENT,
       ST_MAGN_CURRENT, ST_ELEC_CURRENT_MOMENT, ST_MAGN_CURRENT_MOMENT, ST_VOLTAGE,
       ST_THEVENIN_VOLTAGE, ST_NORTON_CURRENT, ST_UNDEFINED };

/* Observer formats/domains/quantities (observer.h:40-88). */
enum { OF_ASCII = 0, OF_BINARY, OF_HDF5 };
enum { OD_TIME = 0, OD_FREQ };
enum { OQ_WF = 0, OQ_E, OQ_H, OQ_EH };

#define VH_TAG 32
#define VH_PATH 1024

typedef struct {
  char name[VH_TAG];
  int type;
  real eps_r, sigma, mu_r;
  int numPoles;
  double complex *residues, *poles;
  char fileName[VH_PATH];
  real alpha, beta, gamma;
  double complex *dalpha, *dbeta;
} vh_medium;

/* One recursive-convolution filter (filter.h:40-46): y = asymp*x + sum Re(zeta_k). */
typedef struct {
  int numPoles;
  double asymp;
  double complex *q0, *q1, *q2;
  /* pole-residue form it came from */
  double complex *poles, *residues;
} vh_rcfilter;

typedef struct {
  char name[VH_TAG];
  int type;
  int numLayers, order;
  real n_eff, refCoeff, kmax;
  char fileName[VH_PATH];
  real S_TM[2][2], S_TE[2][2];
  /* SIBC: 4x4 recursive-convolution matrix, row-major [p][q]. */
  int hasRc;
  vh_rcfilter rc[16];
} vh_boundary;

typedef struct { int mbbox[6]; int medium; FaceMask mask; } vh_block;

/* A voxel object (MT <name> VOXEL "<file>"): a key array and its key table. */
typedef struct {
  char name[VH_TAG];
  int bytes, n[3];            /* bytes per key; voxels per axis */
  double d[3], ref[3];        /* voxel size and reference point, m */
  unsigned char *data;        /* n[0]*n[1]*n[2] little-endian keys, k fastest */
  uint32_t *keys;             /* sorted keys ... */
  int *keyMedium;             /* ... and their media */
  int numKeys;
} vh_voxel;

/* One placement of a voxel object (MB <node> <object> [theta [phi]]). */
typedef struct { int object, node[3]; double theta, phi; } vh_voxel_place;

typedef struct { int mbbox[6]; int boundary; int orient; real angle; int gbbox[6]; } vh_surface;

typedef struct { int mbbox[6]; int wire; } vh_line;

typedef struct {
  char name[VH_TAG];
  int type;
  real size, delay, width, frequency;
  char fileName[VH_PATH];
  unsigned long tableSize, lastIdx;
  real *table;   /* tableSize x 3: time, value, second derivative */
} vh_waveform;

typedef struct {
  char name[VH_TAG];
  int type, pol, normal, field;
  bool isSoft;
  real size, delay, resistance;
  int waveform;
  int mbbox[6], gbbox[6];
  int flim[6][6];
} vh_source;

typedef struct {
  char name[VH_TAG];
  int mbbox[6];
  int waveform;
  bool isActive[6];
  real theta, phi, eta, size, delay;
  int gbbox[6];
  int flim[6][6][6];
  real kinc[3], Finc[6], ijk0[3], r0[3];
  real phaseVelocity;
  /* 1-D auxiliary grid (cubic grids only, planewave.c:1226-1296). */
  int nx, xb;
  real *Eyi, *Hzi, *Pyi, *PPyi, *Bzi, *adx, *bdx, *ahx, *bhx;
  real betaEyi, gammaHzi;
} vh_planewave;

typedef struct {
  char name[VH_TAG];
  int format, domain, quantity;
  int mbbox[6], step[3];
  int waveform;
  bool isInternal;
  int numComp;
  int gbbox[6];
  FILE *outputFile;
  real *dft_re, *dft_im;     /* [numComp][numFreq] running DFT (observer.c:929-930) */
  int waveformObserver;      /* index of the FD waveform observer used to normalise */
} vh_observer;

typedef struct {
  int flim[3][6];
  int medium;
  size_t n[3];                /* cells per component */
  double complex *J[3];       /* [cells][numPoles] */
  real *lastE[3];
} vh_debye;

typedef struct {
  int gbbox[6];               /* +1 in the normal direction, sibc.c:253-267 */
  int orientation, normal;
  real cosa, sina;
  int boundary;
  size_t nfaces;
  double complex *zeta;       /* [face][16][maxPoles] */
  double *old;                /* [face][16] */
  int maxPoles;
  real *Etan;                 /* [face][4] */
  bool *isAdjA, *isAdjB;      /* [face][4] */
} vh_sibc;

typedef struct vh_solve {
  char err[1024];
  void *jb;                   /* error unwind target while parsing/initialising */
  int deviceDft;              /* EH frequency observers are accumulated elsewhere */
  char inputDir[VH_PATH];
  char outputDir[VH_PATH];
  int fullPrecision;          /* write observer data with %24.16e instead of %16.8e */
  int openOutputs;            /* write observer files as the run goes */
  int binaryStreamed;         /* excite.dat and impulse.dat rows are written by
                                 vh_observe_binary_step as runs complete, not by
                                 vh_observe_step */

  /* physical constants (physical.c:37-48), in `real` */
  real pi, c0, eps0, mu0, eta0;

  /* simulation */
  unsigned long numTimeSteps;
  real courantNumber;

  /* parser state */
  int isFound[27];
  char comment[1024];
  int meshVersion[3];
  bool foundSource, foundOutput;

  /* grid */
  int numCells[3], numLines[3];
  int mbox[6], gibox[6], gobox[6], ggbox[6];
  int gfilim[6][6], gfolim[6][6], gfglim[6][6];
  real *xlines, *ylines, *zlines;
  real *dex, *dey, *dez, *dhx, *dhy, *dhz;
  real *idex, *idey, *idez, *idhx, *idhy, *idhz;
  real dt, dmin[3], dmax[3], duni[3];
  int gridType;
  size_t ncell;               /* numCells product */
  real *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;
  real *alphaEx, *alphaEy, *alphaEz, *betaEx, *betaEy, *betaEz;
  real *gammaHx, *gammaHy, *gammaHz;

  /* media, boundaries, blocks, surfaces, wires, lines */
  vh_medium *media; int numMedia, capMedia;
  vh_boundary *bounds; int numBounds, capBounds;
  int extBoundary[6];
  vh_block *blocks; int numBlocks, capBlocks;
  vh_voxel *voxels; int numVoxels, capVoxels;
  vh_voxel_place *voxelPlaces; int numVoxelPlaces, capVoxelPlaces;
  vh_surface *surfaces; int numSurfaces, capSurfaces;
  char (*wireNames)[VH_TAG]; int *wireTypes; int numWires, capWires;
  vh_line *lines; int numLinesTW, capLinesTW;

  /* PML (pml.c:44-84) */
  int pbox[6][6], fplim[6][6][6];
  real *PPx[6], *PPy[6], *PPz[6], *Px[6], *Py[6], *Pz[6], *Bx[6], *By[6], *Bz[6];
  real *adx, *bdx, *gdx, *kdx, *ady, *bdy, *gdy, *kdy, *adz, *bdz, *gdz, *kdz;
  real *ahx, *bhx, *ghx, *khx, *ahy, *bhy, *ghy, *khy, *ahz, *bhz, *ghz, *khz;
  real *ibdx, *ibdy, *ibdz, *ibhx, *ibhy, *ibhz;

  /* Mur (mur.c:41-53) */
  int murbox[6][6], fmlim[6][6][6];
  real zeta[6];

  /* Debye blocks and SIBC surfaces in list order */
  vh_debye *debye; int numDebye;
  vh_sibc *sibc; int numSibc;

  /* waveforms, sources, plane waves */
  vh_waveform *waveforms; int numWaveforms, capWaveforms;
  vh_source *sources; int numSources, capSources;
  vh_planewave *pws; int numPws, capPws;
  bool useAuxGrid;

  /* observers */
  vh_observer *obs; int numObs, capObs;
  bool isOT, isOF;
  unsigned long startTimeStep, stopTimeStep, numOutTimeSteps;
  real startTime, stopTime;
  real startFreq, stopFreq, stepFreq;
  unsigned long numFreq;
  real *omega;
  FILE *impulseFile, *exciteFile;
  real obsDt;                 /* dt cast through float, observer.c:413 */
} vh_solve;

/* Flat index into a k-fastest [NX][NY][NZ] array. */
#define VH_IDX(s, i, j, k) ((((size_t)(i)) * (size_t)(s)->numCells[1] + (size_t)(j)) * (size_t)(s)->numCells[2] + (size_t)(k))

/* Lifecycle. */
vh_solve *vh_create(void);
void vh_destroy(vh_solve *s);
const char *vh_error(const vh_solve *s);
void vh_set_dirs(vh_solve *s, const char *inputDir, const char *outputDir);

/* Parse a mesh file (mesh.c:269-405). Returns 0 on success, -1 on error. */
int vh_read_mesh(vh_solve *s, const char *path);

/* Run every init step in the reference's order (vulture.c:100-138). */
int vh_init(vh_solve *s);

/* Observer helpers shared by the CPU restatement and the device driver. */
int vh_open_observers(vh_solve *s);
/* Free the volume arrays (fields, coefficients, UPML and Debye state) once a
 * device holds them; observers, waveforms and grid lines stay. */
void vh_release_grid_arrays(vh_solve *s);
void vh_observe_step(vh_solve *s, unsigned long n, real t, const double *pointVals, const float *binVals);
void vh_observe_binary_step(vh_solve *s, unsigned long n, real t, const float *binVals);
void vh_observer_point_values(const vh_solve *s, double *vals);
void vh_close_observers(vh_solve *s);
int vh_num_point_observers(const vh_solve *s, int domain);

/* Waveform value at time t (waveform.c:590-638). */
real vh_waveform_value(vh_solve *s, real t, int waveform, real delay);

/* Helpers from the reference used by several modules. */
int vh_bbox_type(const int bbox[6]);
void vh_set_field_limits(const int bbox[6], int fieldLimits[6][6], const bool includeBoundary[6]);
void vh_field_physical_location(const vh_solve *s, real r[3], int field, int i, int j, int k);
void vh_field_index_location(real ijk[3], int field, int i, int j, int k);
bool vh_is_pmc_edge(const vh_solve *s, int direction, int index);
int vh_outer_type(const vh_solve *s, int face);

#ifdef __cplusplus
}
#endif

#endif
