// Internal declarations of the MI355X FDTD core library (not part of the C-ABI).
//
// Device layout: every 3-D array is x-fastest, index (k*NY + j)*PX + i, with
// PX the x extent rounded up to 128 bytes so rows start on cache-line
// boundaries.  Extents NX,NY,NZ are the reference's allocation extents
// (gobox plus one ghost layer each side), so grid indices i,j,k are the
// reference's own.
//
// With the fused leapfrog sweep the fields are held twice.  A step reads the
// fields of step n from one set and writes those of step n+1 to the other
// (the sweeps, Mur E); the corrections after each sweep (Debye, SIBC,
// sources, TF/SF, ghosts) then modify the new set in place.  Nothing a
// kernel reads from the old set is written during the step, so a sweep may
// recompute a neighbour's new value from the old fields at any time, in any
// block, without racing the block that stores it.  The separate sweeps
// update one set in place, as the reference does.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include "vgpu.h"

namespace vgpu {

// ---- kernel argument blocks (passed by value) ----

// Full argument block of the E sweep, kept in device memory and read through
// a pointer by the shell kernel and the fused Debye step.
template <typename T> struct EArgs {
  int NX, NY, NZ, PX;
  long long plane;
  T *E[3];               // new E (written)
  const T *Ein[3];       // old E (read)
  const T *H[3];         // old H
  const T *alpha[3], *beta[3];
  const T *idh[3];
  int in[3][6];          // inner-grid limits per E component
  unsigned pmlmask;
  int pl[6][3][6];       // UPML slab limits [region][component][face]
  T *PP[6][3], *P[6][3]; // UPML auxiliaries, x-fastest within each slab box
  T *PPpool, *Ppool;     // the pools holding them (same layout: a slab is at the same offset in both)
  const T *ad[3], *bd[3], *ah[3], *ibh[3];
  // coefficient classes: per-cell class, [class][8] alpha/beta (6 used), [class][3] Debye medium
  const uint16_t *ecls;
  const T *etab;
  const int *edslot;
  // layer-0 Debye sites applied in the sweep
  int dfuse;
  const int *dbase;      // [3][NZ][NY][NCH] first site of each 64-cell chunk of the fast box
  T *dlast, *dJ;
  long long dstride;
  const double *dalpha, *dbeta;
  const int *dnp;
  int dmaxP, dcplx;
  unsigned *vis[3];      // update-once check: visit counts per component (null in ordinary runs)
};

template <typename T> struct HArgs {
  int NX, NY, NZ, PX;
  long long plane;
  T *H[3];               // new H (written)
  const T *Hin[3];       // old H (read)
  const T *E[3];         // new E
  const T *gamma[3];
  const T *ide[3];
  int in[3][3][6];       // ordinary-update boxes per H component (inner grid, two Mur faces)
  unsigned pmlmask;
  int pl[6][3][6];
  T *B[6][3];
  T *Bpool;              // the pool holding the B slabs
  const T *ad[3], *ah[3], *bh[3], *ibd[3];
  const uint16_t *hcls;  // coefficient classes: per-cell class, [class][4] gamma (3 used)
  const T *htab;
  unsigned *vis[3];      // update-once check (null in ordinary runs)
};

// Lean by-value arguments of the fast sweep kernels.
template <typename T> struct FastE {
  int NX, NY, NZ, PX;
  long long plane;
  int box[6];            // fast box (every E component inner-grid)
  int kchunk, nchx;      // planes per block; 64-cell x chunks from box[0]
  int gy, gz;            // y blocks (blockDim.y rows), z chunks (blocks = nchx*gy*gz, launched in XCD order)
  T *E[3];               // new E (written)
  const T *Ein[3];       // old E
  const T *H[3];         // old H
  const T *idh[3];
  const uint16_t *ecls;  // coefficient classes
  const T *etab;
  const T *alpha[3], *beta[3];  // per-cell arrays when there are no classes
  int dfuse;             // apply layer-0 Debye sites
  const EArgs<T> *ap;    // full argument block in device memory (Debye state)
  unsigned *vis[3];      // update-once check (null in ordinary runs)
};

template <typename T> struct FastH {
  int NX, NY, NZ, PX;
  long long plane;
  int box[6];
  int kchunk, nchx;
  int gy, gz;
  T *H[3];               // new H (written)
  const T *Hin[3];       // old H
  const T *E[3];         // new E
  const T *ide[3];
  const uint16_t *hcls;
  const T *htab;
  const T *gamma[3];
  unsigned *vis[3];      // update-once check (null in ordinary runs)
};

// Fused leapfrog sweep over the fused box (E and H of every cell inside take
// the ordinary update).  A wave owns a tile of 63 x-cells by R rows and
// marches over kz planes; it computes E on 64 x 1+R cells per plane (its
// tile and the x and y neighbours its H needs) and H of the plane below.
template <typename T> struct LeapArgs {
  int NX, NY, NZ, PX;
  long long plane;
  int box[6];            // fused box
  int ntx, ngy, ntz, kz; // x tiles (63 cells), y tiles (4 waves of R rows per block), z chunks of kz planes
  T *E[3];               // new E, H (written)
  T *H[3];
  const T *Ein[3];       // old E, H
  const T *Hin[3];
  const T *idh[3], *ide[3];
  const uint16_t *ecls, *hcls;  // coefficient classes
  const T *etab, *htab;
  const T *alpha[3], *beta[3], *gamma[3];  // per-cell arrays when there are no classes
  int discard;           // first element of the discard area past every field array's cells
  unsigned *vis[6];      // update-once check (null in ordinary runs)
};

// Elements past the cells of every field array that the fused sweep writes
// the values of cells a lane does not own to (never read).
constexpr int FIELD_DISCARD = 65536;

// H cells of the fused box whose new value the fused sweep cannot give
// (an E input of theirs is corrected after the sweep -- a Debye site, a
// source or TF/SF row, an SIBC face -- or is not an ordinary update): one
// 64-cell x chunk of one row per entry, with a mask of its cells.  They are
// updated again after the E corrections, from the old H and the final E.
struct FixEnt {
  int i0, j, k, pad_;    // first cell of the chunk
  unsigned long long mask;  // cells i0 + lane to update
};
template <typename T> struct FixArgs {
  long long n;
  const FixEnt *ent;
  int NY, PX;
  long long plane;
  T *H[3];
  const T *Hin[3];
  const T *E[3];
  const T *ide[3];
  const uint16_t *hcls;
  const T *htab;
  const T *gamma[3];
};

// A piece of a sweep's shell (the cells outside its fast box): a box whose
// rows share, per component, one sequence of x runs, each run one region
// (ordinary update or one UPML slab) over a range of x.  A block covers 256
// consecutive cells of the piece's x-y plane (x fastest) and marches over
// `kc` planes; block b of the piece is x-y block b % nxyb of z chunk
// b / nxyb.
constexpr int SHELL_RUNS = 8;
constexpr int SHELL_KC = 4;  // planes a shell thread marches over
struct ShellRun {
  int xa, xb;            // inclusive x range
  int reg, pad_;         // 6 ordinary, 0..5 UPML slab
  long long p0;          // UPML: slab index of (xa, box[2], box[4])
  long long psy, psz;    // slab index strides of j and k (1 for i)
  void *a, *b;           // UPML auxiliaries of the slab: E sweep PP and P, H sweep B
};
struct ShellPiece {
  int box[6];            // inclusive cell limits
  int ni, nj, nk, cells; // extents, cell count
  int nxy, nxyb, kc, nblk;  // x-y plane lanes (nip per row), its 256-lane blocks, planes per block, blocks
  int nrun[3];           // runs per component (cells outside every run are not updated)
  int nip, xo;           // lanes per row (a whole number of 128-byte lines) and the row's first cell's lane
  int pad_;
  ShellRun run[3][SHELL_RUNS];
};

// One block of a shell launch: its piece and the piece's block (piece < 0:
// nothing to do).

struct ShellWork {
  int piece, block;
};

// Mur tangential-E faces (mur.c:135-395): one box per (face, component).
template <typename T> struct MurArgs {
  int NX, NY, PX;
  long long plane;
  T *E[3];               // new E (written on the face)
  const T *Ein[3];       // old E
  const T *H[3];         // old H
  const T *alpha[3], *beta[3];
  const T *idh[3];
  int box[12][6];        // limits of face f component slot s at [2f+s]
  int comp[12];          // component of each slot (-1 unused)
  T zeta[6];
  const uint16_t *ecls;  // coefficient classes (null: read the per-cell alpha/beta arrays)
  const T *etab;
  unsigned *vis[3];      // update-once check (null in ordinary runs)
};

// Debye sites of one layer (debye.c:243-307): a layer holds at most one site
// per E cell, layers run in block order.
template <typename T> struct DebyeArgs {
  long long n;           // sites in this layer
  const long long *key;  // per site: medium << 42 | device linear index << 2 | component
  const T *sbeta;        // per site: beta of the site's edge
  T *E[3];
  T *lastE;              // per site
  T *J;                  // [pole][site] (real poles) or [pole][site][2]
  long long stride;      // sites in the whole state block (J pole stride)
  long long base;        // first site of this layer in the state block
  const double *dalpha, *dbeta;  // [medium][maxPoles][2]
  const int *npoles;     // [medium]
  int maxPoles;
  int cplx;
};

// A block of the SIBC E launch: up to 128 faces of one surface, so the
// surface, its filters and its matrices are uniform across the block.
struct SibcBlock {
  int surf, pad_;
  long long f0, f1;      // first face of the block, end of the surface's faces
};

// SIBC faces (sibc.c:443-586): per-face geometry, per-surface constants.
template <typename T> struct SibcEArgs {
  long long nfaces;
  const SibcBlock *blocks;
  const int *ijk;        // [face][3]
  const int *surf;       // [face]
  const unsigned char *adjA, *adjB;  // [face][4]
  const int *normal;     // [surface]
  const double *A, *B;   // [surface][16]
  const int *filtset;    // [surface] -> filter set
  const int *npoles;     // [set][16]
  const double *asymp;   // [set][16]
  const double *q;       // [set][16][maxPoles][3][2]
  int maxPoles;
  int real;              // all coefficients real: zeta holds real states only
  int factored;          // one state per (input, pole): zeta [(input*maxPoles + pole)][face]
  const int *fbase;      // factored: [set][input] filter whose coefficients drive the input's states, or -1
  const double *fratio;  // factored: [set][16][maxPoles] each filter's residue over its base filter's
  double *zeta;          // [(filter*maxPoles + pole)][face][2], or [..][face] when real
  double *old;           // [input][face]: the previous step's filter input
  T *Etan;               // [face][4]
  T *E[3];
  const T *H[3];
  int NY, PX;
  long long plane;
};

// Gathered SIBC H corrections (sibc.c:616-776): one row per touched H cell,
// entries in the reference's application order.
// Device form (16 bytes).  bits: kind (bit 0: 0 edge term, 1 parallel-
// adjacent term), Etan slots ea (bits 1-2) and eb (3-4), negative sign (5),
// full edge weight (6: weight 1, else 1/2 = 0.5*(1 + PMC edge)), !isAdjA
// (7), !isAdjB (8), axis of the inverse lengths (9-10); ia/ib index the
// inverse primary edge lengths of that axis.
struct SibcHEnt {
  int face;
  int bits;
  int ia, ib;
};

template <typename T> struct SibcHArgs {
  long long nrows;
  const long long *rowkey;   // cell << 3 | zero << 2 | component; zero: normal H of a face, set to 0
  const int *rowstart;   // [nrows+1]
  const T *rowg;         // the row cell's gamma
  const SibcHEnt *ent;
  const T *Etan;
  T *H[3];
  const T *ide[3];
};

// Sources and TF/SF corrections applied after the volume update, one row
// per touched cell, entries in reference order (sources first, then plane
// waves face by face, planewave.c:576-1007, source.c:602-753).
struct CorrEntry {
  int kind;              // 0 field source, 1 density source, 2 TF/SF aux grid, 3 TF/SF analytic
  int idx;               // source or plane-wave index
  int inc;               // incident field component (TF/SF)
  int auxid;             // floor(d) (aux grid)
  double sign;           // TF/SF sign
  double il;             // inverse edge length (TF/SF)
  double geo;            // dp (aux grid) or d/v (analytic)
};

// Device form of a correction entry (24 bytes).  code: kind (bits 0-1),
// incident component (bits 2-4), negative sign (bit 5), source or plane-wave
// index (bits 8-31).
struct CorrEnt {
  int code;
  int auxid;
  double geo;            // dp (aux grid) or d/v (analytic)
  double il;             // inverse edge length (TF/SF)
};

struct WaveformDev {
  int type;
  double size, delay, width, frequency;
  const double *table;   // [n][3]
  unsigned long n;
};

struct SourceDev {
  int kind, soft, waveform;
  double size, delay;
};

struct PwDev {
  double Finc[6];
  int waveform;
  double delay;
  int use_aux, nx, xb;
  double beta, gamma;
};

template <typename T> struct CorrArgs {
  long long nrows;
  const long long *rowkey;   // cell << 2 | component, ascending
  const int *rowstart;
  const CorrEnt *ent;
  const T *rowcoef;      // the row cell's beta (E) or gamma (H)
  const T *rowalpha;     // the row cell's alpha (E only)
  T *F[3];               // E or H components
  int electric;
  const unsigned long long *step;  // device step counter: n = *step + soff
  int soff;
  double dt, half;       // t = (n + half) * dt: timeE (half 0) or timeH (half 0.5)
  const WaveformDev *wf;
  const SourceDev *src;
  const PwDev *pw;
  const T *aux;          // aux grids [pw][2][auxmax]: Eyi then Hzi
  int auxmax;
  double eta0;
};

template <typename T> struct AuxArgs {
  int npw;
  const PwDev *pw;
  T *aux;                // [pw][2][auxmax]
  T *pml;                // [pw][3][10]: PPyi, Pyi, Bzi
  const T *prof;         // [pw][4][10]: adx, bdx, ahx, bhx
  int auxmax;
  const WaveformDev *wf;
  const unsigned long long *step;  // t = (*step + soff) * dt (timeE)
  int soff;
  double dt;
};

// Point observers and DFT (observer.c:622-650, 857-899).
template <typename T> struct ObsArgs {
  int nobs;
  const long long *cell; // [obs]
  const unsigned long long *first, *last;
  const int *nfreq;
  const long long *dftoff;   // offset of the observer's [6][nfreq] block
  const double *omega;       // concatenated per observer
  double *td;                // [obs][nsteps][6]
  unsigned long long tdsteps;
  double *dre, *dim;
  const T *F[6];
  const unsigned long long *step;  // n = *step + soff, t = n * dt
  int soff;
  double dt;
};

template <typename T> struct VolObsArgs {
  const long long *nodecell; // [node] for all volume observers
  const int *nodeobs;
  long long nnodes;
  const unsigned long long *first, *last;
  float *stage;              // [row][node][6], row = step modulo ring
  const T *F[6];
  const unsigned long long *step;  // n = *step + soff
  int soff;
  unsigned long long ring;   // staging rows: two runs' worth
};

// ---- packed launches over several solves (vgpu_run_group) ----

constexpr int PACK_MAX = VGPU_GROUP_MAX;  // solves in one group run

// One solve's by-value argument blocks for the packed launches, for field set
// 0 (the one the separate sweeps read and write), kept in device memory.  A
// packed launch passes the step offset itself.
template <typename T> struct PackArgs {
  FastE<T> fe;
  FastH<T> fh;
  const ShellPiece *epc, *hpc;
  const ShellWork *ewk, *hwk;
  const EArgs<T> *ea;
  const HArgs<T> *ha;
  DebyeArgs<T> da;
  CorrArgs<T> ce, ch;    // E and H corrections
  ObsArgs<T> oa;
  AuxArgs<T> xa;
  VolObsArgs<T> va;
  int headGx;            // x blocks of the solve's own k_head launch
};

// Argument of a packed launch: the solves it covers, in launch order, and
// the first launch block (or block row) of each: solve d owns blocks
// start[d] .. start[d+1]-1 and runs them as its own launch would run blocks
// 0 .. start[d+1]-start[d]-1.
struct PackMap {
  int n;                         // solves in this launch
  int soff;                      // step offset within the run
  unsigned start[PACK_MAX + 1];
  const void *args[PACK_MAX];    // each solve's PackArgs
  long long base[PACK_MAX];      // Debye launches: first site of the layer
  long long cnt[PACK_MAX];       // Debye: sites in the layer; H corrections: correction blocks
};

template <typename T> struct GhostArgs {
  int NX, NY, NZ, PX;
  long long plane;
  T *F;
  int axis, dst, src;
  T sign;
  int lo[3], hi[3];
};

// ---- context ----

struct Region { int lim[6]; };

// Fast boxes of the two sweeps: where every E (H) component takes the
// ordinary update (the intersection of the inner-grid limits gfilim), kept
// one cell inside the allocation so the fast kernels never test bounds.
// With lw > 0 the x range is cut to whole lines of lw elements (when that
// keeps most of it): then no line is shared by the fast kernel and the
// shell, and the shell's row ends fill their lines.
inline void fast_boxes(const int gfilim[6][6], const int N[3], int fastE[6], int fastH[6], int lw = 0)
{
  for (int e = 0; e < 2; ++e) {
    int *fb = e == 0 ? fastE : fastH;
    for (int a = 0; a < 3; ++a) {
      fb[2 * a] = 1;
      fb[2 * a + 1] = N[a] - 2;
      for (int d = 0; d < 3; ++d) {
        fb[2 * a] = fb[2 * a] > gfilim[3 * e + d][2 * a] ? fb[2 * a] : gfilim[3 * e + d][2 * a];
        fb[2 * a + 1] = fb[2 * a + 1] < gfilim[3 * e + d][2 * a + 1] ? fb[2 * a + 1] : gfilim[3 * e + d][2 * a + 1];
      }
    }
    if (lw > 0) {
      const int x0 = (fb[0] + lw - 1) / lw * lw, x1 = (fb[1] + 1) / lw * lw - 1;
      if (fb[1] - fb[0] + 1 >= 4 * lw && x1 - x0 + 1 >= 2 * lw) {
        fb[0] = x0;
        fb[1] = x1;
      }
    }
  }
}

// Fused box of the leapfrog sweep: where E and H both take the ordinary
// update, its x range cut to whole 63-cell tiles when that leaves at most 16
// columns to the shells.
inline void leap_box(const int gfilim[6][6], const int N[3], int fb[6])
{
  int fe[6], fh[6];
  fast_boxes(gfilim, N, fe, fh, 0);
  for (int a = 0; a < 3; ++a) {
    fb[2 * a] = fe[2 * a] > fh[2 * a] ? fe[2 * a] : fh[2 * a];
    fb[2 * a + 1] = fe[2 * a + 1] < fh[2 * a + 1] ? fe[2 * a + 1] : fh[2 * a + 1];
  }
  const int w = fb[1] - fb[0] + 1;
  if (w >= 63 && w % 63 <= 16) fb[1] = fb[0] + w / 63 * 63 - 1;
}

struct Ctx {
  std::string err;
  vgpu_prec prec;
  int device;
  size_t esz;            // bytes per stored real
  int NX, NY, NZ, PX;
  long long plane;
  size_t ncell;          // PX*NY*NZ
  vgpu_grid_desc grid;
  std::vector<double> xl, yl, zl, dex, dey, dez, idex, idey, idez, idhx, idhy, idhz;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> evE, evH;
  double last_total_ms = 0, last_e_ms = 0, last_h_ms = 0;
  bool finalized = false;
  size_t bytes = 0;

  // device allocations
  std::vector<void *> allocs;
  // the field sets (Ex..Hz each): set `par` holds the current fields; with
  // the fused sweep there are two and every step flips it, otherwise F1 is F
  void *F[6] = {}, *F1[6] = {};
  int par = 0;
  void *fld(int set, int d) const { return set ? F1[d] : F[d]; }
  int next(int set) const { return leap ? set ^ 1 : set; }  // the set a step writes
  void *alpha[3] = {}, *beta[3] = {}, *gamma[3] = {};
  void *ide[3] = {}, *idh[3] = {};
  bool media_set = false;

  // Coefficient classes (the reference's indexed-media representation,
  // grid.h:108-135): every cell's (alpha, beta) triple of E components and
  // gamma triple of H components is one entry of a small table.  Built on
  // the host in vgpu_set_media, refined with the layer-0 Debye block of each
  // E component in vgpu_finalize, uploaded as uint16 per cell when both
  // tables fit.
  std::vector<uint32_t> ecls, hcls;          // host, per device cell
  std::vector<double> etab, htab;            // host: [class][6] alpha/beta, [class][3] gamma (device: padded to 8, 4)
  std::vector<int> edslot;                   // [class][3] layer-0 Debye medium or -1
  bool indexed = false;
  void *d_ecls = nullptr, *d_hcls = nullptr, *d_etab = nullptr, *d_htab = nullptr, *d_edslot = nullptr;
  int nEcls = 0, nHcls = 0;
  // layer-0 Debye sites handled inside the E sweep
  bool debyeFused = false;
  long long debyeL0 = 0;
  // fused leapfrog sweep: the fast boxes are its fused box, H cells to
  // update again after the E corrections
  bool leap = false;
  int leapR = 4, leapKZ = 8;
  long long nFix = 0, fixCells = 0;
  void *d_fix = nullptr;
  std::vector<long long> eTouched;           // E values (cell * 3 + component) changed after the E sweep
  int NCH = 0;                               // 64-cell x chunks of the fast E box
  int fastE[6] = {}, fastH[6] = {};          // boxes of the sweeps' fast kernels
  void *d_eargs[2] = {}, *d_hargs[2] = {};   // argument blocks in device memory, by the set read
  void *d_epcs = nullptr, *d_ework = nullptr;   // E shell: pieces, launch blocks
  void *d_hpcs = nullptr, *d_hwork = nullptr;   // H shell
  int nEwork = 0, nHwork = 0, nEpcs = 0, nHpcs = 0;
  // the same shells at the most planes per block, for the packed launches of
  // group runs, which hold many grids' blocks
  void *d_epcsPk = nullptr, *d_eworkPk = nullptr, *d_hpcsPk = nullptr, *d_hworkPk = nullptr;
  int nEworkPk = 0, nHworkPk = 0;
  long long shellCells[2] = {};              // cells of the E and H shell pieces
  void *d_dbase = nullptr;                   // [3][NZ][NY][NCH] + 1: first layer-0 site

  // UPML
  unsigned pmlmask = 0;
  int fplim[6][6][6];
  void *aux1[6][3] = {}, *aux2[6][3] = {}, *auxB[6][3] = {};
  // UPML auxiliaries live in three pools (E: PP and P with the same layout,
  // H: B); aux1/aux2/auxB point into them, each slab 256-byte aligned.
  void *auxPool[3] = {};
  void *prof[6][3] = {};   // ad, bd, ah, bh, ibd, ibh per axis

  // Mur
  unsigned murmask = 0;
  double zeta[6] = {};
  int fmlim[6][6][6];

  // ghosts
  int ghost[6] = {};

  // Debye blocks (host side until finalize)
  struct DebyeBlock { int flim[3][6]; int np; std::vector<double> da, db; };
  std::vector<DebyeBlock> debye;
  int debyeLayers = 0;
  std::vector<long long> debyeLayerStart;
  long long debyeSites = 0;
  int debyeMaxPoles = 0, debyeCplx = 0;
  void *d_dkey = nullptr, *d_dsbeta = nullptr, *d_dlast = nullptr, *d_dJ = nullptr;
  void *d_dalpha = nullptr, *d_dbeta = nullptr, *d_dnp = nullptr;

  // SIBC
  struct SibcSurf { vgpu_sibc_desc d; std::vector<unsigned char> adjA, adjB;
                    std::vector<double> q0[16], q1[16], q2[16]; };
  std::vector<SibcSurf> sibc;
  long long sibcFaces = 0;
  int sibcMaxPoles = 1;
  bool sibcReal = false;                     // all filter coefficients real: real filter state
  bool sibcFactored = false;                 // FP32 form: one state per (input, pole)
  void *d_sfbase = nullptr, *d_sfratio = nullptr, *d_sblocks = nullptr;
  int sibcBlocks = 0;
  void *d_sijk = nullptr, *d_ssurf = nullptr, *d_sadjA = nullptr, *d_sadjB = nullptr, *d_snormal = nullptr;
  void *d_sA = nullptr, *d_sB = nullptr, *d_sset = nullptr, *d_snp = nullptr, *d_sasymp = nullptr, *d_sq = nullptr;
  void *d_szeta = nullptr, *d_sold = nullptr, *d_setan = nullptr;
  long long sibcHRows = 0;
  void *d_shkey = nullptr, *d_shstart = nullptr, *d_shent = nullptr, *d_shg = nullptr;

  // waveforms, sources, plane waves
  std::vector<WaveformDev> wf;
  std::vector<std::vector<double>> wftab;
  struct Src { int field, kind, soft, waveform; double size, delay; int flim[6]; };
  std::vector<Src> src;
  struct Pw { vgpu_pw_desc d; double prof[4][10]; };
  std::vector<Pw> pw;
  void *d_wf = nullptr, *d_src = nullptr, *d_pw = nullptr, *d_aux = nullptr, *d_auxpml = nullptr, *d_auxprof = nullptr;
  int auxmax = 0;
  long long corrRows[2] = {0, 0};
  void *d_ckey[2] = {}, *d_cstart[2] = {}, *d_cent[2] = {}, *d_ccoef[2] = {}, *d_calpha[2] = {};

  // observers
  struct PObs { int i, j, k; uint64_t first, last; std::vector<double> omega; };
  std::vector<PObs> pobs;
  struct VObs { int box[6]; int step[3]; uint64_t first, last; long long nodes; };
  std::vector<VObs> vobs;
  uint64_t totalSteps = 0, maxRun = 0, lastRunSteps = 0, lastRunFirst = 0;
  // device step counter (read by the time-dependent kernels) and the host's
  // view of its value once the enqueued work has run
  void *d_step = nullptr;
  void *d_pack = nullptr;  // PackArgs for group runs (separate sweeps only)
  uint64_t stepHost = 0;
  // a captured graph of graphSteps whole steps (replayed by vgpu_run)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  uint64_t graphSteps = 0;
  int graphPar = 0;                          // field set current when the graph was captured
  uint64_t graphRuns = 0;                    // runs replayed from it
  bool graphTried = false, lastRunTimed = false;
  // group runs led by this context: the last group's members (context
  // serials, never reused) and length, and a graph captured once the same
  // group runs twice in a row
  uint64_t serial = 0;
  std::vector<uint64_t> groupKey;
  uint64_t groupSteps = 0;
  hipGraph_t groupGraph = nullptr;
  hipGraphExec_t groupExec = nullptr;
  bool groupTried = false;
  // replaced group graphs, freed once the event recorded after their last
  // launch has completed
  struct RetiredGraph { hipGraph_t graph; hipGraphExec_t exec; hipEvent_t done; };
  std::vector<RetiredGraph> retired;
  void *d_ocell = nullptr, *d_ofirst = nullptr, *d_olast = nullptr, *d_onf = nullptr, *d_ooff = nullptr;
  void *d_oomega = nullptr, *d_otd = nullptr, *d_odre = nullptr, *d_odim = nullptr;
  long long vnodes = 0;
  void *d_vcell = nullptr, *d_vobs = nullptr, *d_vfirst = nullptr, *d_vlast = nullptr, *d_vstage = nullptr;
  std::vector<long long> vobsNodeStart;
  // Volume samples stream out without a sync per run: the device staging
  // holds 2*maxRun rows (row = step modulo that), each run's rows are copied
  // asynchronously into its slot of a pinned host ring after its kernels,
  // and a record of each of the last two runs (first step, length, start and
  // end events) lets the host finish one run while the next one runs.
  uint64_t vring = 0;
  float *h_vring = nullptr;  // [2][maxRun][vnodes][6], pinned
  struct RunRec { uint64_t first = 0, n = 0; bool valid = false; hipEvent_t start = nullptr, end = nullptr; };
  RunRec runs[2];
  uint64_t runCount = 0;

  long long lin(int i, int j, int k) const { return ((long long)k * NY + j) * PX + i; }
};

int set_error(Ctx *c, int code, const char *fmt, ...);
void *dev_alloc(Ctx *c, size_t bytes);

// Update-once check: visit-count arrays per component (device layout) and
// sweep argument blocks in device memory that carry them.
struct VisitRun {
  unsigned *vis[6];
  void *eargs, *hargs;
};

// Per-precision step driver, instantiated for double and float.
// Enqueues nsteps steps that start at the device step counter, then advances
// it; with timeKernels the volume sweeps of the first steps are timed.  With
// chk it enqueues one step of the field sweeps only, counting updates.
template <typename T>
int run_steps(Ctx *c, uint64_t nsteps, hipStream_t s, bool timeKernels, const VisitRun *chk = nullptr);
// Uploads the sweeps' argument blocks (once, at finalize).
template <typename T> int prepare_sweeps(Ctx *c);
// Uploads the context's argument blocks for packed launches (once, at finalize).
template <typename T> int upload_pack(Ctx *c);
// Enqueues nsteps steps of n contexts in lockstep on stream s, with packed
// launches, then advances each one's step counter.
template <typename T> int run_group_steps(Ctx *const *cs, int n, uint64_t nsteps, hipStream_t s);
// Update-once check of the field sweeps against the outer field limits.  Its
// buffers come from `scratch` (scratchBytes long) when given, else they are
// allocated for the call.
template <typename T>
int check_limits(Ctx *c, const int folim[6][6], vgpu_limits_report *rep, unsigned char *counts, void *scratch,
                 size_t scratchBytes);
template <typename T> size_t check_limits_bytes(const Ctx *c, bool counts, size_t off[4]);

}  // namespace vgpu
