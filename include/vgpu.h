/*
 * vgpu.h - C-ABI of the MI355X FDTD time-stepping core.
 *
 * This is the drop-in boundary for the reference solver's update loop,
 * propagate() (src/simulation.c:136-230).  A host that has done the
 * reference's set-up (parser, grid, media, UPML profiles, Debye blocks, SIBC
 * surfaces, plane waves, sources, observers) hands the resulting arrays to
 * one vgpu context and calls vgpu_run() in place of the time loop.  Every
 * array crosses in the reference's own layout: 3-D arrays are k-fastest
 * A[i][j][k] over the allocation extents nx*ny*nz (gobox + ghost cells,
 * src/grid.c:357-387), 1-D arrays are per-line over the same extents.  The
 * library transposes to its x-fastest device layout on upload.
 *
 * All entry points return 0 on success and a negative code on error;
 * vgpu_last_error() gives the message (the reference reports the same
 * conditions through message(MSG_ERROR) and exit(6), src/message.c:70-108).
 * A context is one solve; contexts are independent and may be used from
 * different host threads, one thread per context at a time.
 *
 * Plain C: no HIP, torch or C++ types appear here.  Streams are passed as
 * void* (a hipStream_t, or NULL for the context's own stream).
 */
#ifndef VGPU_H
#define VGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vgpu_ctx vgpu_ctx;

/* Storage precision of fields, coefficients, UPML and Debye state.  SIBC
 * recursive-convolution state and DFT accumulators are FP64 in both. */
typedef enum { VGPU_FP64 = 0, VGPU_FP32 = 1 } vgpu_prec;

enum { VGPU_OK = 0, VGPU_EINVAL = -1, VGPU_ENOMEM = -2, VGPU_EDEVICE = -3, VGPU_ESTATE = -4 };

/* Grid description: replaces the globals of src/grid.c:55-126. */
typedef struct {
  int nx, ny, nz;            /* allocation extents incl. ghost layers (numCells) */
  int gibox[6], gobox[6];    /* inner (mesh) and outer (mesh + PML) boxes */
  int gfilim[6][6];          /* inner-grid field limits [component][face] (Mur faces excluded) */
  double dt;                 /* time step */
  double c0, eps0, mu0, eta0;/* physical constants as the host computed them */
  int grid_type;             /* 0 cubic, 1 uniform, 2 non-uniform (src/grid.c:1549-1590) */
  const double *xlines, *ylines, *zlines;  /* mesh line coordinates (numLines each) */
  int nlines[3];
  const double *dex, *dey, *dez;           /* primary edge lengths [nx], [ny], [nz] */
  const double *idex, *idey, *idez;        /* inverse primary edge lengths */
  const double *idhx, *idhy, *idhz;        /* inverse secondary edge lengths */
} vgpu_grid_desc;

/* Create a context for one solve (replaces the allocation half of initGrid,
 * src/grid.c:327-355, and clearGrid).  Fields start at zero. */
int vgpu_create(const vgpu_grid_desc *grid, vgpu_prec prec, int device, vgpu_ctx **out);

/* Update coefficients alpha/beta (E) and gamma (H), k-fastest [nx][ny][nz]
 * (src/grid.c:97-105 filled by setMediumOnGrid, src/grid.c:1726-1873). */
int vgpu_set_media(vgpu_ctx *c, const double *alpha[3], const double *beta[3], const double *gamma[3]);
/* The same from float arrays: the reference's own build keeps the coefficient
 * globals in `real` = float (src/fdtd_types.h:36-38), and this takes them as
 * they are, with no double copy on the host (each value is widened exactly). */
int vgpu_set_media_f32(vgpu_ctx *c, const float *alpha[3], const float *beta[3], const float *gamma[3]);

/* UPML slabs (src/pml.c:44-84, updates at 859-1176).  region_mask bit r set
 * when face r is PML; fplim the slab field limits; per-axis profile arrays
 * a_d, b_d, a_h, b_h, 1/b_d (ib_d), 1/b_h (ib_h) over the allocation extents. */
typedef struct {
  unsigned region_mask;
  int fplim[6][6][6];
  const double *ad[3], *bd[3], *ah[3], *bh[3], *ibd[3], *ibh[3];
} vgpu_pml_desc;
int vgpu_set_pml(vgpu_ctx *c, const vgpu_pml_desc *pml);

/* First-order Mur faces (src/mur.c:41-131): face_mask bit f set when face f
 * is Mur, zeta the per-face factors, fmlim the face field limits. */
int vgpu_set_mur(vgpu_ctx *c, unsigned face_mask, const double zeta[6], const int fmlim[6][6][6]);

/* Ghost-cell update per face (src/surface.c:569-812): 0 none (PML/Mur),
 * 1 PEC, 2 PMC, 3 periodic. */
int vgpu_set_ghosts(vgpu_ctx *c, const int face_type[6]);

/* One Debye block (src/debye.c:44-55, 76-196), in block order.  flim is the
 * E field limits [EX..EZ][face]; dalpha/dbeta are npoles complex values as
 * interleaved (re, im) doubles (src/medium.c:482-504). */
int vgpu_add_debye_block(vgpu_ctx *c, const int flim[3][6], int npoles, const double *dalpha,
                         const double *dbeta);

/* One SIBC surface (src/sibc.c:44-56, 211-397), in surface order.  gbbox is
 * the grid face box with +1 along the normal; A/B the orientation matrices;
 * is_adj_a/b the [face][4] adjacency flags in i,j,k face order; pmc_index[d]
 * the two indices along axis d whose edges lie on a PMC/periodic face (-1
 * for none, src/surface.c:1022-1052); filters the 4x4 recursive-convolution
 * matrix row-major (src/filter.c:347-368) with q0,q1,q2 as interleaved
 * complex [npoles]. */
typedef struct {
  int gbbox[6];
  int normal;
  double A[4][4], B[4][4];
  const unsigned char *is_adj_a, *is_adj_b;
  int pmc_index[3][2];
  int npoles[16];
  double asymp[16];
  const double *q0[16], *q1[16], *q2[16];
} vgpu_sibc_desc;
int vgpu_add_sibc_surface(vgpu_ctx *c, const vgpu_sibc_desc *sibc);

/* Waveforms (src/waveform.c:256-347, 413-638) in waveform-number order;
 * type as in src/waveform.h:43-57; table is [n][3] (time, value, spline
 * second derivative) for the external type. */
int vgpu_add_waveform(vgpu_ctx *c, int type, double size, double delay, double width, double frequency,
                      const double *table, unsigned long n);

/* One source after initSources' conversions (src/source.c:469-599):
 * field 0..5, kind 0 = field (E/H += or =), 1 = current density (-beta*s or
 * -gamma*s); flim the field limits of that component. */
int vgpu_add_source(vgpu_ctx *c, int field, int kind, int soft, double size, int waveform, double delay,
                    const int flim[6]);

/* One TF/SF plane wave (src/planewave.c:54-92, 277-554, 1226-1296). */
typedef struct {
  int flim[6][6][6];         /* [face][component][face] correction limits */
  unsigned char active[6];
  double kinc[3], Finc[6], ijk0[3], r0[3];
  double phase_velocity;     /* analytic incident field (non-cubic grids) */
  int waveform;
  double delay;
  int use_aux;               /* 1-D auxiliary grid (cubic grids) */
  int aux_nx, aux_xb;
  double beta_eyi, gamma_hzi;
  const double *aux_adx, *aux_bdx, *aux_ahx, *aux_bhx;  /* [10] */
} vgpu_pw_desc;
int vgpu_add_planewave(vgpu_ctx *c, const vgpu_pw_desc *pw);

/* Point observer at grid node (i,j,k), sampled before each step's updates
 * for steps first..last (src/observer.c:622-650).  With nfreq > 0 a running
 * FP64 DFT over omega[nfreq] is kept (src/observer.c:857-899).  Returns the
 * observer id (>= 0) or a negative error. */
int vgpu_add_point_observer(vgpu_ctx *c, int i, int j, int k, uint64_t first, uint64_t last, int nfreq,
                            const double *omega);

/* Volume observer over grid box with node steps (TDOM_BINARY,
 * src/observer.c:1164-1198).  Returns the observer id. */
int vgpu_add_volume_observer(vgpu_ctx *c, const int gbbox[6], const int step[3], uint64_t first, uint64_t last);

/* Build device-side structures after all set-up calls.  max_steps_per_run
 * sizes the volume-observer staging buffers. */
int vgpu_finalize(vgpu_ctx *c, uint64_t total_steps, uint64_t max_steps_per_run);

/* Advance steps [first, first+nsteps) (replaces the body of propagate(),
 * src/simulation.c:157-222).  Asynchronous on the stream; returns after
 * enqueueing.  stream NULL uses the context's stream. */
int vgpu_run(vgpu_ctx *c, uint64_t first, uint64_t nsteps, void *stream);
int vgpu_synchronize(vgpu_ctx *c);

/* Advance up to VGPU_GROUP_MAX solves by nsteps each in lockstep, context i
 * over steps [first[i], first[i]+nsteps): one packed launch per step kernel
 * over all of them (fast and shell sweeps, Debye layers, corrections,
 * observers) instead of one launch per solve, for grids too small to fill
 * the device alone.  Each context ends as after vgpu_run(ctxs[i], first[i],
 * nsteps, NULL), bit for bit, with its own run record (vgpu_run_result).
 * Runs on the first context's stream after the work queued on every
 * member's stream; each member's stream waits for the group's end.  Every
 * member must use the separate sweeps (not VGPU_LEAP, not a fused Debye
 * layer) and share device and precision.  No counterpart in the reference,
 * which runs one solve per process (src/simulation.c:136-230); it serves
 * many small solves on one GPU. */
#define VGPU_GROUP_MAX 16
int vgpu_run_group(vgpu_ctx *const *ctxs, int n, const uint64_t *first, uint64_t nsteps);

/* Non-blocking completion test of the last vgpu_run: 1 done, 0 still
 * running, <0 device error.  No counterpart in the reference, whose loop is
 * synchronous; it lets one host thread keep many solves in flight. */
int vgpu_query(vgpu_ctx *c);

/* Point observer samples: td[n][6] for every step sampled so far (FP64),
 * dft_re/dft_im[6][nfreq] (either may be NULL). */
int vgpu_read_point_observer(vgpu_ctx *c, int id, double *td, uint64_t nsteps, double *dft_re, double *dft_im);

/* Volume observer samples of the last vgpu_run: [step][node][6] floats in
 * the k, j, i node order of impulse.dat.  Returns the node count via *nodes. */
int vgpu_read_volume_observer(vgpu_ctx *c, int id, float *out, uint64_t nsteps, uint64_t *nodes);

/* Completion of one of the last two runs, without waiting for later ones.
 * The reference writes volume samples to impulse.dat inside its time loop
 * (src/observer.c:1164-1198); here each vgpu_run copies its samples into
 * pinned host memory asynchronously, after its kernels on its stream, so a
 * host can write them out while the next run proceeds.  For the run that
 * began at step `first`: wait 1 blocks until it has ended, wait 0 does not.
 * Returns 1 when it has ended (then *ms is its device time and *samples its
 * volume observer samples, [step][node][6] floats over every volume observer
 * in id order and the k, j, i node order of impulse.dat, or NULL without
 * volume observers; valid until the second vgpu_run after that run is
 * called), 0 when it has not (wait 0), <0 on error or when `first` names
 * neither of the last two runs.  Any of ms, samples, nsteps may be NULL. */
int vgpu_run_result(vgpu_ctx *c, uint64_t first, int wait, double *ms, const float **samples, uint64_t *nsteps);

/* Copy one field component (0..5) back in the host k-fastest layout. */
int vgpu_download_field(vgpu_ctx *c, int component, double *host);

/* Device bytes held by the context after vgpu_finalize. */
size_t vgpu_footprint_bytes(const vgpu_ctx *c);

/* Work counts of a finalized context, for throughput and roofline
 * accounting (no counterpart in the reference, whose timer reports only
 * seconds per iteration per cell, timer.c:95-99). */
typedef struct {
  uint64_t cells;              /* gobox cells (mesh + UPML) updated per step */
  uint64_t mesh_cells;         /* gibox cells */
  uint64_t debye_sites;        /* electric edges carrying a Debye correction */
  int debye_poles;             /* poles per Debye site (maximum) */
  int debye_complex;           /* 1 if any pole is complex */
  uint64_t sibc_faces;         /* SIBC surface faces */
  uint64_t correction_rows[2]; /* source / TF/SF correction rows, E and H */
  uint64_t bytes;              /* device footprint */
  uint64_t fast_cells[2];      /* cells of the E and H fast sweep kernels */
  uint64_t debye_fused_sites;  /* Debye sites applied inside the fast E kernel */
  int indexed_media;           /* 1: coefficients read as 16-bit classes */
  int debye_fused;             /* 1: first Debye layer fused into the E sweep */
  int sibc_state;              /* SIBC filter state: 0 complex per filter, 1 real per filter, 2 real per input */
  int leap;                    /* 1: fused leapfrog sweep (E then H in one pass over the fast box) */
  uint64_t fix_cells;          /* H cells of the fused box updated again after the E corrections */
  uint64_t fix_chunks;         /* the 64-cell row chunks holding them */
  uint64_t graph_runs;         /* runs replayed from a captured graph (on unless VGPU_GRAPHS=0) */
} vgpu_stats;
int vgpu_get_stats(const vgpu_ctx *c, vgpu_stats *out);

/* Time of the kernels of the last vgpu_run, from events on its stream (ms),
 * and the average duration of the E and H volume kernels in it (ms). */
int vgpu_last_run_timing(const vgpu_ctx *c, double *total_ms, double *volume_e_ms, double *volume_h_ms);

/* Update-once check of the field sweeps.  Replaces the reference's
 * limit-checking build (CMakeLists.txt:40,89-95: fields start at -1, each
 * update asserts the old value and marks the cell, grid.h:181-191, and
 * checkGrid, grid.c:951-1025, reports fields of the outer limits left
 * unmarked after the run).  Here one step of the field sweeps (Mur E, the E
 * sweep over inner grid and UPML, the H sweep over inner grid, Mur normal H
 * and UPML) runs in a counting form that adds one to a per-cell count for
 * each update instead of storing it, and each component's counts are
 * tallied against folim[component][face] (the outer field limits, gfolim of
 * grid.c:404).  Fields, UPML auxiliaries and Debye state are unchanged.
 * counts, when not NULL, receives the visits per cell (saturated at 255) as
 * 6 * nx * ny * nz bytes, index ((component * nz + k) * ny + j) * nx + i.
 * Synchronizes the device; call it between runs. */
typedef struct {
  uint64_t missed[6];          /* cells inside folim that no sweep updated */
  uint64_t repeated[6];        /* cells updated more than once */
  uint64_t outside[6];         /* cells outside folim that a sweep updated */
  int64_t first[6];            /* lowest (k * ny + j) * nx + i of an offending cell, -1 if none */
} vgpu_limits_report;
int vgpu_check_limits(vgpu_ctx *c, const int folim[6][6], vgpu_limits_report *out, unsigned char *counts);

/* The update-once check without allocating device memory.  No counterpart in
 * the reference: it serves callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and synchronization as vgpu_check_limits; its buffers (the visit
 * counts, 4 bytes per component and device cell, and the staging of
 * `counts`) come from a fixed scratch area of vgpu_noalloc_scratch_bytes()
 * bytes that the library holds on each device it runs on until the process
 * ends.  A grid whose check needs more returns VGPU_EINVAL and says how much
 * in vgpu_last_error.  All calls on one device share its scratch: calls that
 * overlap on one device race, so serialize them. */
#define VGPU_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int vgpu_check_limits_noalloc(vgpu_ctx *c, const int folim[6][6], vgpu_limits_report *out, unsigned char *counts);
size_t vgpu_noalloc_scratch_bytes(void);

const char *vgpu_last_error(const vgpu_ctx *c);
void vgpu_destroy(vgpu_ctx *c);

#ifdef __cplusplus
}
#endif

#endif
