/*
 * vh_gpu.h - the host's device driver and solve server.
 *
 * The device driver hands a set-up vh_solve to the vgpu C-ABI and runs its
 * time loop (the drop-in for propagate(), simulation.c:136-230); the server
 * keeps many such solves resident and in flight on one GPU at once.
 */
#ifndef VH_GPU_H
#define VH_GPU_H

#include "vh.h"
#include "../../include/vgpu.h"

/* ---- device driver (vh_gpu.c) ---- */
vgpu_ctx *vh_gpu_setup(vh_solve *s, int prec, int device, unsigned long chunk);
int vh_gpu_run(vh_solve *s, int prec, int device, unsigned long chunk, double *loop_ms);
int vh_run_mesh_gpu(const char *mesh, const char *outdir, int prec, int device, int fullPrecision, long steps,
                    unsigned long chunk, double *loop_ms, char *err, int errlen);

/* Several small solves run together, one packed launch per step kernel
 * over all of them (vgpu_run_group); at most VGPU_GROUP_MAX. */
int vh_gpu_run_group(vh_solve **ss, int n, int prec, int device, unsigned long chunk, double *loop_ms);
int vh_run_meshes_gpu_group(const char *const *meshes, const char *const *outdirs, int n, int prec, int device,
                            int fullPrecision, unsigned long chunk, double *loop_ms, char *err, int errlen);

typedef struct vh_dsolve vh_dsolve;
vh_dsolve *vh_dsolve_open(vh_solve *s, int prec, int device, unsigned long chunk);
long vh_dsolve_enqueue(vh_dsolve *d);
long vh_dsolve_enqueue_group(vh_dsolve **ds, int n);
int vh_dsolve_ready(vh_dsolve *d);
int vh_dsolve_complete(vh_dsolve *d);
unsigned long vh_dsolve_remaining(const vh_dsolve *d);
int vh_dsolve_done(const vh_dsolve *d);
/* 1 when the solve can step in a group (vh_dsolve_enqueue_group): its
 * context uses the separate sweeps, not the fused leapfrog or Debye forms. */
int vh_dsolve_packable(const vh_dsolve *d);
int vh_dsolve_collect(vh_dsolve *d, double ***td);
void vh_dsolve_write(vh_dsolve *d, double **td);
double vh_dsolve_loop_ms(const vh_dsolve *d);
vgpu_ctx *vh_dsolve_context(vh_dsolve *d);
void vh_dsolve_close(vh_dsolve *d);

/* Parse, set up and upload a mesh without running it; vh_gpu_close frees both. */
int vh_gpu_open(const char *mesh, int prec, int device, long steps, unsigned long chunk, vh_solve **solve,
                vgpu_ctx **ctx, char *err, int errlen);
void vh_gpu_close(vh_solve *s, vgpu_ctx *c);

/* out[0] gobox cells, out[1] mesh cells, out[2] time steps; numCells. */
void vh_solve_sizes(const vh_solve *s, long long out[3]);
void vh_solve_shape(const vh_solve *s, long long out[3]);
void vh_solve_field_limits(const vh_solve *s, int out[6][6]);

#endif
