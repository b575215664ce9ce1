/*
 * vh_serve.h - solve server: many solves resident and in flight on one GPU
 * (no counterpart in the reference, which runs one solve per process).
 */
#ifndef VH_SERVE_H
#define VH_SERVE_H

#include <stddef.h>

typedef struct vh_server vh_server;

/* Times are seconds on the server's monotonic clock (vh_server_now). */
typedef struct {
  int id;
  int status;               /* 0 pending, 1 done, -1 failed */
  double t_submit;          /* request accepted */
  double t_ready;           /* host set-up finished */
  double t_admit;           /* uploaded, first chunk may run */
  double t_gpu_done;        /* last step finished on the device */
  double t_done;            /* observer files written */
  double gpu_ms;            /* device time of its own kernels (events on its stream) */
  long long cells;          /* gobox cells */
  unsigned long steps;
  size_t bytes;             /* device footprint */
  char err[256];
} vh_request_info;

/* device: GPU ordinal; prec: VGPU_FP64/VGPU_FP32; budgetBytes: device memory
 * the server may hold for solves; maxActive: solves in flight at once;
 * workers: host set-up / output threads; chunk: steps per enqueue. */
vh_server *vh_server_create(int device, int prec, int fullPrecision, double budgetBytes, int maxActive, int workers,
                            unsigned long chunk);
/* Queue a mesh (observer files go to outdir).  Returns the request id. */
int vh_server_submit(vh_server *sv, const char *mesh, const char *outdir);
/* Block until request id has finished; 0 done, -1 failed. */
int vh_server_wait(vh_server *sv, int id, vh_request_info *info);
/* Wait for every request submitted so far; returns the number that failed. */
int vh_server_drain(vh_server *sv);
double vh_server_now(vh_server *sv);
/* Drain, stop the threads and free the server. */
void vh_server_destroy(vh_server *sv);

#endif
