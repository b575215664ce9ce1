/*
 * Driver for the CPU restatement (test infrastructure only): parse a mesh,
 * set up, run the time loop and write the reference's observer files.
 *
 *   cpu_ref [-f] [-n steps] [-w warmup] [-t threads] [-o outdir] mesh
 *
 * -f writes observer data with all 17 significant digits.  -w leaves the
 * first `warmup` steps out of the timed loop (they still run).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "vh.h"

void cpu_ref_propagate(vh_solve *s, unsigned long first, unsigned long count);

/* Run a whole mesh; returns 0 on success.  `seconds` receives the wall time
 * of the time loop alone. */
int cpu_ref_run(const char *mesh, const char *outdir, int fullPrecision, long steps, long warmup, int threads,
                double *seconds, double *cells, char *err, int errlen)
{
  vh_solve *s = vh_create();
  if (!s) return -1;
  vh_set_dirs(s, NULL, outdir);
  s->fullPrecision = fullPrecision;
  if (vh_read_mesh(s, mesh) || (steps >= 0 && (s->numTimeSteps = (unsigned long)steps, 0)) || vh_init(s) ||
      vh_open_observers(s)) {
    if (err) snprintf(err, (size_t)errlen, "%s", vh_error(s));
    vh_destroy(s);
    return -1;
  }
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  unsigned long w = warmup > 0 ? (unsigned long)warmup : 0;
  if (w > s->numTimeSteps) w = s->numTimeSteps;
  cpu_ref_propagate(s, 0, w);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  cpu_ref_propagate(s, w, s->numTimeSteps - w);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (cells)
    *cells = (double)(s->gobox[XHI] - s->gobox[XLO]) * (double)(s->gobox[YHI] - s->gobox[YLO]) *
             (double)(s->gobox[ZHI] - s->gobox[ZLO]);
  if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  vh_destroy(s);
  return 0;
}

#ifndef CPU_REF_NO_MAIN
int main(int argc, char **argv)
{
  int full = 0, threads = 0;
  long steps = -1, warmup = 0;
  const char *out = NULL, *mesh = NULL;
  for (int a = 1; a < argc; a++) {
    if (!strcmp(argv[a], "-f")) full = 1;
    else if (!strcmp(argv[a], "-n") && a + 1 < argc) steps = atol(argv[++a]);
    else if (!strcmp(argv[a], "-w") && a + 1 < argc) warmup = atol(argv[++a]);
    else if (!strcmp(argv[a], "-t") && a + 1 < argc) threads = atoi(argv[++a]);
    else if (!strcmp(argv[a], "-o") && a + 1 < argc) out = argv[++a];
    else mesh = argv[a];
  }
  if (!mesh) {
    fprintf(stderr, "usage: cpu_ref [-f] [-n steps] [-w warmup] [-t threads] [-o outdir] mesh\n");
    return 2;
  }
  char err[1024] = "";
  double sec = 0.0, cells = 0.0;
  if (cpu_ref_run(mesh, out, full, steps, warmup, threads, &sec, &cells, err, sizeof(err))) {
    fprintf(stderr, "cpu_ref: %s\n", err);
    return 6;
  }
  printf("loop_seconds %.6f\ncells %.0f\n", sec, cells);
  return 0;
}
#endif
