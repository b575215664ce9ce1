/*
 * vgvulture: Vulture mesh in, Vulture observer files out, time loop on one
 * MI355X.
 *
 *   vgvulture [-f] [-p fp64|fp32] [-n steps] [-c chunk] [-o outdir] mesh
 *   vgvulture -L [-p fp64|fp32] mesh
 *
 * -L runs the update-once check of the field sweeps instead of the loop
 * (the reference's limit-checking build, grid.h:181-191 and checkGrid,
 * grid.c:951-1025): it warns for every cell of the outer field limits that
 * no sweep updates, as the reference does, and reports cells updated twice
 * or outside the limits; the exit status is 6 when any is found.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vh_gpu.h"

static const char *FIELD[6] = {"EX", "EY", "EZ", "HX", "HY", "HZ"};

static int check_limits(const char *mesh, int prec)
{
  vh_solve *s = NULL;
  vgpu_ctx *c = NULL;
  char err[1024] = "";
  if (vh_gpu_open(mesh, prec, 0, 1, 1, &s, &c, err, sizeof(err))) {
    fprintf(stderr, "vgvulture: %s\n", err);
    return 6;
  }
  int lim[6][6];
  long long n[3];
  vh_solve_field_limits(s, lim);
  vh_solve_shape(s, n);
  const size_t plane = (size_t)n[0] * n[1], cells = plane * n[2];
  unsigned char *counts = malloc(6 * cells);
  vgpu_limits_report rep;
  int rc = counts ? vgpu_check_limits(c, (const int(*)[6])lim, &rep, counts) : -1;
  if (rc) {
    fprintf(stderr, "vgvulture: update-once check failed: %s\n", counts ? vgpu_last_error(c) : "out of memory");
    free(counts);
    vh_gpu_close(s, c);
    return 6;
  }
  unsigned long long missed = 0, bad = 0;
  for (int f = 0; f < 6; f++) {
    const int *b = lim[f];
    for (int i = b[0]; i <= b[1]; i++)
      for (int j = b[2]; j <= b[3]; j++)
        for (int k = b[4]; k <= b[5]; k++)
          if (counts[f * cells + (size_t)k * plane + (size_t)j * n[0] + i] == 0)
            printf("*** Warning: %s[%d][%d][%d] not updated\n", FIELD[f], i, j, k);
    missed += rep.missed[f];
    bad += rep.repeated[f] + rep.outside[f];
    printf("%s: missed %llu repeated %llu outside %llu\n", FIELD[f], (unsigned long long)rep.missed[f],
           (unsigned long long)rep.repeated[f], (unsigned long long)rep.outside[f]);
  }
  free(counts);
  vh_gpu_close(s, c);
  if (missed) printf("** Found %llu field elements that have not been updated! **\n", missed);
  if (bad) printf("** Found %llu field updates repeated or outside the field limits! **\n", bad);
  return missed || bad ? 6 : 0;
}

int main(int argc, char **argv)
{
  int full = 0, prec = 0, limits = 0;
  long steps = -1;
  unsigned long chunk = 0;
  const char *out = NULL, *mesh = NULL;
  for (int a = 1; a < argc; a++) {
    if (!strcmp(argv[a], "-f")) full = 1;
    else if (!strcmp(argv[a], "-L")) limits = 1;
    else if (!strcmp(argv[a], "-p") && a + 1 < argc) prec = !strcmp(argv[++a], "fp32");
    else if (!strcmp(argv[a], "-n") && a + 1 < argc) steps = atol(argv[++a]);
    else if (!strcmp(argv[a], "-c") && a + 1 < argc) chunk = strtoul(argv[++a], NULL, 10);
    else if (!strcmp(argv[a], "-o") && a + 1 < argc) out = argv[++a];
    else mesh = argv[a];
  }
  if (!mesh) {
    fprintf(stderr, "usage: vgvulture [-f] [-p fp64|fp32] [-n steps] [-c chunk] [-o outdir] mesh\n"
                    "       vgvulture -L [-p fp64|fp32] mesh\n");
    return 2;
  }
  if (limits) return check_limits(mesh, prec);
  char err[1024] = "";
  double ms = 0;
  if (vh_run_mesh_gpu(mesh, out, prec, 0, full, steps, chunk, &ms, err, sizeof(err))) {
    fprintf(stderr, "vgvulture: %s\n", err);
    return 6;
  }
  printf("loop_ms %.3f\n", ms);
  return 0;
}
