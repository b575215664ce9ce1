/*
 * vgvulture: Vulture mesh in, Vulture observer files out, time loop on one
 * MI355X.
 *
 *   vgvulture [-f] [-p fp64|fp32] [-n steps] [-c chunk] [-o outdir] mesh
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int vh_run_mesh_gpu(const char *mesh, const char *outdir, int prec, int device, int fullPrecision, long steps,
                    unsigned long chunk, double *loop_ms, char *err, int errlen);

int main(int argc, char **argv)
{
  int full = 0, prec = 0;
  long steps = -1;
  unsigned long chunk = 0;
  const char *out = NULL, *mesh = NULL;
  for (int a = 1; a < argc; a++) {
    if (!strcmp(argv[a], "-f")) full = 1;
    else if (!strcmp(argv[a], "-p") && a + 1 < argc) prec = !strcmp(argv[++a], "fp32");
    else if (!strcmp(argv[a], "-n") && a + 1 < argc) steps = atol(argv[++a]);
    else if (!strcmp(argv[a], "-c") && a + 1 < argc) chunk = strtoul(argv[++a], NULL, 10);
    else if (!strcmp(argv[a], "-o") && a + 1 < argc) out = argv[++a];
    else mesh = argv[a];
  }
  if (!mesh) {
    fprintf(stderr, "usage: vgvulture [-f] [-p fp64|fp32] [-n steps] [-c chunk] [-o outdir] mesh\n");
    return 2;
  }
  char err[1024] = "";
  double ms = 0;
  if (vh_run_mesh_gpu(mesh, out, prec, 0, full, steps, chunk, &ms, err, sizeof(err))) {
    fprintf(stderr, "vgvulture: %s\n", err);
    return 6;
  }
  printf("loop_ms %.3f\n", ms);
  return 0;
}
