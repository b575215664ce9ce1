/* Drives the solve server over the meshes named on the command line (test
 * infrastructure): submit all (SERVE_DRIVER_GAP_MS apart), wait, print one
 * line per request. */
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>
#include "vh_serve.h"

int main(int argc, char **argv)
{
  if (argc < 4) {
    fprintf(stderr, "usage: serve_driver max_active outdir mesh...\n");
    return 2;
  }
  vh_server *sv = vh_server_create(0, 1, 1, 1e12, atoi(argv[1]), 3, 7);
  int n = argc - 3;
  int *ids = calloc((size_t)n, sizeof(int));
  char dir[4096];
  const char *gap = getenv("SERVE_DRIVER_GAP_MS");
  for (int q = 0; q < n; q++) {
    if (q && gap) usleep((useconds_t)(1000 * atoi(gap)));
    snprintf(dir, sizeof(dir), "%s/%d", argv[2], q);
    ids[q] = vh_server_submit(sv, argv[3 + q], dir);
  }
  int bad = 0;
  for (int q = 0; q < n; q++) {
    vh_request_info info;
    int rc = vh_server_wait(sv, ids[q], &info);
    printf("%d %d %lld %lu %.6f %s\n", q, rc, info.cells, info.steps, info.t_gpu_done, info.err);
    if (rc) bad++;
  }
  vh_server_destroy(sv);
  free(ids);
  return bad ? 1 : 0;
}
