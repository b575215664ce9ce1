/*
 * Runs the meshes named on the command line as one group through
 * vh_run_meshes_gpu_group, linked against the host-only stand-in for the
 * device library: the host side of group runs (chunking, members leaving
 * the group, completion and file output) without a GPU.
 *
 * usage: group_driver CHUNK OUTDIR MESH...   (prints "rc message")
 */
#include <stdio.h>
#include <stdlib.h>
#include "vh_gpu.h"

int main(int argc, char **argv)
{
  if (argc < 3) {
    fprintf(stderr, "usage: %s CHUNK OUTDIR MESH...\n", argv[0]);
    return 2;
  }
  const unsigned long chunk = strtoul(argv[1], NULL, 10);
  const int n = argc - 3;
  const char **outs = calloc((size_t)n + 1, sizeof(char *));
  for (int i = 0; i < n; i++) {
    char *o = malloc(1024);
    snprintf(o, 1024, "%s/%d", argv[2], i);
    outs[i] = o;
  }
  char err[1024] = "";
  double ms = 0;
  const int rc = vh_run_meshes_gpu_group((const char *const *)(argv + 3), outs, n, 1, 0, 1, chunk, &ms, err,
                                         (int)sizeof(err));
  printf("%d %s\n", rc, rc ? err : "ok");
  for (int i = 0; i < n; i++) free((char *)outs[i]);
  free(outs);
  return rc ? 1 : 0;
}
