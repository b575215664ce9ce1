/*
 * Host-only stand-in for libvgpu, for testing the host's device driver and
 * solve server on a machine without a GPU (test infrastructure).  Every
 * entry point of include/vgpu.h is accepted; runs "complete" after a few
 * vgpu_query or vgpu_run_result polls so the server's asynchronous paths are
 * exercised; observer reads return zeros.
 */
#include <stdlib.h>
#include <string.h>
#include "../../include/vgpu.h"

/* The set-up of the last context, kept when vgpu_stub_capture is set, for
 * single-threaded checkers that replay the library's host-side planning
 * (tests/native/shell_check.cpp). */
int vgpu_stub_capture;
vgpu_grid_desc vgpu_stub_grid;
unsigned vgpu_stub_pmlmask, vgpu_stub_murmask;
int vgpu_stub_fplim[6][6][6], vgpu_stub_fmlim[6][6][6];

struct vgpu_ctx {
  int finalized, npt, nvol, polls;
  size_t bytes;
  uint64_t total;
  int nf[256];
  uint64_t vnodes[256];
  uint64_t maxrun, runs, rfirst[2], rn[2];
  int rpolls[2], rvalid[2];
  float *zeros;             /* samples of a run: maxrun x all volume nodes x 6 */
};

int vgpu_create(const vgpu_grid_desc *g, vgpu_prec prec, int device, vgpu_ctx **out)
{
  (void)device;
  vgpu_ctx *c = calloc(1, sizeof(*c));
  if (vgpu_stub_capture) {
    vgpu_stub_grid = *g;
    vgpu_stub_pmlmask = vgpu_stub_murmask = 0;
    memset(vgpu_stub_fplim, 0, sizeof(vgpu_stub_fplim));
    memset(vgpu_stub_fmlim, 0, sizeof(vgpu_stub_fmlim));
  }
  c->bytes = (size_t)g->nx * g->ny * g->nz * (prec == VGPU_FP64 ? 8 : 4) * 15;
  *out = c;
  return 0;
}
int vgpu_set_media(vgpu_ctx *c, const double *a[3], const double *b[3], const double *g[3])
{ (void)c; (void)a; (void)b; (void)g; return 0; }
int vgpu_set_media_f32(vgpu_ctx *c, const float *a[3], const float *b[3], const float *g[3])
{ (void)c; (void)a; (void)b; (void)g; return 0; }
int vgpu_set_pml(vgpu_ctx *c, const vgpu_pml_desc *p)
{
  (void)c;
  if (vgpu_stub_capture) {
    vgpu_stub_pmlmask = p->region_mask & 0x3Fu;
    memcpy(vgpu_stub_fplim, p->fplim, sizeof(vgpu_stub_fplim));
  }
  return 0;
}
int vgpu_set_mur(vgpu_ctx *c, unsigned m, const double z[6], const int f[6][6][6])
{
  (void)c; (void)z;
  if (vgpu_stub_capture) {
    vgpu_stub_murmask = m & 0x3Fu;
    memcpy(vgpu_stub_fmlim, f, sizeof(vgpu_stub_fmlim));
  }
  return 0;
}
int vgpu_set_ghosts(vgpu_ctx *c, const int t[6]) { (void)c; (void)t; return 0; }
int vgpu_add_debye_block(vgpu_ctx *c, const int f[3][6], int n, const double *a, const double *b)
{ (void)c; (void)f; (void)n; (void)a; (void)b; return 0; }
int vgpu_add_sibc_surface(vgpu_ctx *c, const vgpu_sibc_desc *d) { (void)c; (void)d; return 0; }
int vgpu_add_waveform(vgpu_ctx *c, int t, double s, double d, double w, double f, const double *tab, unsigned long n)
{ (void)c; (void)t; (void)s; (void)d; (void)w; (void)f; (void)tab; (void)n; return 0; }
int vgpu_add_source(vgpu_ctx *c, int f, int k, int s, double z, int w, double d, const int l[6])
{ (void)c; (void)f; (void)k; (void)s; (void)z; (void)w; (void)d; (void)l; return 0; }
int vgpu_add_planewave(vgpu_ctx *c, const vgpu_pw_desc *p) { (void)c; (void)p; return 0; }
int vgpu_add_point_observer(vgpu_ctx *c, int i, int j, int k, uint64_t f, uint64_t l, int nf, const double *om)
{
  (void)i; (void)j; (void)k; (void)f; (void)l; (void)om;
  if (c->npt < 256) c->nf[c->npt] = nf;
  return c->npt++;
}
int vgpu_add_volume_observer(vgpu_ctx *c, const int b[6], const int s[3], uint64_t f, uint64_t l)
{
  (void)f; (void)l;
  uint64_t n = 1;
  for (int a = 0; a < 3; a++) n *= (uint64_t)((b[2 * a + 1] - b[2 * a]) / s[a] + 1);
  if (c->nvol < 256) c->vnodes[c->nvol] = n;
  return c->nvol++;
}
int vgpu_finalize(vgpu_ctx *c, uint64_t total, uint64_t maxrun)
{
  uint64_t nodes = 0;
  for (int v = 0; v < c->nvol && v < 256; v++) nodes += c->vnodes[v];
  c->total = total;
  c->maxrun = maxrun ? maxrun : 1;
  c->zeros = nodes ? calloc(c->maxrun * nodes * 6, sizeof(float)) : NULL;
  c->finalized = 1;
  return 0;
}
int vgpu_run(vgpu_ctx *c, uint64_t first, uint64_t n, void *s)
{
  (void)s;
  const int q = (int)(c->runs++ & 1);
  c->rfirst[q] = first;
  c->rn[q] = n;
  c->rvalid[q] = 1;
  c->rpolls[q] = rand() % 4;
  c->polls = rand() % 4;
  return 0;
}
int vgpu_run_group(vgpu_ctx *const *cs, int n, const uint64_t *first, uint64_t nsteps)
{
  if (n < 1 || n > VGPU_GROUP_MAX) return VGPU_EINVAL;
  for (int i = 0; i < n; i++) vgpu_run(cs[i], first[i], nsteps, NULL);
  return 0;
}
int vgpu_run_result(vgpu_ctx *c, uint64_t first, int wait, double *ms, const float **samples, uint64_t *nsteps)
{
  int q = -1;
  for (int back = 1; back >= 0; back--) {
    if ((uint64_t)back >= c->runs) continue;
    const int slot = (int)((c->runs - 1 - (uint64_t)back) & 1);
    if (c->rvalid[slot] && c->rfirst[slot] == first) q = slot;
  }
  if (q < 0) return VGPU_EINVAL;
  if (!wait && c->rpolls[q]-- > 0) return 0;
  if (ms) *ms = 0;
  if (samples) *samples = c->zeros;
  if (nsteps) *nsteps = c->rn[q];
  return 1;
}
int vgpu_synchronize(vgpu_ctx *c) { c->polls = 0; return 0; }
int vgpu_query(vgpu_ctx *c) { return c->polls-- <= 0 ? 1 : 0; }
int vgpu_read_point_observer(vgpu_ctx *c, int id, double *td, uint64_t n, double *re, double *im)
{
  if (td) memset(td, 0, n * 6 * sizeof(double));
  int nf = (id >= 0 && id < 256) ? c->nf[id] : 0;
  if (re) memset(re, 0, 6 * (size_t)nf * sizeof(double));
  if (im) memset(im, 0, 6 * (size_t)nf * sizeof(double));
  return 0;
}
int vgpu_read_volume_observer(vgpu_ctx *c, int id, float *out, uint64_t n, uint64_t *nodes)
{
  const uint64_t vn = (id >= 0 && id < 256) ? c->vnodes[id] : 0;
  if (nodes) *nodes = vn;
  if (out) memset(out, 0, vn * 6 * n * sizeof(float));
  return 0;
}
int vgpu_download_field(vgpu_ctx *c, int comp, double *h) { (void)c; (void)comp; (void)h; return 0; }
size_t vgpu_footprint_bytes(const vgpu_ctx *c) { return c ? c->bytes : 0; }
int vgpu_get_stats(const vgpu_ctx *c, vgpu_stats *o) { (void)c; memset(o, 0, sizeof(*o)); return 0; }
int vgpu_last_run_timing(const vgpu_ctx *c, double *t, double *e, double *h)
{ (void)c; if (t) *t = 0; if (e) *e = 0; if (h) *h = 0; return 0; }
const char *vgpu_last_error(const vgpu_ctx *c) { (void)c; return ""; }
void vgpu_destroy(vgpu_ctx *c)
{
  if (c) free(c->zeros);
  free(c);
}
