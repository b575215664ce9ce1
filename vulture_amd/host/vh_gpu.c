/*
 * Device driver for a set-up solve: hands every array of a vh_solve to the
 * vgpu C-ABI, runs the time loop on the GPU and writes the observer files.
 * This is the drop-in replacement of propagate() (simulation.c:136-230) for
 * a host that did the reference's set-up.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "vh_internal.h"
#include "vh_gpu.h"

/* SIBC orientation matrices (sibc.c:72-110). */
static const real Acp[4][4] = {{0, 0, 1, 0}, {0, 0, 0, -1}, {-1, 0, 0, 0}, {0, 1, 0, 0}};
static const real Acm[4][4] = {{0, 0, 0, -1}, {0, 0, 1, 0}, {0, 1, 0, 0}, {-1, 0, 0, 0}};
static const real Asp[4][4] = {{-1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, -1, 0}, {0, 0, 0, 1}};
static const real Asm[4][4] = {{0, 1, 0, 0}, {-1, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, -1, 0}};
static const real Bcp[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
static const real Bcm[4][4] = {{0, 1, 0, 0}, {1, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}};
static const real Bsp[4][4] = {{0, 0, -1, 0}, {0, 0, 0, -1}, {1, 0, 0, 0}, {0, 1, 0, 0}};
static const real Bsm[4][4] = {{0, 0, 0, -1}, {0, 0, -1, 0}, {0, 1, 0, 0}, {1, 0, 0, 0}};

static void comb(double C[4][4], real a, const real A[4][4], real b, const real B[4][4])
{
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 4; q++) {
      real v = a * A[p][q] + b * B[p][q];
      C[p][q] = v;
    }
}

static double *cplx(const double complex *z, int n)
{
  double *v = calloc((size_t)(n > 0 ? 2 * n : 2), sizeof(double));
  for (int i = 0; i < n; i++) { v[2 * i] = creal(z[i]); v[2 * i + 1] = cimag(z[i]); }
  return v;
}

static int fail(vh_solve *s, vgpu_ctx *c, const char *what, int rc)
{
  /* without a context (vgpu_create itself failed) the code is all there is */
  const char *why = c ? vgpu_last_error(c)
                      : rc == VGPU_EDEVICE ? "no usable GPU (the HIP runtime found no device or could not select it)"
                      : rc == VGPU_ENOMEM  ? "out of device memory"
                      : rc == VGPU_EINVAL  ? "invalid grid description"
                                           : "";
  snprintf(s->err, sizeof(s->err), "%s failed (%d): %s", what, rc, why);
  return -1;
}

/* Build a context from a set-up solve.  Returns NULL on error (message in s->err). */
vgpu_ctx *vh_gpu_setup(vh_solve *s, int prec, int device, unsigned long chunk)
{
  vgpu_grid_desc g;
  memset(&g, 0, sizeof(g));
  g.nx = s->numCells[0]; g.ny = s->numCells[1]; g.nz = s->numCells[2];
  memcpy(g.gibox, s->gibox, sizeof(g.gibox));
  memcpy(g.gobox, s->gobox, sizeof(g.gobox));
  memcpy(g.gfilim, s->gfilim, sizeof(g.gfilim));
  g.dt = s->dt; g.c0 = s->c0; g.eps0 = s->eps0; g.mu0 = s->mu0; g.eta0 = s->eta0;
  g.grid_type = s->gridType;
  g.xlines = s->xlines; g.ylines = s->ylines; g.zlines = s->zlines;
  for (int a = 0; a < 3; a++) g.nlines[a] = s->numLines[a];
  g.dex = s->dex; g.dey = s->dey; g.dez = s->dez;
  g.idex = s->idex; g.idey = s->idey; g.idez = s->idez;
  g.idhx = s->idhx; g.idhy = s->idhy; g.idhz = s->idhz;
  vgpu_ctx *c = NULL;
  int rc = vgpu_create(&g, prec ? VGPU_FP32 : VGPU_FP64, device, &c);
  if (rc) { fail(s, NULL, "vgpu_create", rc); return NULL; }
  const double *A[3] = {s->alphaEx, s->alphaEy, s->alphaEz};
  const double *B[3] = {s->betaEx, s->betaEy, s->betaEz};
  const double *G[3] = {s->gammaHx, s->gammaHy, s->gammaHz};
  {
    /* VH_MEDIA_F32=1 hands the coefficients over as float arrays, as the
     * reference's float build holds them (the FP32 storage form rounds them to
     * float anyway, so its results are the same bits either way) */
    const char *mf = getenv("VH_MEDIA_F32");
    if (mf && mf[0] == '1') {
      float *buf = malloc(9 * s->ncell * sizeof(float));
      if (!buf) { snprintf(s->err, sizeof(s->err), "out of host memory"); vgpu_destroy(c); return NULL; }
      const double *src[9] = {A[0], A[1], A[2], B[0], B[1], B[2], G[0], G[1], G[2]};
      for (int q = 0; q < 9; q++)
        for (size_t n = 0; n < s->ncell; n++) buf[q * s->ncell + n] = (float)src[q][n];
      const float *fa[3] = {buf, buf + s->ncell, buf + 2 * s->ncell};
      const float *fb[3] = {buf + 3 * s->ncell, buf + 4 * s->ncell, buf + 5 * s->ncell};
      const float *fg[3] = {buf + 6 * s->ncell, buf + 7 * s->ncell, buf + 8 * s->ncell};
      rc = vgpu_set_media_f32(c, fa, fb, fg);
      free(buf);
    } else {
      rc = vgpu_set_media(c, A, B, G);
    }
    if (rc) goto bad;
  }
  {
    vgpu_pml_desc p;
    memset(&p, 0, sizeof(p));
    for (int f = 0; f < 6; f++)
      if (vh_outer_type(s, f) == BT_PML) p.region_mask |= 1u << f;
    memcpy(p.fplim, s->fplim, sizeof(p.fplim));
    const double *ad[3] = {s->adx, s->ady, s->adz}, *bd[3] = {s->bdx, s->bdy, s->bdz};
    const double *ah[3] = {s->ahx, s->ahy, s->ahz}, *bh[3] = {s->bhx, s->bhy, s->bhz};
    const double *ibd[3] = {s->ibdx, s->ibdy, s->ibdz}, *ibh[3] = {s->ibhx, s->ibhy, s->ibhz};
    for (int a = 0; a < 3; a++) { p.ad[a] = ad[a]; p.bd[a] = bd[a]; p.ah[a] = ah[a]; p.bh[a] = bh[a]; p.ibd[a] = ibd[a]; p.ibh[a] = ibh[a]; }
    if ((rc = vgpu_set_pml(c, &p))) goto bad;
  }
  {
    unsigned mm = 0;
    int ghost[6];
    for (int f = 0; f < 6; f++) {
      int t = vh_outer_type(s, f);
      if (t == BT_MUR) mm |= 1u << f;
      ghost[f] = t == BT_PEC ? 1 : t == BT_PMC ? 2 : t == BT_PERIODIC ? 3 : 0;
    }
    double zeta[6];
    for (int f = 0; f < 6; f++) zeta[f] = s->zeta[f];
    if ((rc = vgpu_set_mur(c, mm, zeta, (const int(*)[6][6])s->fmlim))) goto bad;
    if ((rc = vgpu_set_ghosts(c, ghost))) goto bad;
  }
  for (int d = 0; d < s->numDebye; d++) {
    vh_medium *m = &s->media[s->debye[d].medium];
    double *da = cplx(m->dalpha, m->numPoles), *db = cplx(m->dbeta, m->numPoles);
    rc = vgpu_add_debye_block(c, (const int(*)[6])s->debye[d].flim, m->numPoles, da, db);
    free(da);
    free(db);
    if (rc) goto bad;
  }
  for (int n = 0; n < s->numSibc; n++) {
    vh_sibc *sb = &s->sibc[n];
    vh_boundary *bd = &s->bounds[sb->boundary];
    vgpu_sibc_desc d;
    memset(&d, 0, sizeof(d));
    memcpy(d.gbbox, sb->gbbox, sizeof(d.gbbox));
    d.normal = sb->normal;
    if (sb->orientation == 1) { comb(d.A, sb->cosa, Acp, sb->sina, Asp); comb(d.B, sb->cosa, Bcp, sb->sina, Bsp); }
    else { comb(d.A, sb->cosa, Acm, sb->sina, Asm); comb(d.B, sb->cosa, Bcm, sb->sina, Bsm); }
    unsigned char *aa = malloc(sb->nfaces * 4), *ab = malloc(sb->nfaces * 4);
    for (size_t q = 0; q < sb->nfaces * 4; q++) { aa[q] = sb->isAdjA[q]; ab[q] = sb->isAdjB[q]; }
    d.is_adj_a = aa;
    d.is_adj_b = ab;
    for (int dir = 0; dir < 3; dir++) {
      int lo = 2 * dir, hi = lo + 1;
      int tl = vh_outer_type(s, lo), th = vh_outer_type(s, hi);
      d.pmc_index[dir][0] = (tl == BT_PMC || tl == BT_PERIODIC) ? s->gibox[lo] : -1;
      d.pmc_index[dir][1] = (th == BT_PMC || th == BT_PERIODIC) ? s->gibox[hi] : -1;
    }
    double *bufs[48] = {0};
    for (int f = 0; f < 16; f++) {
      vh_rcfilter *rf = &bd->rc[f];
      d.npoles[f] = rf->numPoles;
      d.asymp[f] = rf->asymp;
      bufs[3 * f] = cplx(rf->q0, rf->numPoles);
      bufs[3 * f + 1] = cplx(rf->q1, rf->numPoles);
      bufs[3 * f + 2] = cplx(rf->q2, rf->numPoles);
      d.q0[f] = bufs[3 * f]; d.q1[f] = bufs[3 * f + 1]; d.q2[f] = bufs[3 * f + 2];
    }
    rc = vgpu_add_sibc_surface(c, &d);
    for (int q = 0; q < 48; q++) free(bufs[q]);
    free(aa);
    free(ab);
    if (rc) goto bad;
  }
  for (int w = 0; w < s->numWaveforms; w++) {
    vh_waveform *wf = &s->waveforms[w];
    rc = vgpu_add_waveform(c, wf->type, wf->size, wf->delay, wf->width, wf->frequency, wf->table, wf->tableSize);
    if (rc < 0) goto bad;
  }
  for (int n = 0; n < s->numSources; n++) {
    vh_source *src = &s->sources[n];
    int kind;
    if (src->type == ST_EFIELD || src->type == ST_HFIELD) kind = 0;
    else if (src->type == ST_ELEC_CURR_DENSITY || src->type == ST_MAGN_CURR_DENSITY) kind = 1;
    else continue;
    rc = vgpu_add_source(c, src->field, kind, src->isSoft, src->size, src->waveform, src->delay, src->flim[src->field]);
    if (rc < 0) goto bad;
  }
  for (int n = 0; n < s->numPws; n++) {
    vh_planewave *pw = &s->pws[n];
    vgpu_pw_desc d;
    memset(&d, 0, sizeof(d));
    memcpy(d.flim, pw->flim, sizeof(d.flim));
    for (int f = 0; f < 6; f++) d.active[f] = pw->isActive[f];
    for (int a = 0; a < 3; a++) { d.kinc[a] = pw->kinc[a]; d.ijk0[a] = pw->ijk0[a]; d.r0[a] = pw->r0[a]; }
    for (int f = 0; f < 6; f++) d.Finc[f] = pw->Finc[f];
    d.phase_velocity = pw->phaseVelocity;
    d.waveform = pw->waveform;
    d.delay = pw->delay;
    d.use_aux = s->useAuxGrid;
    if (s->useAuxGrid) {
      d.aux_nx = pw->nx; d.aux_xb = pw->xb;
      d.beta_eyi = pw->betaEyi; d.gamma_hzi = pw->gammaHzi;
      d.aux_adx = pw->adx; d.aux_bdx = pw->bdx; d.aux_ahx = pw->ahx; d.aux_bhx = pw->bhx;
    }
    rc = vgpu_add_planewave(c, &d);
    if (rc < 0) goto bad;
  }
  for (int o = 0; o < s->numObs; o++) {
    vh_observer *ob = &s->obs[o];
    if (ob->quantity != OQ_EH) continue;
    if (ob->format == OF_ASCII) {
      int nf = ob->domain == OD_FREQ ? (int)s->numFreq : 0;
      rc = vgpu_add_point_observer(c, ob->gbbox[XLO], ob->gbbox[YLO], ob->gbbox[ZLO], s->startTimeStep,
                                   s->stopTimeStep, nf, s->omega);
    } else {
      rc = vgpu_add_volume_observer(c, ob->gbbox, ob->step, s->startTimeStep, s->stopTimeStep);
    }
    if (rc < 0) goto bad;
  }
  if ((rc = vgpu_finalize(c, s->numTimeSteps, chunk))) goto bad;
  return c;
bad:
  fail(s, c, "vgpu set-up", rc);
  vgpu_destroy(c);
  return NULL;
}

/* A solve on the device, advanced chunk by chunk: the context, the mapping
 * of the host's observers to device observers, and the host buffer for
 * binary volume samples (drained after every chunk). */
struct vh_dsolve {
  vh_solve *s;
  vgpu_ctx *c;
  unsigned long NT, chunk;
  unsigned long next;          /* steps completed */
  unsigned long enq;           /* steps enqueued */
  unsigned long qfirst[2], qn[2];  /* runs in flight, oldest first */
  int qlen;
  int nbin;                    /* binary observers: rows streamed as runs complete */
  int packable;                /* the context can take part in group runs (separate sweeps) */
  int *pid, *vid;
  uint64_t *vnodes;
  size_t volNodes;
  double loop_ms;
};

void vh_dsolve_close(vh_dsolve *d)
{
  if (!d) return;
  if (d->c) vgpu_destroy(d->c);
  free(d->pid);
  free(d->vid);
  free(d->vnodes);
  free(d);
}

/* Build the device context for a set-up solve (uploads everything).
 * Returns NULL on error (message in s->err). */
vh_dsolve *vh_dsolve_open(vh_solve *s, int prec, int device, unsigned long chunk)
{
  unsigned long NT = s->numTimeSteps;
  if (chunk == 0 || chunk > NT) chunk = NT ? NT : 1;
  vh_dsolve *d = calloc(1, sizeof(*d));
  if (!d) return NULL;
  d->s = s;
  d->NT = NT;
  d->chunk = chunk;
  d->c = vh_gpu_setup(s, prec, device, chunk);
  if (!d->c) { free(d); return NULL; }
  /* Map observers to device ids in the order they were added. */
  int npt = 0, nvol = 0;
  d->pid = calloc((size_t)s->numObs + 1, sizeof(int));
  d->vid = calloc((size_t)s->numObs + 1, sizeof(int));
  d->vnodes = calloc((size_t)s->numObs + 1, sizeof(uint64_t));
  for (int o = 0; o < s->numObs; o++) {
    d->pid[o] = d->vid[o] = -1;
    if (s->obs[o].quantity != OQ_EH) continue;
    if (s->obs[o].format == OF_ASCII) d->pid[o] = npt++;
    else d->vid[o] = nvol++;
  }
  for (int o = 0; o < s->numObs; o++)
    if (d->vid[o] >= 0) {
      vgpu_read_volume_observer(d->c, d->vid[o], NULL, 0, &d->vnodes[d->vid[o]]);
      d->volNodes += d->vnodes[d->vid[o]];
    }
  for (int o = 0; o < s->numObs; o++) d->nbin += s->obs[o].format == OF_BINARY;
  s->binaryStreamed = d->nbin > 0;
  vgpu_stats st;
  d->packable = vgpu_get_stats(d->c, &st) == 0 && !st.leap && !st.debye_fused;
  return d;
}

vgpu_ctx *vh_dsolve_context(vh_dsolve *d) { return d->c; }
unsigned long vh_dsolve_remaining(const vh_dsolve *d) { return d->NT - d->enq; }
int vh_dsolve_done(const vh_dsolve *d) { return d->next >= d->NT; }
int vh_dsolve_packable(const vh_dsolve *d) { return d->packable; }

/* Enqueue the next chunk of steps on the context's stream (asynchronous),
 * keeping at most two chunks in flight, so the device runs one while the
 * host completes the other.  Returns the steps enqueued (0 when two are in
 * flight or every step is enqueued), <0 on error. */
long vh_dsolve_enqueue(vh_dsolve *d)
{
  if (d->qlen >= 2 || d->enq >= d->NT) return 0;
  unsigned long n = d->NT - d->enq < d->chunk ? d->NT - d->enq : d->chunk;
  int rc = vgpu_run(d->c, d->enq, n, NULL);
  if (rc) return fail(d->s, d->c, "vgpu_run", rc);
  d->qfirst[d->qlen] = d->enq;
  d->qn[d->qlen] = n;
  d->qlen++;
  d->enq += n;
  return (long)n;
}

/* Enqueue the next chunk of n solves together, in lockstep, with one packed
 * launch per step kernel over all of them (vgpu_run_group): each advances by
 * the steps the member with the fewest left still has, at most the chunk.
 * Each member's chunk is then completed on its own (vh_dsolve_complete), as
 * after vh_dsolve_enqueue.  Returns the steps enqueued (0 when a member has
 * two chunks in flight or no steps left), <0 on error. */
long vh_dsolve_enqueue_group(vh_dsolve **ds, int n)
{
  if (n < 1 || n > VGPU_GROUP_MAX) return -1;
  vgpu_ctx *cs[VGPU_GROUP_MAX];
  uint64_t first[VGPU_GROUP_MAX];
  unsigned long k = 0;
  for (int i = 0; i < n; i++) {
    vh_dsolve *d = ds[i];
    if (d->qlen >= 2 || d->enq >= d->NT) return 0;
    const unsigned long left = d->NT - d->enq < d->chunk ? d->NT - d->enq : d->chunk;
    if (i == 0 || left < k) k = left;
    cs[i] = d->c;
    first[i] = d->enq;
  }
  int rc = vgpu_run_group(cs, n, first, k);
  if (rc) {
    for (int i = 0; i < n; i++)  /* the member the device layer named */
      if (vgpu_last_error(cs[i])[0]) return fail(ds[i]->s, cs[i], "vgpu_run_group", rc);
    return fail(ds[0]->s, cs[0], "vgpu_run_group", rc);
  }
  for (int i = 0; i < n; i++) {
    vh_dsolve *d = ds[i];
    d->qfirst[d->qlen] = d->enq;
    d->qn[d->qlen] = k;
    d->qlen++;
    d->enq += k;
  }
  return (long)k;
}

/* 1 when the oldest chunk in flight has finished (or none is in flight), 0
 * while it runs, <0 on a device error. */
int vh_dsolve_ready(vh_dsolve *d)
{
  if (!d->qlen) return 1;
  int rc = vgpu_run_result(d->c, d->qfirst[0], 0, NULL, NULL, NULL);
  return rc < 0 ? fail(d->s, d->c, "vgpu_run", rc) : rc;
}

/* Complete the oldest chunk in flight without waiting for the one behind it:
 * add its device time and write its excite.dat lines and impulse.dat rows
 * straight from the pinned copy of its samples (observer.c:1164-1198 writes
 * them inside the loop; the files come out the same).  Returns 0 or <0. */
int vh_dsolve_complete(vh_dsolve *d)
{
  if (!d->qlen) return 0;
  vh_solve *s = d->s;
  double ms = 0;
  const float *smp = NULL;
  uint64_t n = 0;
  int rc = vgpu_run_result(d->c, d->qfirst[0], 1, &ms, &smp, &n);
  if (rc < 0) return fail(s, d->c, "vgpu_run", rc);
  d->loop_ms += ms;
  if (d->nbin) {
    if (!smp && d->volNodes) { snprintf(s->err, sizeof(s->err), "volume observer samples missing"); return -1; }
    for (uint64_t q = 0; q < n; q++) {
      const unsigned long step = d->qfirst[0] + q;
      vh_observe_binary_step(s, step, step * s->dt, smp + q * d->volNodes * 6);
    }
  }
  d->next += n;
  d->qfirst[0] = d->qfirst[1];
  d->qn[0] = d->qn[1];
  d->qlen--;
  return 0;
}

/* After the last chunk: read the point observers (time samples and device
 * DFTs) and free the device context, so its memory is available at once. */
int vh_dsolve_collect(vh_dsolve *d, double ***tdOut)
{
  vh_solve *s = d->s;
  unsigned long nout = s->stopTimeStep - s->startTimeStep + 1;
  double **td = calloc((size_t)s->numObs + 1, sizeof(double *));
  for (int o = 0; o < s->numObs; o++) {
    if (d->pid[o] < 0) continue;
    td[o] = malloc(nout * 6 * sizeof(double));
    vh_observer *ob = &s->obs[o];
    double *re = NULL, *im = NULL;
    if (ob->domain == OD_FREQ) {
      re = malloc(6 * s->numFreq * sizeof(double));
      im = malloc(6 * s->numFreq * sizeof(double));
    }
    vgpu_read_point_observer(d->c, d->pid[o], td[o], nout, re, im);
    if (re) {
      for (size_t q = 0; q < 6 * s->numFreq; q++) { ob->dft_re[q] = re[q]; ob->dft_im[q] = im[q]; }
      free(re);
      free(im);
    }
  }
  s->deviceDft = 1;
  vgpu_destroy(d->c);
  d->c = NULL;
  *tdOut = td;
  return 0;
}

/* Replay the collected samples through the host observer writers (the
 * reference's per-step file output, observer.c:622-650) and free them. */
void vh_dsolve_write(vh_dsolve *d, double **td)
{
  vh_solve *s = d->s;
  double *pv = calloc((size_t)s->numObs * 6 + 6, sizeof(double));
  for (unsigned long n = 0; n < d->NT; n++) {
    real t = n * s->dt;
    bool ot = n >= s->startTimeStep && n <= s->stopTimeStep;
    for (int o = 0; o < s->numObs; o++)
      if (d->pid[o] >= 0 && ot)
        for (int q = 0; q < 6; q++) pv[6 * o + q] = td[o][(n - s->startTimeStep) * 6 + q];
    vh_observe_step(s, n, t, pv, NULL);
  }
  for (int o = 0; o < s->numObs; o++) free(td[o]);
  free(td);
  free(pv);
}

double vh_dsolve_loop_ms(const vh_dsolve *d) { return d->loop_ms; }

/* Run the whole time loop on the device and write the observer files.
 * loop_ms receives the device time of the time loop alone. */
int vh_gpu_run(vh_solve *s, int prec, int device, unsigned long chunk, double *loop_ms)
{
  vh_dsolve *d = vh_dsolve_open(s, prec, device, chunk);
  if (!d) return -1;
  while (!vh_dsolve_done(d)) {
    long e;
    while ((e = vh_dsolve_enqueue(d)) > 0) {}
    if (e < 0 || vh_dsolve_complete(d)) { vh_dsolve_close(d); return -1; }
  }
  if (loop_ms) *loop_ms = d->loop_ms;
  double **td = NULL;
  vh_dsolve_collect(d, &td);
  vh_dsolve_write(d, td);
  vh_dsolve_close(d);
  return 0;
}

/* Run n set-up solves to their ends together, packed (vh_dsolve_enqueue_group),
 * and write each one's observer files; a solve leaves the group when all its
 * steps are enqueued.  loop_ms receives the host's wall time of the loop, from
 * the first enqueue to the last completion.  On error the message is in the
 * failing solve's err. */
int vh_gpu_run_group(vh_solve **ss, int n, int prec, int device, unsigned long chunk, double *loop_ms)
{
  if (n < 1 || n > VGPU_GROUP_MAX) return -1;
  vh_dsolve *ds[VGPU_GROUP_MAX] = {0};
  int rc = 0;
  for (int i = 0; i < n && !rc; i++)
    if (!(ds[i] = vh_dsolve_open(ss[i], prec, device, chunk))) rc = -1;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  while (!rc) {
    for (;;) {
      vh_dsolve *live[VGPU_GROUP_MAX];
      int m = 0;
      for (int i = 0; i < n; i++)
        if (ds[i]->enq < ds[i]->NT) live[m++] = ds[i];
      if (!m) break;
      const long e = vh_dsolve_enqueue_group(live, m);
      if (e < 0) rc = -1;
      if (e <= 0) break;
    }
    for (int i = 0; i < n && !rc; i++)
      if (ds[i]->qlen && vh_dsolve_complete(ds[i])) rc = -1;
    int all = 1;
    for (int i = 0; i < n; i++) all = all && vh_dsolve_done(ds[i]);
    if (all) break;
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (loop_ms) *loop_ms = (t1.tv_sec - t0.tv_sec) * 1e3 + (t1.tv_nsec - t0.tv_nsec) * 1e-6;
  for (int i = 0; i < n && !rc; i++) {
    double **td = NULL;
    vh_dsolve_collect(ds[i], &td);
    vh_dsolve_write(ds[i], td);
  }
  for (int i = 0; i < n; i++) vh_dsolve_close(ds[i]);
  return rc;
}

/* Parse, set up, run on the device and write observer files: the whole
 * solver behind one call (the reference's main(), vulture.c:77-197, with the
 * time loop on the GPU).  Returns 0 on success. */
int vh_run_mesh_gpu(const char *mesh, const char *outdir, int prec, int device, int fullPrecision, long steps,
                    unsigned long chunk, double *loop_ms, char *err, int errlen)
{
  vh_solve *s = vh_create();
  if (!s) return -1;
  vh_set_dirs(s, NULL, outdir);
  s->fullPrecision = fullPrecision;
  int rc = vh_read_mesh(s, mesh);
  if (!rc && steps >= 0) s->numTimeSteps = (unsigned long)steps;
  if (!rc) rc = vh_init(s);
  if (!rc) rc = vh_open_observers(s);
  if (!rc) rc = vh_gpu_run(s, prec, device, chunk, loop_ms);
  if (rc && err) snprintf(err, (size_t)errlen, "%s", vh_error(s));
  vh_destroy(s);
  return rc;
}

/* vh_run_mesh_gpu for n meshes run together as one packed group
 * (vh_gpu_run_group).  err receives the first failing mesh's message. */
int vh_run_meshes_gpu_group(const char *const *meshes, const char *const *outdirs, int n, int prec, int device,
                            int fullPrecision, unsigned long chunk, double *loop_ms, char *err, int errlen)
{
  if (n < 1 || n > VGPU_GROUP_MAX) {
    if (err) snprintf(err, (size_t)errlen, "a group holds 1 to %d solves, not %d", VGPU_GROUP_MAX, n);
    return -1;
  }
  vh_solve *ss[VGPU_GROUP_MAX] = {0};
  int rc = 0;
  for (int i = 0; i < n && !rc; i++) {
    vh_solve *s = ss[i] = vh_create();
    if (!s) {
      rc = -1;
      break;
    }
    vh_set_dirs(s, NULL, outdirs[i]);
    s->fullPrecision = fullPrecision;
    rc = vh_read_mesh(s, meshes[i]);
    if (!rc) rc = vh_init(s);
    if (!rc) rc = vh_open_observers(s);
  }
  if (!rc) rc = vh_gpu_run_group(ss, n, prec, device, chunk, loop_ms);
  if (rc && err) {
    snprintf(err, (size_t)errlen, "group run failed");
    for (int i = 0; i < n; i++)
      if (ss[i] && vh_error(ss[i])[0]) {
        snprintf(err, (size_t)errlen, "%s: %s", meshes[i], vh_error(ss[i]));
        break;
      }
  }
  for (int i = 0; i < n; i++)
    if (ss[i]) vh_destroy(ss[i]);
  return rc;
}

/* Parse and set up a mesh and build its device context without running it,
 * for callers that drive vgpu_run themselves (timing, serving).  No observer
 * files are opened.  Returns 0 and sets *solve and *ctx on success. */
int vh_gpu_open(const char *mesh, int prec, int device, long steps, unsigned long chunk, vh_solve **solve,
                vgpu_ctx **ctx, char *err, int errlen)
{
  *solve = NULL;
  *ctx = NULL;
  vh_solve *s = vh_create();
  if (!s) return -1;
  vh_set_dirs(s, NULL, NULL);
  int rc = vh_read_mesh(s, mesh);
  if (!rc && steps >= 0) s->numTimeSteps = (unsigned long)steps;
  if (!rc) rc = vh_init(s);
  vgpu_ctx *c = NULL;
  if (!rc) {
    unsigned long NT = s->numTimeSteps;
    if (chunk == 0 || chunk > NT) chunk = NT ? NT : 1;
    c = vh_gpu_setup(s, prec, device, chunk);
    if (!c) rc = -1;
  }
  if (rc) {
    if (err) snprintf(err, (size_t)errlen, "%s", vh_error(s));
    vh_destroy(s);
    return rc;
  }
  *solve = s;
  *ctx = c;
  return 0;
}

void vh_gpu_close(vh_solve *s, vgpu_ctx *c)
{
  if (c) vgpu_destroy(c);
  if (s) vh_destroy(s);
}

/* out[0] = gobox cells (mesh + PML), out[1] = mesh cells, out[2] = time steps. */
void vh_solve_sizes(const vh_solve *s, long long out[3])
{
  out[0] = (long long)(s->gobox[XHI] - s->gobox[XLO]) * (s->gobox[YHI] - s->gobox[YLO]) *
           (s->gobox[ZHI] - s->gobox[ZLO]);
  out[1] = (long long)(s->gibox[XHI] - s->gibox[XLO]) * (s->gibox[YHI] - s->gibox[YLO]) *
           (s->gibox[ZHI] - s->gibox[ZLO]);
  out[2] = (long long)s->numTimeSteps;
}

/* Array extents numCells[3] (gobox + ghost layers) of the host layout. */
void vh_solve_shape(const vh_solve *s, long long out[3])
{
  for (int a = 0; a < 3; a++) out[a] = s->numCells[a];
}

/* Outer field limits gfolim[component][face] (grid.c:404), the limits the
 * update-once check (vgpu_check_limits) tallies against. */
void vh_solve_field_limits(const vh_solve *s, int out[6][6])
{
  for (int c = 0; c < 6; c++)
    for (int f = 0; f < 6; f++) out[c][f] = s->gfolim[c][f];
}
