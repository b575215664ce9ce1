/*
 * Solve server: many Vulture solves resident and in flight on one GPU.
 *
 * Requests (a mesh file and an output directory) arrive at any time.  Host
 * worker threads parse and set them up (the reference's initialisation,
 * vulture.c:87-147) and later write their observer files; one scheduler
 * thread admits set-up solves while device memory and active slots allow,
 * uploads them, and keeps every active solve's next chunk of time steps
 * enqueued on that solve's own stream, so the kernels of different solves
 * overlap on the device.  A finished solve's device memory is freed as soon
 * as its observer data is read back, and the next waiting solve is admitted
 * in its place (continuous batching).  Results are the same as a solve run
 * alone: each solve has its own context and stream, and the step kernels
 * are the same ones.
 *
 * Order: set-up solves are admitted least work (cells x steps) first, and of
 * the active solves only the `running` ones with the least work left keep
 * chunks in flight (two each); the others wait resident until they are among
 * the least.  With two running solves (the default) the two shortest
 * remaining ones overlap on the device and every other waits, preempted at
 * chunk boundaries: a short request that arrives behind long ones shares the
 * device with one other solve at most, without stream priorities.
 *
 * Packing: a running place whose solve is small (under packCells gobox
 * cells, a grid whose own launches cannot fill the device) takes up to
 * VGPU_GROUP_MAX of the small solves next in that order with it, and they
 * step together in lockstep, one packed launch per step kernel over all of
 * them (vh_dsolve_enqueue_group, vgpu_run_group).  Each still completes its
 * chunks on its own, with the same results as alone.  While the first
 * running place is small, the other places go to small solves only.
 *
 * The reference has no counterpart (one solve per process, simulation.c).
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#include "vh_internal.h"
#include "vh_gpu.h"
#include "vh_serve.h"

enum { RQ_QUEUED, RQ_SETUP, RQ_READY, RQ_ADMITTED, RQ_UPLOAD, RQ_ACTIVE, RQ_WRITE, RQ_WRITING, RQ_DONE, RQ_FAILED };

typedef struct {
  vh_request_info info;
  char mesh[VH_PATH], outdir[VH_PATH];
  int state;
  vh_solve *s;
  vh_dsolve *d;
  size_t est;
  double **td;
} request;

struct vh_server {
  int device, prec, fullPrecision, maxActive, nworkers, lookahead, running;
  long long packCells;  /* solves of fewer gobox cells share packed launches (0: none) */
  int smallAlone;       /* a running small solve shares the device with small ones only */
  int nextWorker, uploaders;  /* the first `uploaders` workers never set up */
  unsigned long chunk;
  size_t budget, used;
  pthread_mutex_t mu;
  pthread_cond_t cvWork, cvDone;
  request **rq;
  int nrq, cap;
  int active, ready, stop;
  pthread_t *workers, sched;
  struct timespec t0;
};

static double now(const vh_server *sv)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)(t.tv_sec - sv->t0.tv_sec) + 1e-9 * (double)(t.tv_nsec - sv->t0.tv_nsec);
}

double vh_server_now(vh_server *sv) { return now(sv); }

/* Device bytes a set-up solve will need: fields, coefficient arrays and
 * classes over the padded allocation, UPML auxiliaries, Debye state and
 * observer buffers, with a margin for work lists.  Used only to admit. */
static size_t estimate_bytes(const vh_solve *s, int prec)
{
  const size_t esz = prec == VGPU_FP64 ? 8 : 4;
  const size_t align = 128 / esz;
  const size_t px = ((size_t)s->numCells[0] + align - 1) / align * align;
  const size_t ncell = px * (size_t)s->numCells[1] * (size_t)s->numCells[2];
  size_t b = ncell * (15 * esz + 4);
  const size_t go = (size_t)(s->gobox[XHI] - s->gobox[XLO]) * (s->gobox[YHI] - s->gobox[YLO]) *
                    (s->gobox[ZHI] - s->gobox[ZLO]);
  const size_t gi = (size_t)(s->gibox[XHI] - s->gibox[XLO]) * (s->gibox[YHI] - s->gibox[YLO]) *
                    (s->gibox[ZHI] - s->gibox[ZLO]);
  b += (go - gi) * 9 * esz;
  for (int d = 0; d < s->numDebye; d++) {
    size_t n = 0;
    for (int f = 0; f < 3; f++) {
      const int *l = s->debye[d].flim[f];
      n += (size_t)(l[XHI] - l[XLO] + 1) * (l[YHI] - l[YLO] + 1) * (l[ZHI] - l[ZLO] + 1);
    }
    b += n * (esz * 3 + 16);
  }
  b += (size_t)s->numObs * (s->numTimeSteps * 6 * 8 + 12 * s->numFreq * 8);
  return b + b / 8 + ((size_t)64 << 20);
}

/* Steps per chunk of a request: the server's chunk, or fewer for large grids,
 * so that a chunk is at most about 1e9 cell-updates (about 40 ms on the
 * device).  A running solve gives way to a shorter one at its next chunk
 * boundary, and large grids would otherwise hold the device for a long chunk
 * after losing their turn.  Even, so captured graphs replay on either field
 * set; at least 20 steps. */
static unsigned long chunk_of(const vh_server *sv, const request *r)
{
  const double cap = 1e9 / (double)(r->info.cells > 0 ? r->info.cells : 1);
  unsigned long c = sv->chunk;
  if ((double)c > cap) c = (unsigned long)cap & ~1ul;
  return c < 20 ? 20 : c;
}

/* Work of a request (gobox cells x steps) and what is left of it. */
static double work_of(const request *r) { return (double)r->info.cells * (double)r->info.steps; }

static int cmp_work_left(const void *a, const void *b)
{
  const request *x = *(request *const *)a, *y = *(request *const *)b;
  const double wx = (double)x->info.cells * (double)vh_dsolve_remaining(x->d);
  const double wy = (double)y->info.cells * (double)vh_dsolve_remaining(y->d);
  if (wx != wy) return wx < wy ? -1 : 1;
  return x->info.id < y->info.id ? -1 : x->info.id > y->info.id;
}

/* The request in state `from` with the least work (arrival order breaks
 * ties), moved to state `to`: uploads go shortest first, as admission does. */
static request *take_least(vh_server *sv, int from, int to)
{
  request *best = NULL;
  for (int q = 0; q < sv->nrq; q++)
    if (sv->rq[q]->state == from && (!best || work_of(sv->rq[q]) < work_of(best))) best = sv->rq[q];
  if (best) best->state = to;
  return best;
}

static request *take(vh_server *sv, int from, int to)
{
  for (int q = 0; q < sv->nrq; q++)
    if (sv->rq[q]->state == from) {
      sv->rq[q]->state = to;
      return sv->rq[q];
    }
  return NULL;
}

static void fail_request(vh_server *sv, request *r, const char *msg)
{
  snprintf(r->info.err, sizeof(r->info.err), "%s", msg);
  for (size_t q = strlen(r->info.err); q > 0 && (r->info.err[q - 1] == '\n' || r->info.err[q - 1] == ' '); q--)
    r->info.err[q - 1] = 0;
  r->info.status = -1;
  r->info.t_done = now(sv);
  r->state = RQ_FAILED;
  pthread_cond_broadcast(&sv->cvDone);
}

/* Host workers: set-up (parse, initialise, open observer files), upload of
 * admitted solves, and observer-file writing of finished ones. */
static void *worker(void *arg)
{
  vh_server *sv = arg;
  pthread_mutex_lock(&sv->mu);
  /* a few workers only upload and write, so an admitted solve never waits
   * behind the parsing of large meshes */
  const int me = sv->nextWorker++;
  while (!sv->stop) {
    request *r;
    if ((r = take(sv, RQ_WRITE, RQ_WRITING))) {
      pthread_mutex_unlock(&sv->mu);
      vh_dsolve_write(r->d, r->td);
      r->td = NULL;
      vh_dsolve_close(r->d);
      r->d = NULL;
      vh_destroy(r->s);
      r->s = NULL;
      pthread_mutex_lock(&sv->mu);
      r->info.t_done = now(sv);
      r->info.status = 1;
      r->state = RQ_DONE;
      pthread_cond_broadcast(&sv->cvDone);
    } else if ((r = take_least(sv, RQ_ADMITTED, RQ_UPLOAD))) {
      pthread_mutex_unlock(&sv->mu);
      vh_dsolve *d = vh_dsolve_open(r->s, sv->prec, sv->device, chunk_of(sv, r));
      if (d) vh_release_grid_arrays(r->s);
      pthread_mutex_lock(&sv->mu);
      if (!d) {
        sv->used -= r->est;
        sv->active--;
        fail_request(sv, r, vh_error(r->s));
        vh_destroy(r->s);
        r->s = NULL;
      } else {
        r->d = d;
        const size_t real = vgpu_footprint_bytes(vh_dsolve_context(d));
        sv->used = sv->used - r->est + real;
        r->est = real;
        r->info.bytes = real;
        r->info.t_admit = now(sv);
        r->state = RQ_ACTIVE;
      }
    } else if (me >= sv->uploaders && sv->ready < sv->lookahead && (r = take(sv, RQ_QUEUED, RQ_SETUP))) {
      pthread_mutex_unlock(&sv->mu);
      vh_solve *s = vh_create();
      int rc = s ? 0 : -1;
      if (!rc) {
        vh_set_dirs(s, NULL, r->outdir);
        s->fullPrecision = sv->fullPrecision;
        rc = vh_read_mesh(s, r->mesh);
        if (!rc) rc = vh_init(s);
        if (!rc) rc = vh_open_observers(s);
      }
      pthread_mutex_lock(&sv->mu);
      if (rc) {
        fail_request(sv, r, s ? vh_error(s) : "out of memory");
        vh_destroy(s);
      } else {
        r->s = s;
        r->est = estimate_bytes(s, sv->prec);
        long long sz[3];
        vh_solve_sizes(s, sz);
        r->info.cells = sz[0];
        r->info.steps = (unsigned long)sz[2];
        r->info.t_ready = now(sv);
        r->state = RQ_READY;
        sv->ready++;
      }
    } else {
      pthread_cond_wait(&sv->cvWork, &sv->mu);
      continue;
    }
    pthread_cond_broadcast(&sv->cvWork);
  }
  pthread_mutex_unlock(&sv->mu);
  return NULL;
}

/* A solve that failed on the device: its slot and memory are freed and its
 * request fails with the solve's message. */
static void drop_failed(vh_server *sv, request *r)
{
  pthread_mutex_lock(&sv->mu);
  sv->used -= r->est;
  sv->active--;
  fail_request(sv, r, vh_error(r->s));
  vh_dsolve_close(r->d);
  r->d = NULL;
  vh_destroy(r->s);
  r->s = NULL;
  pthread_cond_broadcast(&sv->cvWork);
  pthread_mutex_unlock(&sv->mu);
}

/* Scheduler: admission (least work first), completion, chunk enqueueing. */
static void *scheduler(void *arg)
{
  vh_server *sv = arg;
  request **act = calloc((size_t)sv->maxActive + 1, sizeof(request *));
  char *live = calloc((size_t)sv->maxActive + 1, 1);
  for (;;) {
    pthread_mutex_lock(&sv->mu);
    if (sv->stop) {
      pthread_mutex_unlock(&sv->mu);
      break;
    }
    int admitted = 0;
    while (sv->active < sv->maxActive) {
      /* the set-up solve with the least work (arrival order breaks ties) */
      request *best = NULL;
      for (int q = 0; q < sv->nrq; q++) {
        request *r = sv->rq[q];
        if (r->state == RQ_READY && (!best || work_of(r) < work_of(best))) best = r;
      }
      if (!best || (sv->used + best->est > sv->budget && sv->active > 0)) break;
      best->state = RQ_ADMITTED;
      sv->ready--;
      sv->used += best->est;
      sv->active++;
      admitted = 1;
    }
    int n = 0;
    for (int q = 0; q < sv->nrq && n < sv->maxActive; q++)
      if (sv->rq[q]->state == RQ_ACTIVE) act[n++] = sv->rq[q];
    if (admitted) pthread_cond_broadcast(&sv->cvWork);
    pthread_mutex_unlock(&sv->mu);
    /* least work left first; only the first `running` get new chunks */
    qsort(act, (size_t)n, sizeof(request *), cmp_work_left);

    int busy = 0;
    /* each solve's oldest chunk, once finished; a solve that is done goes to
     * its writer, one that failed is dropped */
    for (int q = 0; q < n; q++) {
      request *r = act[q];
      int rc = 0;
      const int ready = vh_dsolve_ready(r->d);
      live[q] = 0;
      if (ready < 0) rc = -1;
      else if (ready == 1) {
        rc = vh_dsolve_complete(r->d);
        if (!rc && vh_dsolve_done(r->d)) {
          r->info.gpu_ms = vh_dsolve_loop_ms(r->d);
          double **td = NULL;
          vh_dsolve_collect(r->d, &td);
          pthread_mutex_lock(&sv->mu);
          r->td = td;
          r->info.t_gpu_done = now(sv);
          sv->used -= r->est;
          sv->active--;
          r->state = RQ_WRITE;
          pthread_cond_broadcast(&sv->cvWork);
          pthread_mutex_unlock(&sv->mu);
          busy = 1;
          continue;
        }
      }
      if (rc) {
        drop_failed(sv, r);
        continue;
      }
      live[q] = ready == 1 ? 1 : 2;  /* 2: oldest chunk still running */
    }
    /* new chunks for the first `running` places in least-work-left order: a
     * place is one solve, whose oldest chunk has finished, or a pack of up to
     * VGPU_GROUP_MAX small solves taken in that order, stepped together with
     * packed launches (vh_dsolve_enqueue_group); two chunks stay in flight */
    int firstSmall = 0;
    for (int q = 0, places = 0; q < n && places < sv->running; q++) {
      if (!live[q]) continue;
      request *r = act[q];
      long e = 0;
      /* with smallAlone, small solves (latency-bound, slowed most by a large
       * grid's long kernels between theirs) share the device only with each
       * other */
      if (places == 0) firstSmall = r->info.cells < sv->packCells;
      else if (sv->smallAlone && firstSmall && r->info.cells >= sv->packCells) break;
      places++;
      if (r->info.cells < sv->packCells) {
        vh_dsolve *pk[VGPU_GROUP_MAX];
        int idx[VGPU_GROUP_MAX], m = 0;
        for (int u = q; u < n && m < VGPU_GROUP_MAX; u++)
          if (live[u] && act[u]->info.cells < sv->packCells && vh_dsolve_remaining(act[u]->d) > 0 &&
              vh_dsolve_packable(act[u]->d)) {
            idx[m] = u;
            pk[m++] = act[u]->d;
          }
        if (m > 1) {
          for (int u = 0; u < m; u++) live[idx[u]] = 0;
          while ((e = vh_dsolve_enqueue_group(pk, m)) > 0) busy = 1;
          if (e < 0)
            for (int u = 0; u < m; u++) drop_failed(sv, act[idx[u]]);
          continue;
        }
      }
      const int ready = live[q] == 1;
      live[q] = 0;
      if (ready)
        while ((e = vh_dsolve_enqueue(r->d)) > 0) busy = 1;
      if (e < 0) drop_failed(sv, r);
    }
    if (!busy) usleep(n ? 20 : 200);
  }
  free(act);
  free(live);
  return NULL;
}

vh_server *vh_server_create(int device, int prec, int fullPrecision, double budgetBytes, int maxActive, int workers,
                            unsigned long chunk)
{
  vh_server *sv = calloc(1, sizeof(*sv));
  if (!sv) return NULL;
  sv->device = device;
  sv->prec = prec;
  sv->fullPrecision = fullPrecision;
  sv->budget = (size_t)budgetBytes;
  sv->maxActive = maxActive > 0 ? maxActive : 8;
  sv->nworkers = workers > 0 ? workers : 4;
  /* set-up solves kept ready beyond the active ones, so that a short request
   * arriving behind long ones is set up and can be admitted first */
  sv->lookahead = 2 * sv->maxActive < 32 ? 2 * sv->maxActive : 32;
  /* solves with chunks in flight at once; the rest wait resident.  Two
   * overlap on the device; one lowers the tail latency a little more and
   * shares nothing, four raise throughput and the tail with it;
   * VH_SERVE_RUNNING overrides */
  sv->running = sv->maxActive < 2 ? sv->maxActive : 2;
  const char *rn = getenv("VH_SERVE_RUNNING");
  if (rn && atoi(rn) > 0) sv->running = atoi(rn);
  sv->uploaders = sv->nworkers > 6 ? 3 : sv->nworkers > 4 ? 2 : 0;
  /* grids too small to fill the device alone share packed launches;
   * VH_SERVE_PACK_CELLS overrides (0 turns packing off) */
  sv->packCells = 1000000;
  const char *pc = getenv("VH_SERVE_PACK_CELLS");
  if (pc) sv->packCells = atoll(pc);
  /* a running small solve shares the device with small ones only: a large
   * grid's long kernels between its short ones slow it most (32 requests:
   * p99 latency / isolated 2.07 -> 1.91 at the same throughput);
   * VH_SERVE_SMALL_ALONE=0 lets it share with any */
  const char *sa = getenv("VH_SERVE_SMALL_ALONE");
  sv->smallAlone = !(sa && atoi(sa) == 0);
  sv->chunk = chunk ? chunk : 200;
  clock_gettime(CLOCK_MONOTONIC, &sv->t0);
  pthread_mutex_init(&sv->mu, NULL);
  pthread_cond_init(&sv->cvWork, NULL);
  pthread_cond_init(&sv->cvDone, NULL);
  sv->workers = calloc((size_t)sv->nworkers, sizeof(pthread_t));
  for (int w = 0; w < sv->nworkers; w++) pthread_create(&sv->workers[w], NULL, worker, sv);
  pthread_create(&sv->sched, NULL, scheduler, sv);
  return sv;
}

int vh_server_submit(vh_server *sv, const char *mesh, const char *outdir)
{
  request *r = calloc(1, sizeof(*r));
  if (!r) return -1;
  snprintf(r->mesh, sizeof(r->mesh), "%s", mesh);
  snprintf(r->outdir, sizeof(r->outdir), "%s", outdir ? outdir : "");
  pthread_mutex_lock(&sv->mu);
  if (sv->nrq == sv->cap) {
    int cap = sv->cap ? 2 * sv->cap : 64;
    request **n = realloc(sv->rq, (size_t)cap * sizeof(request *));
    if (!n) {
      pthread_mutex_unlock(&sv->mu);
      free(r);
      return -1;
    }
    sv->rq = n;
    sv->cap = cap;
  }
  r->info.id = sv->nrq;
  r->info.t_submit = now(sv);
  r->state = RQ_QUEUED;
  sv->rq[sv->nrq++] = r;
  pthread_cond_broadcast(&sv->cvWork);
  pthread_mutex_unlock(&sv->mu);
  return r->info.id;
}

int vh_server_wait(vh_server *sv, int id, vh_request_info *info)
{
  pthread_mutex_lock(&sv->mu);
  if (id < 0 || id >= sv->nrq) {
    pthread_mutex_unlock(&sv->mu);
    return -1;
  }
  request *r = sv->rq[id];
  while (r->state != RQ_DONE && r->state != RQ_FAILED) pthread_cond_wait(&sv->cvDone, &sv->mu);
  if (info) *info = r->info;
  const int rc = r->state == RQ_DONE ? 0 : -1;
  pthread_mutex_unlock(&sv->mu);
  return rc;
}

int vh_server_drain(vh_server *sv)
{
  int bad = 0;
  for (int q = 0;; q++) {
    pthread_mutex_lock(&sv->mu);
    const int n = sv->nrq;
    pthread_mutex_unlock(&sv->mu);
    if (q >= n) break;
    if (vh_server_wait(sv, q, NULL)) bad++;
  }
  return bad;
}

void vh_server_destroy(vh_server *sv)
{
  if (!sv) return;
  vh_server_drain(sv);
  pthread_mutex_lock(&sv->mu);
  sv->stop = 1;
  pthread_cond_broadcast(&sv->cvWork);
  pthread_mutex_unlock(&sv->mu);
  pthread_join(sv->sched, NULL);
  for (int w = 0; w < sv->nworkers; w++) pthread_join(sv->workers[w], NULL);
  for (int q = 0; q < sv->nrq; q++) free(sv->rq[q]);
  free(sv->rq);
  free(sv->workers);
  pthread_mutex_destroy(&sv->mu);
  pthread_cond_destroy(&sv->cvWork);
  pthread_cond_destroy(&sv->cvDone);
  free(sv);
}
