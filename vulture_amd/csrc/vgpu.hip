// C-ABI of the MI355X FDTD core (include/vgpu.h): context management,
// uploads from the reference's k-fastest layout to the device x-fastest
// layout, construction of the surface work lists (Debye sites, SIBC faces and
// gathered H corrections, source and TF/SF correction rows, observers) and
// dispatch to the per-precision step driver.
//
// This translation unit is compiled with -ffp-contract=off: the TF/SF
// geometry (distances, interpolation weights) is computed here on the host
// and must round exactly as the reference's set-up does.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <array>
#include <map>
#include <atomic>
#include <mutex>
#include <shared_mutex>
#include <unordered_map>
#include "vgpu_kernels.h"

namespace vgpu {

int set_error(Ctx *c, int code, const char *fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

void *dev_alloc(Ctx *c, size_t bytes)
{
  if (bytes == 0) bytes = 16;
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  if (hipMemset(p, 0, bytes) != hipSuccess) {
    hipFree(p);
    return nullptr;
  }
  c->allocs.push_back(p);
  c->bytes += bytes;
  return p;
}

template <typename V> static void *upload_vec(Ctx *c, const std::vector<V> &v)
{
  void *p = dev_alloc(c, v.size() * sizeof(V));
  if (p && !v.empty()) hipMemcpy(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice);
  return p;
}

// Upload a host k-fastest [NX][NY][NZ] double array into an x-fastest device
// array of the context's precision.
template <typename S> static int upload_field(Ctx *c, void *dst, const S *src)
{
  const size_t n = c->ncell;
  std::vector<unsigned char> buf(n * c->esz, 0);
  const long long NY = c->NY, NZ = c->NZ;
  for (int k = 0; k < c->NZ; ++k)
    for (int j = 0; j < c->NY; ++j) {
      const size_t row = ((size_t)k * c->NY + j) * c->PX;
      for (int i = 0; i < c->NX; ++i) {
        const double v = (double)src[((size_t)i * NY + j) * NZ + k];
        if (c->esz == 8) ((double *)buf.data())[row + i] = v;
        else ((float *)buf.data())[row + i] = (float)v;
      }
    }
  if (hipMemcpy(dst, buf.data(), n * c->esz, hipMemcpyHostToDevice) != hipSuccess)
    return set_error(c, VGPU_EDEVICE, "upload failed");
  return VGPU_OK;
}

// Coefficient classes: the distinct N-tuples of per-cell coefficients, in
// first-seen order over the host layout, and the class of every device cell
// (padding cells stay in class 0).
template <int N> struct TupleHash {
  size_t operator()(const std::array<uint64_t, N> &k) const
  {
    uint64_t h = 1469598103934665603ull;
    for (int q = 0; q < N; ++q) h = (h ^ k[q]) * 1099511628211ull + (k[q] >> 29);
    return (size_t)h;
  }
};

template <int N, typename S>
static void build_classes(Ctx *c, const S *const src[N], std::vector<uint32_t> &cls, std::vector<double> &tab)
{
  cls.assign(c->ncell, 0);
  tab.clear();
  std::unordered_map<std::array<uint64_t, N>, uint32_t, TupleHash<N>> ids;
  std::array<uint64_t, N> prev{};
  uint32_t prevId = UINT32_MAX;
  const size_t NY = c->NY, NZ = c->NZ;
  for (int i = 0; i < c->NX; ++i)
    for (int j = 0; j < c->NY; ++j)
      for (int k = 0; k < c->NZ; ++k) {
        const size_t h = ((size_t)i * NY + j) * NZ + k;
        std::array<uint64_t, N> key;
        for (int q = 0; q < N; ++q) {
          const double v = (double)src[q][h];  // exact for float input
          memcpy(&key[q], &v, 8);
        }
        uint32_t id;
        if (prevId != UINT32_MAX && key == prev) {
          id = prevId;
        } else {
          auto it = ids.find(key);
          if (it == ids.end()) {
            id = (uint32_t)ids.size();
            ids.emplace(key, id);
            for (int q = 0; q < N; ++q) tab.push_back((double)src[q][h]);
          } else {
            id = it->second;
          }
          prev = key;
          prevId = id;
        }
        cls[c->lin(i, j, k)] = id;
      }
}

static void *upload_line(Ctx *c, const double *src, int n)
{
  void *p = dev_alloc(c, (size_t)n * c->esz);
  if (!p) return nullptr;
  if (c->esz == 8) {
    hipMemcpy(p, src, (size_t)n * 8, hipMemcpyHostToDevice);
  } else {
    std::vector<float> f(n);
    for (int i = 0; i < n; ++i) f[i] = (float)src[i];
    hipMemcpy(p, f.data(), (size_t)n * 4, hipMemcpyHostToDevice);
  }
  return p;
}

static void *upload_reals(Ctx *c, const std::vector<double> &v)
{
  return upload_line(c, v.data(), (int)v.size());
}

static bool box_empty(const int *b) { return b[1] < b[0] || b[3] < b[2] || b[5] < b[4]; }

static long long box_cells(const int *b)
{
  if (box_empty(b)) return 0;
  return (long long)(b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1);
}

// ---- TF/SF geometry (planewave.c:557-573, 1375-1419; grid.c:1397-1445) ----

static void field_index_location(double r[3], int field, int i, int j, int k)
{
  r[0] = i; r[1] = j; r[2] = k;
  switch (field) {
  case 0: r[0] = i + 0.5; break;
  case 1: r[1] = j + 0.5; break;
  case 2: r[2] = k + 0.5; break;
  case 3: r[1] = j + 0.5; r[2] = k + 0.5; break;
  case 4: r[0] = i + 0.5; r[2] = k + 0.5; break;
  case 5: r[0] = i + 0.5; r[1] = j + 0.5; break;
  }
}

static void field_physical_location(const Ctx *c, double r[3], int field, int i, int j, int k)
{
  const int im = i - c->grid.gibox[0], jm = j - c->grid.gibox[2], km = k - c->grid.gibox[4];
  r[0] = c->xl[im]; r[1] = c->yl[jm]; r[2] = c->zl[km];
  switch (field) {
  case 0: r[0] = c->xl[im] + 0.5 * c->dex[i]; break;
  case 1: r[1] = c->yl[jm] + 0.5 * c->dey[j]; break;
  case 2: r[2] = c->zl[km] + 0.5 * c->dez[k]; break;
  case 3: r[1] = c->yl[jm] + 0.5 * c->dey[j]; r[2] = c->zl[km] + 0.5 * c->dez[k]; break;
  case 4: r[0] = c->xl[im] + 0.5 * c->dex[i]; r[2] = c->zl[km] + 0.5 * c->dez[k]; break;
  case 5: r[0] = c->xl[im] + 0.5 * c->dex[i]; r[1] = c->yl[jm] + 0.5 * c->dey[j]; break;
  }
}

struct RowEntry { long long key; long long seq; CorrEntry e; };

}  // namespace vgpu

using namespace vgpu;

struct vgpu_ctx : public Ctx {};

// Scratch of the entry point that allocates nothing (vgpu_check_limits_noalloc):
// one file-scope device array, present on every device the library runs on
// until the process ends.  Host code never uses its name as an address: the
// device address of each device's copy comes from hipGetSymbolAddress, taken
// once per device with that device current (scratch_setup, called by
// vgpu_create) and kept in g_scratch.  Every call on one device shares that
// device's copy; calls that overlap on one device race.
static __device__ __align__(256) unsigned char scratch[VGPU_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void *g_scratch[64];

// The device address of `device`'s scratch, looked up once (device current).
static int scratch_setup(int device)
{
  if (device < 0 || device >= 64) return VGPU_EINVAL;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return VGPU_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
    (void)hipGetLastError();
    return VGPU_EDEVICE;
  }
  g_scratch[device] = p;
  return VGPU_OK;
}

static void *scratch_of_device(int device)
{
  if (device < 0 || device >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[device];
}

// A stream capture must not overlap a synchronization of the whole device
// (HIP follows CUDA here: synchronizing a device while any of its streams is
// being captured is invalid, and hipFree and hipDeviceSynchronize synchronize
// the device).  One solve's thread may capture while other threads set up,
// read back or destroy other solves, so the entry points that synchronize
// the device or the null stream -- allocation, synchronous copies, frees --
// hold this lock shared, and a capture holds it alone.
static std::shared_mutex g_capture_mu;
using DeviceCalls = std::shared_lock<std::shared_mutex>;

template <typename S> static int set_media(Ctx *c, const S *const alpha[3], const S *const beta[3],
                                           const S *const gamma[3])
{
  if (!c || c->finalized) return VGPU_ESTATE;
  for (int d = 0; d < 3; ++d)
    if (!alpha[d] || !beta[d] || !gamma[d]) return set_error(c, VGPU_EINVAL, "missing coefficient array");
  const DeviceCalls guard(g_capture_mu);
  for (int d = 0; d < 3; ++d) {
    int r;
    if ((r = upload_field(c, c->alpha[d], alpha[d])) || (r = upload_field(c, c->beta[d], beta[d])) ||
        (r = upload_field(c, c->gamma[d], gamma[d])))
      return r;
  }
  const S *e6[6] = {alpha[0], beta[0], alpha[1], beta[1], alpha[2], beta[2]};
  build_classes<6, S>(c, e6, c->ecls, c->etab);
  build_classes<3, S>(c, gamma, c->hcls, c->htab);
  c->edslot.assign(c->etab.size() / 6 * 3, -1);
  c->media_set = true;
  return VGPU_OK;
}

extern "C" {

static void destroy_ctx(vgpu_ctx *c);

// The last run may have gone to a caller's stream (vgpu_run's stream
// argument): wait for its end event as well as for the context's stream
// before reading results or freeing memory.  An event never recorded
// counts as complete.
static void wait_last_run(Ctx *c)
{
  hipEventSynchronize(c->ev1);
  hipStreamSynchronize(c->stream);
}

int vgpu_create(const vgpu_grid_desc *g, vgpu_prec prec, int device, vgpu_ctx **out)
{
  if (!g || !out) return VGPU_EINVAL;
  const DeviceCalls guard(g_capture_mu);
  *out = nullptr;
  if (g->nx < 3 || g->ny < 3 || g->nz < 3) return VGPU_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return VGPU_EDEVICE;
  // this device's no-allocation scratch address, with the device current
  if (scratch_setup(device) != VGPU_OK) return VGPU_EDEVICE;
  vgpu_ctx *c = new vgpu_ctx();
  static std::atomic<uint64_t> serials{0};
  c->serial = ++serials;
  c->prec = prec;
  c->device = device;
  c->esz = prec == VGPU_FP64 ? 8 : 4;
  c->grid = *g;
  c->NX = g->nx; c->NY = g->ny; c->NZ = g->nz;
  const int align = (int)(128 / c->esz);
  c->PX = (c->NX + align - 1) / align * align;
  c->plane = (long long)c->PX * c->NY;
  c->ncell = (size_t)c->plane * c->NZ;
  auto cp = [](std::vector<double> &v, const double *p, int n) { v.assign(p, p + n); };
  cp(c->xl, g->xlines, g->nlines[0]); cp(c->yl, g->ylines, g->nlines[1]); cp(c->zl, g->zlines, g->nlines[2]);
  cp(c->dex, g->dex, c->NX); cp(c->dey, g->dey, c->NY); cp(c->dez, g->dez, c->NZ);
  cp(c->idex, g->idex, c->NX); cp(c->idey, g->idey, c->NY); cp(c->idez, g->idez, c->NZ);
  cp(c->idhx, g->idhx, c->NX); cp(c->idhy, g->idhy, c->NY); cp(c->idhz, g->idhz, c->NZ);
  c->grid.xlines = c->grid.ylines = c->grid.zlines = nullptr;
  c->grid.dex = c->grid.dey = c->grid.dez = nullptr;
  c->grid.idex = c->grid.idey = c->grid.idez = nullptr;
  c->grid.idhx = c->grid.idhy = c->grid.idhz = nullptr;
  {
    const int N[3] = {c->NX, c->NY, c->NZ};
    // The fused leapfrog sweep with VGPU_LEAP=1 (not with the fused Debye
    // layer, VGPU_DEBYE_FUSION=1, which only the separate sweeps have; its
    // kernel indexes cells with 32-bit offsets).  The separate E and H
    // sweeps are the default: on the Debye volume at 256^3 the fused sweep
    // and the H cells it leaves to k_h_fix measured slower.
    const char *lp = getenv("VGPU_LEAP"), *df = getenv("VGPU_DEBYE_FUSION");
    c->leap = lp && lp[0] == '1' && !(df && df[0] == '1') && c->ncell + FIELD_DISCARD < (1ull << 31);
    if (c->leap) {
      int fb[6];
      leap_box(g->gfilim, N, fb);
      for (int q = 0; q < 6; ++q) c->fastE[q] = c->fastH[q] = fb[q];
      c->leapR = 4;
      if (const char *ev = getenv("VGPU_LEAP_ROWS")) c->leapR = atoi(ev) <= 1 ? 1 : atoi(ev) == 2 ? 2 : 4;  // tuning override
      if (const char *ev = getenv("VGPU_LEAP_KZ")) c->leapKZ = std::max(1, std::min(atoi(ev), 64));
    } else {
      // x range of the fast boxes on whole 128-byte lines (VGPU_ALIGN_FAST=0 keeps the exact boxes)
      const char *al = getenv("VGPU_ALIGN_FAST");
      fast_boxes(g->gfilim, N, c->fastE, c->fastH, al && al[0] == '0' ? 0 : (int)(128 / c->esz));
    }
  }
  c->NCH = c->fastE[0] <= c->fastE[1] ? (c->fastE[1] - c->fastE[0] + 64) / 64 : 0;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return VGPU_EDEVICE;
  }
  hipEventCreate(&c->ev0);
  hipEventCreate(&c->ev1);
  for (auto &rr : c->runs) {
    hipEventCreate(&rr.start);
    hipEventCreate(&rr.end);
  }
  // VGPU_FIELD_STAGGER=<bytes> starts component d's array d times that many
  // bytes (rounded to 256) into its allocation: a probe of how the arrays'
  // relative placement in the HBM channels affects the sweeps (default 0).
  const char *sg = getenv("VGPU_FIELD_STAGGER");
  const size_t stagger = sg ? ((size_t)strtoull(sg, nullptr, 10) + 255) / 256 * 256 : 0;
  for (int d = 0; d < 6; ++d) {
    const size_t off = (size_t)d * stagger, bytes = (c->ncell + FIELD_DISCARD) * c->esz + off;
    void *f = dev_alloc(c, bytes);
    void *f1 = c->leap ? dev_alloc(c, bytes) : f;
    if (!f || !f1) {
      destroy_ctx(c);
      return VGPU_ENOMEM;
    }
    c->F[d] = (char *)f + off;
    c->F1[d] = (char *)f1 + off;
  }
  for (int d = 0; d < 3; ++d) {
    c->alpha[d] = dev_alloc(c, c->ncell * c->esz);
    c->beta[d] = dev_alloc(c, c->ncell * c->esz);
    c->gamma[d] = dev_alloc(c, c->ncell * c->esz);
    if (!c->alpha[d] || !c->beta[d] || !c->gamma[d]) {
      destroy_ctx(c);
      return VGPU_ENOMEM;
    }
  }
  c->ide[0] = upload_reals(c, c->idex); c->ide[1] = upload_reals(c, c->idey); c->ide[2] = upload_reals(c, c->idez);
  c->idh[0] = upload_reals(c, c->idhx); c->idh[1] = upload_reals(c, c->idhy); c->idh[2] = upload_reals(c, c->idhz);
  for (int r = 0; r < 6; ++r)
    for (int f = 0; f < 6; ++f)
      for (int q = 0; q < 6; ++q) { c->fplim[r][f][q] = 0; c->fmlim[r][f][q] = 0; }
  const char *ev = getenv("VGPU_KERNEL_EVENTS");
  int nev = ev ? atoi(ev) : 0;
  for (int q = 0; q < 2 * nev; ++q) {
    hipEvent_t e1, e2;
    hipEventCreate(&e1);
    hipEventCreate(&e2);
    c->evE.push_back(e1);
    c->evH.push_back(e2);
  }
  *out = c;
  return VGPU_OK;
}

int vgpu_set_media(vgpu_ctx *c, const double *alpha[3], const double *beta[3], const double *gamma[3])
{
  return set_media<double>(c, alpha, beta, gamma);
}

int vgpu_set_media_f32(vgpu_ctx *c, const float *alpha[3], const float *beta[3], const float *gamma[3])
{
  return set_media<float>(c, alpha, beta, gamma);
}

int vgpu_set_pml(vgpu_ctx *c, const vgpu_pml_desc *p)
{
  const DeviceCalls guard(g_capture_mu);
  if (!c || c->finalized || !p) return VGPU_ESTATE;
  c->pmlmask = p->region_mask & 0x3Fu;
  memcpy(c->fplim, p->fplim, sizeof(c->fplim));
  // Slab offsets within the pools, in elements, each slab 256-byte aligned.
  // The shell kernels index the pools with 32-bit element offsets.
  const long long al = 256 / (long long)c->esz;
  long long offE[6][3] = {}, offH[6][3] = {}, totE = 0, totH = 0;
  for (int r = 0; r < 6; ++r) {
    if (!((c->pmlmask >> r) & 1u)) continue;
    for (int d = 0; d < 3; ++d) {
      offE[r][d] = totE;
      offH[r][d] = totH;
      totE += (box_cells(c->fplim[r][d]) + al - 1) / al * al;
      totH += (box_cells(c->fplim[r][3 + d]) + al - 1) / al * al;
    }
  }
  if (totE > 0x7fffffffLL || totH > 0x7fffffffLL)
    return set_error(c, VGPU_EINVAL, "UPML auxiliaries exceed 2^31 elements");
  if (c->pmlmask) {
    for (int q = 0; q < 3; ++q) {
      c->auxPool[q] = dev_alloc(c, (size_t)(q < 2 ? totE : totH) * c->esz);
      if (!c->auxPool[q]) return set_error(c, VGPU_ENOMEM, "UPML allocation");
    }
    for (int r = 0; r < 6; ++r) {
      if (!((c->pmlmask >> r) & 1u)) continue;
      for (int d = 0; d < 3; ++d) {
        c->aux1[r][d] = (char *)c->auxPool[0] + offE[r][d] * c->esz;
        c->aux2[r][d] = (char *)c->auxPool[1] + offE[r][d] * c->esz;
        c->auxB[r][d] = (char *)c->auxPool[2] + offH[r][d] * c->esz;
      }
    }
  }
  const int n[3] = {c->NX, c->NY, c->NZ};
  for (int d = 0; d < 3; ++d) {
    const double *src[6] = {p->ad[d], p->bd[d], p->ah[d], p->bh[d], p->ibd[d], p->ibh[d]};
    for (int q = 0; q < 6; ++q) {
      if (!src[q]) return set_error(c, VGPU_EINVAL, "missing UPML profile array");
      c->prof[q][d] = upload_line(c, src[q], n[d]);
    }
  }
  return VGPU_OK;
}

int vgpu_set_mur(vgpu_ctx *c, unsigned face_mask, const double zeta[6], const int fmlim[6][6][6])
{
  if (!c || c->finalized) return VGPU_ESTATE;
  c->murmask = face_mask & 0x3Fu;
  for (int f = 0; f < 6; ++f) c->zeta[f] = zeta[f];
  memcpy(c->fmlim, fmlim, sizeof(c->fmlim));
  return VGPU_OK;
}

int vgpu_set_ghosts(vgpu_ctx *c, const int face_type[6])
{
  if (!c || c->finalized) return VGPU_ESTATE;
  for (int f = 0; f < 6; ++f) {
    if (face_type[f] < 0 || face_type[f] > 3) return set_error(c, VGPU_EINVAL, "bad ghost type");
    c->ghost[f] = face_type[f];
  }
  return VGPU_OK;
}

int vgpu_add_debye_block(vgpu_ctx *c, const int flim[3][6], int npoles, const double *dalpha, const double *dbeta)
{
  if (!c || c->finalized) return VGPU_ESTATE;
  if (npoles < 0 || (npoles > 0 && (!dalpha || !dbeta))) return set_error(c, VGPU_EINVAL, "bad Debye block");
  Ctx::DebyeBlock b;
  memcpy(b.flim, flim, sizeof(b.flim));
  b.np = npoles;
  b.da.assign(dalpha, dalpha + 2 * npoles);
  b.db.assign(dbeta, dbeta + 2 * npoles);
  c->debye.push_back(b);
  return VGPU_OK;
}

int vgpu_add_sibc_surface(vgpu_ctx *c, const vgpu_sibc_desc *d)
{
  if (!c || c->finalized || !d) return VGPU_ESTATE;
  Ctx::SibcSurf s;
  s.d = *d;
  const long long nf = (long long)(d->gbbox[1] - d->gbbox[0]) * (d->gbbox[3] - d->gbbox[2]) * (d->gbbox[5] - d->gbbox[4]);
  if (nf <= 0 || d->normal < 0 || d->normal > 2) return set_error(c, VGPU_EINVAL, "bad SIBC surface");
  s.adjA.assign(d->is_adj_a, d->is_adj_a + 4 * nf);
  s.adjB.assign(d->is_adj_b, d->is_adj_b + 4 * nf);
  for (int f = 0; f < 16; ++f) {
    const int np = d->npoles[f];
    if (np < 0) return set_error(c, VGPU_EINVAL, "bad SIBC filter");
    if (np > 0) {
      s.q0[f].assign(d->q0[f], d->q0[f] + 2 * np);
      s.q1[f].assign(d->q1[f], d->q1[f] + 2 * np);
      s.q2[f].assign(d->q2[f], d->q2[f] + 2 * np);
    }
    c->sibcMaxPoles = std::max(c->sibcMaxPoles, np);
  }
  s.d.is_adj_a = s.d.is_adj_b = nullptr;
  for (int f = 0; f < 16; ++f) s.d.q0[f] = s.d.q1[f] = s.d.q2[f] = nullptr;
  c->sibc.push_back(std::move(s));
  c->sibcFaces += nf;
  return VGPU_OK;
}

int vgpu_add_waveform(vgpu_ctx *c, int type, double size, double delay, double width, double frequency,
                      const double *table, unsigned long n)
{
  if (!c || c->finalized) return VGPU_ESTATE;
  if (type < 0 || type > 9) return set_error(c, VGPU_EINVAL, "bad waveform type %d", type);
  WaveformDev w{};
  w.type = type; w.size = size; w.delay = delay; w.width = width; w.frequency = frequency;
  w.n = 0;
  c->wftab.emplace_back();
  if (type == 9) {
    if (!table || n < 2) return set_error(c, VGPU_EINVAL, "external waveform needs a table");
    c->wftab.back().assign(table, table + 3 * n);
    w.n = n;
  }
  c->wf.push_back(w);
  return (int)c->wf.size() - 1;
}

int vgpu_add_source(vgpu_ctx *c, int field, int kind, int soft, double size, int waveform, double delay,
                    const int flim[6])
{
  if (!c || c->finalized) return VGPU_ESTATE;
  if (field < 0 || field > 5 || kind < 0 || kind > 1 || waveform < 0 || waveform >= (int)c->wf.size())
    return set_error(c, VGPU_EINVAL, "bad source");
  Ctx::Src s{};
  s.field = field; s.kind = kind; s.soft = soft ? 1 : 0; s.size = size; s.waveform = waveform; s.delay = delay;
  memcpy(s.flim, flim, sizeof(s.flim));
  c->src.push_back(s);
  return (int)c->src.size() - 1;
}

int vgpu_add_planewave(vgpu_ctx *c, const vgpu_pw_desc *d)
{
  if (!c || c->finalized || !d) return VGPU_ESTATE;
  if (d->waveform < 0 || d->waveform >= (int)c->wf.size()) return set_error(c, VGPU_EINVAL, "bad plane-wave waveform");
  Ctx::Pw p{};
  p.d = *d;
  if (d->use_aux) {
    if (d->aux_nx < 4 || d->aux_xb < 1 || d->aux_xb > d->aux_nx || d->aux_nx - d->aux_xb > 10)
      return set_error(c, VGPU_EINVAL, "bad auxiliary grid");
    const double *pr[4] = {d->aux_adx, d->aux_bdx, d->aux_ahx, d->aux_bhx};
    for (int q = 0; q < 4; ++q)
      for (int i = 0; i < 10; ++i) p.prof[q][i] = pr[q] ? pr[q][i] : 0.0;
  }
  p.d.aux_adx = p.d.aux_bdx = p.d.aux_ahx = p.d.aux_bhx = nullptr;
  c->pw.push_back(p);
  return (int)c->pw.size() - 1;
}

int vgpu_add_point_observer(vgpu_ctx *c, int i, int j, int k, uint64_t first, uint64_t last, int nfreq,
                            const double *omega)
{
  if (!c || c->finalized) return VGPU_ESTATE;
  if (i < 0 || j < 0 || k < 0 || i >= c->NX || j >= c->NY || k >= c->NZ || nfreq < 0 || (nfreq && !omega))
    return set_error(c, VGPU_EINVAL, "bad point observer");
  Ctx::PObs o;
  o.i = i; o.j = j; o.k = k; o.first = first; o.last = last;
  if (nfreq) o.omega.assign(omega, omega + nfreq);
  c->pobs.push_back(o);
  return (int)c->pobs.size() - 1;
}

int vgpu_add_volume_observer(vgpu_ctx *c, const int gbbox[6], const int step[3], uint64_t first, uint64_t last)
{
  if (!c || c->finalized) return VGPU_ESTATE;
  for (int d = 0; d < 3; ++d)
    if (step[d] <= 0 || gbbox[2 * d] < 0 || gbbox[2 * d + 1] >= (d == 0 ? c->NX : d == 1 ? c->NY : c->NZ) ||
        gbbox[2 * d] > gbbox[2 * d + 1])
      return set_error(c, VGPU_EINVAL, "bad volume observer");
  Ctx::VObs o{};
  memcpy(o.box, gbbox, sizeof(o.box));
  memcpy(o.step, step, sizeof(o.step));
  o.first = first; o.last = last;
  o.nodes = 1;
  for (int d = 0; d < 3; ++d) o.nodes *= 1 + (gbbox[2 * d + 1] - gbbox[2 * d]) / step[d];
  c->vobs.push_back(o);
  return (int)c->vobs.size() - 1;
}

// ---------------------------------------------------------------------------
// finalize: build device lists
// ---------------------------------------------------------------------------

static bool in_fast(const int *fb, int i, int j, int k)
{
  return i >= fb[0] && i <= fb[1] && j >= fb[2] && j <= fb[3] && k >= fb[4] && k <= fb[5];
}

static int build_debye(Ctx *c)
{
  if (c->debye.empty()) return VGPU_OK;
  // Distinct Debye media (blocks of one medium share their coefficients).
  int maxP = 1, cplx = 0;
  for (auto &b : c->debye) {
    maxP = std::max(maxP, b.np);
    for (int p = 0; p < b.np; ++p)
      if (b.da[2 * p + 1] != 0.0 || b.db[2 * p + 1] != 0.0) cplx = 1;
  }
  c->debyeMaxPoles = maxP;
  c->debyeCplx = cplx;
  std::map<std::vector<double>, int> mediumOf;
  std::vector<int> blockMedium(c->debye.size());
  std::vector<double> da, db;
  std::vector<int> np;
  for (size_t b = 0; b < c->debye.size(); ++b) {
    const auto &blk = c->debye[b];
    std::vector<double> key;
    key.push_back(blk.np);
    key.insert(key.end(), blk.da.begin(), blk.da.begin() + 2 * blk.np);
    key.insert(key.end(), blk.db.begin(), blk.db.begin() + 2 * blk.np);
    auto it = mediumOf.find(key);
    if (it == mediumOf.end()) {
      const int m = (int)np.size();
      it = mediumOf.emplace(key, m).first;
      np.push_back(blk.np);
      da.resize((size_t)(m + 1) * maxP * 2, 0.0);
      db.resize(da.size(), 0.0);
      for (int p = 0; p < blk.np; ++p)
        for (int q = 0; q < 2; ++q) {
          da[((size_t)m * maxP + p) * 2 + q] = blk.da[2 * p + q];
          db[((size_t)m * maxP + p) * 2 + q] = blk.db[2 * p + q];
        }
    }
    blockMedium[b] = it->second;
  }
  // Sites in block order; a cell already covered by an earlier block goes to
  // the next layer so corrections apply in block order (debye.c:243-307
  // loops block by block).
  struct Site { int layer, comp; long long cell; int med; };
  std::vector<Site> sites;
  std::unordered_map<long long, int> seen;
  int layers = 1;
  for (size_t b = 0; b < c->debye.size(); ++b)
    for (int f = 0; f < 3; ++f) {
      const int *l = c->debye[b].flim[f];
      for (int i = l[0]; i <= l[1]; ++i)
        for (int j = l[2]; j <= l[3]; ++j)
          for (int k = l[4]; k <= l[5]; ++k) {
            if (i < 0 || j < 0 || k < 0 || i >= c->NX || j >= c->NY || k >= c->NZ)
              return set_error(c, VGPU_EINVAL, "Debye block outside the grid");
            const long long cell = c->lin(i, j, k);
            int &cnt = seen[cell * 3 + f];
            // layer 0: first block of a cell inside the fast E box (applied
            // by the fast sweep); layer 1: first block elsewhere; then the
            // later blocks of a cell, one layer each
            const int layer = cnt == 0 ? (in_fast(c->fastE, i, j, k) ? 0 : 1) : cnt + 1;
            ++cnt;
            layers = std::max(layers, layer + 1);
            sites.push_back({layer, f, cell, blockMedium[b]});
          }
    }
  seen.clear();
  if (c->leap)
    for (const Site &st : sites) c->eTouched.push_back(st.cell * 3 + st.comp);
  std::sort(sites.begin(), sites.end(), [](const Site &a, const Site &b) {
    if (a.layer != b.layer) return a.layer < b.layer;
    if (a.comp != b.comp) return a.comp < b.comp;
    return a.cell < b.cell;
  });
  const long long n = (long long)sites.size();
  std::vector<int> comp(n), med(n);
  std::vector<long long> cell(n);
  c->debyeLayerStart.assign(layers + 1, n);
  for (long long s = n - 1; s >= 0; --s) c->debyeLayerStart[sites[s].layer] = s;
  for (int l = layers - 1; l >= 0; --l)
    if (c->debyeLayerStart[l] > c->debyeLayerStart[l + 1]) c->debyeLayerStart[l] = c->debyeLayerStart[l + 1];
  for (long long s = 0; s < n; ++s) { comp[s] = sites[s].comp; cell[s] = sites[s].cell; med[s] = sites[s].med; }
  c->debyeLayers = layers;
  c->debyeSites = n;
  c->debyeL0 = layers > 1 ? c->debyeLayerStart[1] : n;

  // Refine the E coefficient classes with each component's layer-0 Debye
  // medium, so the E sweep can apply layer 0 itself.
  std::unordered_map<long long, std::array<int, 3>> slot;
  for (long long s = 0; s < c->debyeL0; ++s) {
    auto it = slot.find(cell[s]);
    if (it == slot.end()) it = slot.emplace(cell[s], std::array<int, 3>{-1, -1, -1}).first;
    it->second[comp[s]] = med[s];
  }
  std::map<std::array<int, 4>, uint32_t> refined;
  for (auto &kv : slot) {
    const uint32_t old = c->ecls[kv.first];
    const std::array<int, 4> key{(int)old, kv.second[0], kv.second[1], kv.second[2]};
    auto it = refined.find(key);
    if (it == refined.end()) {
      const uint32_t id = (uint32_t)(c->etab.size() / 6);
      for (int q = 0; q < 6; ++q) c->etab.push_back(c->etab[(size_t)old * 6 + q]);
      for (int q = 0; q < 3; ++q) c->edslot.push_back(kv.second[q]);
      it = refined.emplace(key, id).first;
    }
    c->ecls[kv.first] = it->second;
  }
  // First layer-0 site of every (component, row, 64-cell x chunk of the fast
  // box), for the fast sweep's in-wavefront site numbering.
  const size_t nrow = (size_t)c->NZ * c->NY * c->NCH;
  std::vector<int> dbase(3 * nrow + 1, 0);
  for (long long s = 0; s < c->debyeL0; ++s) {
    const long long cl = cell[s];
    const int i = (int)(cl % c->PX);
    const long long row = cl / c->PX;  // k*NY + j
    const size_t q = (size_t)comp[s] * nrow + (size_t)row * c->NCH + (i - c->fastE[0]) / 64;
    dbase[q + 1]++;
  }
  if (c->debyeL0 > INT32_MAX) return set_error(c, VGPU_EINVAL, "too many Debye sites");
  for (size_t q = 1; q < dbase.size(); ++q) dbase[q] += dbase[q - 1];
  // dbase[q] is the first site of chunk q; dbase[3*nrow] the layer-0 count.
  // The E sweep numbers layer-0 sites by counting, in x order within each
  // 64-cell chunk, the cells whose class carries a Debye medium for the
  // component.  Check here that this numbering reproduces the site list
  // exactly, so the sweep can never address a site outside it.
  const int *fb = c->fastE;
  for (int f = 0; f < 3; ++f) {
    long long g = dbase[(size_t)f * nrow];
    for (int k = fb[4]; k <= fb[5]; ++k)
      for (int j = fb[2]; j <= fb[3]; ++j)
        for (int i = fb[0]; i <= fb[1]; ++i) {
          const long long cl = c->lin(i, j, k);
          const size_t row = (size_t)f * nrow + ((size_t)k * c->NY + j) * c->NCH + (i - fb[0]) / 64;
          if ((i - fb[0]) % 64 == 0 && dbase[row] != g)
            return set_error(c, VGPU_EINVAL, "Debye site numbering (chunk start)");
          const int ds = c->edslot[(size_t)c->ecls[cl] * 3 + f];
          if (ds < 0) continue;
          if (g >= c->debyeL0 || cell[g] != cl || comp[g] != f || med[g] != ds)
            return set_error(c, VGPU_EINVAL, "Debye site numbering (site %lld)", g);
          ++g;
        }
    const long long end = dbase[(size_t)(f + 1) * nrow];
    if (g != end) return set_error(c, VGPU_EINVAL, "Debye site numbering (component %d count)", f);
  }
  c->d_dbase = upload_vec(c, dbase);

  std::vector<long long> key(n);
  for (long long g = 0; g < n; ++g) {
    if (cell[g] >= (1LL << 40) || med[g] < 0 || med[g] >= (1 << 21))
      return set_error(c, VGPU_EINVAL, "Debye site index out of range");
    key[g] = (long long)med[g] << 42 | cell[g] << 2 | comp[g];
  }
  c->d_dkey = upload_vec(c, key);
  c->d_dsbeta = dev_alloc(c, (size_t)n * c->esz);
  c->d_dlast = dev_alloc(c, (size_t)n * c->esz);
  c->d_dJ = dev_alloc(c, (size_t)n * maxP * (cplx ? 2 : 1) * c->esz);
  c->d_dalpha = upload_vec(c, da);
  c->d_dbeta = upload_vec(c, db);
  c->d_dnp = upload_vec(c, np);
  if (!c->d_dkey || !c->d_dsbeta || !c->d_dlast || !c->d_dJ || !c->d_dbase)
    return set_error(c, VGPU_ENOMEM, "Debye state allocation");
  return VGPU_OK;
}

// Upload the coefficient classes when both tables fit 16-bit indices.
static int build_class_tables(Ctx *c)
{
  c->nEcls = (int)(c->etab.size() / 6);
  c->nHcls = (int)(c->htab.size() / 3);
  c->indexed = c->nEcls <= 65536 && c->nHcls <= 65536 && getenv("VGPU_NO_INDEXED_MEDIA") == nullptr;
  if (c->indexed) {
    std::vector<uint16_t> e(c->ncell), h(c->ncell);
    for (size_t q = 0; q < c->ncell; ++q) { e[q] = (uint16_t)c->ecls[q]; h[q] = (uint16_t)c->hcls[q]; }
    c->d_ecls = upload_vec(c, e);
    c->d_hcls = upload_vec(c, h);
    // padded to ETAB_STRIDE / HTAB_STRIDE entries per class (vector loads)
    std::vector<double> e8((size_t)c->nEcls * ETAB_STRIDE, 0.0), h4((size_t)c->nHcls * HTAB_STRIDE, 0.0);
    for (int k = 0; k < c->nEcls; ++k)
      for (int q = 0; q < 6; ++q) e8[(size_t)k * ETAB_STRIDE + q] = c->etab[(size_t)k * 6 + q];
    for (int k = 0; k < c->nHcls; ++k)
      for (int q = 0; q < 3; ++q) h4[(size_t)k * HTAB_STRIDE + q] = c->htab[(size_t)k * 3 + q];
    c->d_etab = upload_reals(c, e8);
    c->d_htab = upload_reals(c, h4);
    c->d_edslot = upload_vec(c, c->edslot);
    if (!c->d_ecls || !c->d_hcls || !c->d_etab || !c->d_htab || !c->d_edslot)
      return set_error(c, VGPU_ENOMEM, "coefficient class allocation");
  }
  // Applying the first Debye layer inside the fast E kernel saves the
  // separate pass's field traffic but serialises three dependent table
  // look-ups per component on Debye rows; measured on the Debye-sphere
  // workload it is slower, so the separate pass is the default and the
  // fused one is kept behind VGPU_DEBYE_FUSION=1.
  const char *df = getenv("VGPU_DEBYE_FUSION");
  c->debyeFused = c->indexed && c->debyeSites > 0 && df && df[0] == '1';
  std::vector<uint32_t>().swap(c->ecls);
  std::vector<uint32_t>().swap(c->hcls);
  return VGPU_OK;
}

static int build_sibc(Ctx *c)
{
  if (c->sibc.empty()) return VGPU_OK;
  const int maxP = c->sibcMaxPoles;
  const long long NF = c->sibcFaces;
  const int ns = (int)c->sibc.size();
  std::vector<int> ijk(3 * NF), surf(NF), normal(ns), set(ns), npl(16 * ns);
  std::vector<unsigned char> adjA(4 * NF), adjB(4 * NF);
  std::vector<double> A(16 * ns), B(16 * ns), asy(16 * ns), q((size_t)16 * ns * maxP * 6, 0.0);
  // Gathered H corrections: (cell key, sequence) -> entry.
  struct HE { long long key; long long seq; SibcHEnt e; };
  std::vector<HE> hents;
  std::vector<long long> zeroKeys;
  long long f = 0, seq = 0;
  const std::vector<double> *ide[3] = {&c->idex, &c->idey, &c->idez};
  std::vector<SibcBlock> blocks;
  for (int s = 0; s < ns; ++s) {
    const vgpu_sibc_desc &d = c->sibc[s].d;
    {
      const long long nfs = (long long)(d.gbbox[1] - d.gbbox[0]) * (d.gbbox[3] - d.gbbox[2]) * (d.gbbox[5] - d.gbbox[4]);
      for (long long f0 = 0; f0 < nfs; f0 += 128) blocks.push_back({s, 0, f + f0, f + nfs});
    }
    normal[s] = d.normal;
    set[s] = s;
    for (int p = 0; p < 16; ++p) {
      A[16 * s + p] = d.A[p / 4][p % 4];
      B[16 * s + p] = d.B[p / 4][p % 4];
      npl[16 * s + p] = d.npoles[p];
      asy[16 * s + p] = d.asymp[p];
      for (int kp = 0; kp < d.npoles[p]; ++kp)
        for (int r = 0; r < 2; ++r) {
          const size_t base = (((size_t)16 * s + p) * maxP + kp) * 6;
          q[base + 0 + r] = c->sibc[s].q0[p][2 * kp + r];
          q[base + 2 + r] = c->sibc[s].q1[p][2 * kp + r];
          q[base + 4 + r] = c->sibc[s].q2[p][2 * kp + r];
        }
    }
    auto pmc = [&](int dir, int idx) {
      return (d.pmc_index[dir][0] >= 0 && idx == d.pmc_index[dir][0]) ||
             (d.pmc_index[dir][1] >= 0 && idx == d.pmc_index[dir][1]);
    };
    // Faces are visited in the reference's order (k fastest), which fixes
    // the order of the H terms; they are stored x-fastest (then y, then z),
    // so a wavefront of faces of a y- or z-normal wall reads and writes
    // consecutive cells.
    long long lf = 0;
    const long long f0 = f;
    const int ni = d.gbbox[1] - d.gbbox[0], nj = d.gbbox[3] - d.gbbox[2];
    for (int i = d.gbbox[0]; i < d.gbbox[1]; ++i)
      for (int j = d.gbbox[2]; j < d.gbbox[3]; ++j)
        for (int k = d.gbbox[4]; k < d.gbbox[5]; ++k, ++lf) {
          if (i < 1 || j < 1 || k < 1 || i + 1 >= c->NX || j + 1 >= c->NY || k + 1 >= c->NZ)
            return set_error(c, VGPU_EINVAL, "SIBC face too close to the grid edge");
          const long long f =
              f0 + ((long long)(k - d.gbbox[4]) * nj + (j - d.gbbox[2])) * ni + (i - d.gbbox[0]);
          ijk[3 * f] = i; ijk[3 * f + 1] = j; ijk[3 * f + 2] = k;
          surf[f] = s;
          if (c->leap) {
            // the tangential E that k_sibc_e sets to zero on this face
            const long long id = c->lin(i, j, k), P = c->plane, SJ = c->PX;
            const int t1 = (d.normal + 1) % 3, t2 = (d.normal + 2) % 3;
            // each tangential component at the face cell and at its next
            // cell along the other tangential axis
            const long long o1 = d.normal == 0 ? P : d.normal == 1 ? 1 : SJ;
            const long long o2 = d.normal == 0 ? SJ : d.normal == 1 ? P : 1;
            c->eTouched.push_back(id * 3 + t1);
            c->eTouched.push_back((id + o1) * 3 + t1);
            c->eTouched.push_back(id * 3 + t2);
            c->eTouched.push_back((id + o2) * 3 + t2);
          }
          const unsigned char *a = &c->sibc[s].adjA[4 * lf], *b = &c->sibc[s].adjB[4 * lf];
          for (int r = 0; r < 4; ++r) { adjA[4 * f + r] = a[r]; adjB[4 * f + r] = b[r]; }
          // terms in sibc.c:638-757 order: (comp, di, dj, dk, sign, weight dir/index, Etan slot, id index)
          auto wbit = [](double w) { return w == 1.0 ? 1 : 0; };  // weights are 0.5 or 1
          auto edge = [&](int comp, int ci, int cj, int ck, double sign, double w, int slot, int idaxis, int idi) {
            SibcHEnt e{};
            e.face = (int)f;
            e.bits = 0 | slot << 1 | slot << 3 | (sign < 0) << 5 | wbit(w) << 6 | idaxis << 9;
            e.ia = e.ib = idi;
            hents.push_back({c->lin(ci, cj, ck) * 6 + comp, seq++, e});
          };
          auto adj = [&](int comp, int ci, int cj, int ck, double sign, double w, int sa, int sb, int idaxis,
                         int ia, int ib, int r) {
            SibcHEnt e{};
            e.face = (int)f;
            e.bits = 1 | sa << 1 | sb << 3 | (sign < 0) << 5 | wbit(w) << 6 | (a[r] ? 0 : 1) << 7 |
                     (b[r] ? 0 : 1) << 8 | idaxis << 9;
            e.ia = ia;
            e.ib = ib;
            hents.push_back({c->lin(ci, cj, ck) * 6 + comp, seq++, e});
          };
          if (d.normal == 0) {
            const double wyl = 0.5 * (1 + pmc(2, k)), wyh = 0.5 * (1 + pmc(2, k + 1));
            const double wzl = 0.5 * (1 + pmc(1, j)), wzh = 0.5 * (1 + pmc(1, j + 1));
            zeroKeys.push_back(c->lin(i, j, k) * 6 + 3);
            edge(4, i - 1, j, k, +1, wzl, 2, 0, i - 1);
            edge(4, i - 1, j + 1, k, +1, wzh, 2, 0, i - 1);
            edge(4, i, j, k, -1, wzl, 3, 0, i);
            edge(4, i, j + 1, k, -1, wzh, 3, 0, i);
            edge(5, i - 1, j, k, -1, wyl, 0, 0, i - 1);
            edge(5, i - 1, j, k + 1, -1, wyh, 0, 0, i - 1);
            edge(5, i, j, k, +1, wyl, 1, 0, i);
            edge(5, i, j, k + 1, +1, wyh, 1, 0, i);
            adj(3, i, j + 1, k, -1, wzh, 2, 3, 0, i - 1, i, 0);
            adj(3, i, j - 1, k, +1, wzl, 2, 3, 0, i - 1, i, 1);
            adj(3, i, j, k + 1, -1, wyh, 0, 1, 0, i - 1, i, 2);
            adj(3, i, j, k - 1, +1, wyl, 0, 1, 0, i - 1, i, 3);
          } else if (d.normal == 1) {
            const double wzl = 0.5 * (1 + pmc(0, i)), wzh = 0.5 * (1 + pmc(0, i + 1));
            const double wxl = 0.5 * (1 + pmc(2, k)), wxh = 0.5 * (1 + pmc(2, k + 1));
            zeroKeys.push_back(c->lin(i, j, k) * 6 + 4);
            edge(5, i, j - 1, k, +1, wxl, 2, 1, j - 1);
            edge(5, i, j - 1, k + 1, +1, wxh, 2, 1, j - 1);
            edge(5, i, j, k, -1, wxl, 3, 1, j);
            edge(5, i, j, k + 1, -1, wxh, 3, 1, j);
            edge(3, i, j - 1, k, -1, wzl, 0, 1, j - 1);
            edge(3, i + 1, j - 1, k, -1, wzh, 0, 1, j - 1);
            edge(3, i, j, k, +1, wzl, 1, 1, j);
            edge(3, i + 1, j, k, +1, wzh, 1, 1, j);
            // The B-side term of the first two uses the Ex slot (sibc.c:689-690).
            adj(4, i, j, k + 1, +1, wzh, 0, 3, 1, j - 1, j, 0);
            adj(4, i, j, k - 1, -1, wzl, 0, 3, 1, j - 1, j, 1);
            adj(4, i + 1, j, k, +1, wxh, 2, 3, 1, j - 1, j, 2);
            adj(4, i - 1, j, k, -1, wxl, 2, 3, 1, j - 1, j, 3);
          } else {
            const double wxl = 0.5 * (1 + pmc(1, j)), wxh = 0.5 * (1 + pmc(1, j + 1));
            const double wyl = 0.5 * (1 + pmc(0, i)), wyh = 0.5 * (1 + pmc(0, i + 1));
            zeroKeys.push_back(c->lin(i, j, k) * 6 + 5);
            edge(3, i, j, k - 1, +1, wyl, 2, 2, k - 1);
            edge(3, i + 1, j, k - 1, +1, wyh, 2, 2, k - 1);
            edge(3, i, j, k, -1, wyl, 3, 2, k);
            edge(3, i + 1, j, k, -1, wyh, 3, 2, k);
            edge(4, i, j, k - 1, -1, wxl, 0, 2, k - 1);
            edge(4, i, j + 1, k - 1, -1, wxh, 0, 2, k - 1);
            edge(4, i, j, k, +1, wxl, 1, 2, k);
            edge(4, i, j + 1, k, +1, wxh, 1, 2, k);
            adj(5, i + 1, j, k, +1, wyh, 2, 3, 2, k - 1, k, 0);
            adj(5, i - 1, j, k, -1, wyl, 2, 3, 2, k - 1, k, 1);
            adj(5, i, j + 1, k, +1, wxh, 0, 1, 2, k - 1, k, 2);
            adj(5, i, j - 1, k, -1, wxl, 0, 1, 2, k - 1, k, 3);
          }
        }
    f = f0 + lf;
  }
  std::stable_sort(hents.begin(), hents.end(), [](const HE &a, const HE &b) { return a.key < b.key; });
  std::sort(zeroKeys.begin(), zeroKeys.end());
  zeroKeys.erase(std::unique(zeroKeys.begin(), zeroKeys.end()), zeroKeys.end());
  std::vector<int> rstart;
  std::vector<long long> rkey;
  std::vector<SibcHEnt> ents;
  size_t e = 0, zq = 0;
  while (e < hents.size() || zq < zeroKeys.size()) {
    long long key;
    if (zq < zeroKeys.size() && (e >= hents.size() || zeroKeys[zq] <= hents[e].key)) key = zeroKeys[zq];
    else key = hents[e].key;
    const bool isZero = zq < zeroKeys.size() && zeroKeys[zq] == key;
    rkey.push_back(key / 6 << 3 | (isZero ? 4 : 0) | (key % 6 - 3));
    rstart.push_back((int)ents.size());
    while (e < hents.size() && hents[e].key == key) {
      if (!isZero) ents.push_back(hents[e].e);
      ++e;
    }
    if (isZero) ++zq;
  }
  rstart.push_back((int)ents.size());
  c->sibcHRows = (long long)rkey.size();
  c->d_sijk = upload_vec(c, ijk); c->d_ssurf = upload_vec(c, surf);
  c->d_sadjA = upload_vec(c, adjA); c->d_sadjB = upload_vec(c, adjB);
  c->d_snormal = upload_vec(c, normal); c->d_sA = upload_vec(c, A); c->d_sB = upload_vec(c, B);
  c->d_sset = upload_vec(c, set); c->d_snp = upload_vec(c, npl); c->d_sasymp = upload_vec(c, asy);
  c->d_sq = upload_vec(c, q);
  // With every coefficient real (real poles and residues, as in fitted
  // conductor slabs) each state's imaginary part stays exactly zero, so the
  // state is kept real: the same values, half the state traffic.
  c->sibcReal = true;
  for (size_t e = 1; e < q.size(); e += 2)
    if (q[e] != 0.0) c->sibcReal = false;
  // FP32 form with real coefficients: the filters fed by one input usually
  // share their poles (a fitted slab's Z11 and Z12).  A filter's recursive-
  // convolution coefficients q1, q2 are its residue times a pole-only factor,
  // so its state is its residue ratio to a base filter of the same input
  // times the base filter's state: one state per (input, pole) serves every
  // filter of that input.  The reordered rounding stays far inside the FP32
  // bar; the FP64 form keeps the reference's per-filter recursion.
  std::vector<int> fbase(4 * ns, -1);
  std::vector<double> fratio((size_t)16 * ns * maxP, 0.0);
  c->sibcFactored = c->prec == VGPU_FP32 && c->sibcReal;
  for (int st = 0; st < ns && c->sibcFactored; ++st)
    for (int in = 0; in < 4 && c->sibcFactored; ++in) {
      // base: the diagonal filter when every q1, q2 of it is non-zero (a
      // slab's Z11), else the first such filter of the input
      auto usable = [&](int flt) {
        const int np = npl[16 * st + flt];
        if (np == 0) return false;
        for (int kp = 0; kp < np; ++kp) {
          const double *qf = &q[(((size_t)16 * st + flt) * maxP + kp) * 6];
          if (qf[2] == 0.0 || qf[4] == 0.0) return false;
        }
        return true;
      };
      if (usable(5 * in)) fbase[4 * st + in] = 5 * in;
      for (int p = 0; p < 4 && fbase[4 * st + in] < 0; ++p)
        if (usable(4 * p + in)) fbase[4 * st + in] = 4 * p + in;
      for (int p = 0; p < 4; ++p) {
        const int flt = 4 * p + in, np = npl[16 * st + flt];
        if (np == 0) continue;
        const double *qf = &q[((size_t)16 * st + flt) * maxP * 6];
        if (fbase[4 * st + in] < 0) { c->sibcFactored = false; break; }
        const int b = fbase[4 * st + in];
        const double *qb = &q[((size_t)16 * st + b) * maxP * 6];
        if (np != npl[16 * st + b]) { c->sibcFactored = false; break; }
        for (int kp = 0; kp < np && c->sibcFactored; ++kp) {
          const double q0 = qf[6 * kp], q1 = qf[6 * kp + 2], q2 = qf[6 * kp + 4];
          const double b0 = qb[6 * kp], b1 = qb[6 * kp + 2], b2 = qb[6 * kp + 4];
          if (q0 != b0 || b1 == 0.0 || b2 == 0.0) { c->sibcFactored = false; break; }
          const double r1 = q1 / b1, r2 = q2 / b2;
          if (std::fabs(r1 - r2) > 1e-12 * std::max(std::max(std::fabs(r1), std::fabs(r2)), 1e-300)) {
            c->sibcFactored = false;
            break;
          }
          fratio[((size_t)16 * st + flt) * maxP + kp] = r1;
        }
      }
    }
  const size_t states = (size_t)(c->sibcFactored ? 4 : 16) * maxP * NF * (c->sibcReal ? 1 : 2);
  c->d_szeta = dev_alloc(c, states * sizeof(double));
  c->d_sold = dev_alloc(c, (size_t)4 * NF * sizeof(double));
  c->d_sfbase = upload_vec(c, fbase);
  c->d_sfratio = upload_vec(c, fratio);
  if (f != NF) return set_error(c, VGPU_EINVAL, "SIBC face count");
  c->d_sblocks = upload_vec(c, blocks);
  c->sibcBlocks = (int)blocks.size();
  c->d_setan = dev_alloc(c, (size_t)4 * NF * c->esz);
  c->d_shkey = upload_vec(c, rkey);
  c->d_shstart = upload_vec(c, rstart);
  c->d_shent = upload_vec(c, ents);
  c->d_shg = dev_alloc(c, rkey.size() * c->esz);
  if (!c->d_szeta || !c->d_sold || !c->d_setan || !c->d_shkey || !c->d_shstart || !c->d_shent || !c->d_shg)
    return set_error(c, VGPU_ENOMEM, "SIBC state allocation");
  return VGPU_OK;
}

// TF/SF face table: face, component, incident component, offset, sign
// (planewave.c:596-1003, eqns 5.48-5.59).
static const int TFSF_E[12][5] = {{2, 0, 5, -1, -1}, {2, 2, 3, -1, +1}, {3, 0, 5, 0, +1}, {3, 2, 3, 0, -1},
                                  {4, 0, 4, -1, +1}, {4, 1, 3, -1, -1}, {5, 0, 4, 0, -1}, {5, 1, 3, 0, +1},
                                  {0, 1, 5, -1, +1}, {0, 2, 4, -1, -1}, {1, 1, 5, 0, -1}, {1, 2, 4, 0, +1}};
static const int TFSF_H[12][5] = {{2, 5, 0, +1, -1}, {2, 3, 2, +1, +1}, {3, 5, 0, 0, +1}, {3, 3, 2, 0, -1},
                                  {4, 4, 0, +1, +1}, {4, 3, 1, +1, -1}, {5, 4, 0, 0, -1}, {5, 3, 1, 0, +1},
                                  {0, 5, 1, +1, +1}, {0, 4, 2, +1, -1}, {1, 5, 1, 0, -1}, {1, 4, 2, 0, +1}};

static int build_corrections(Ctx *c)
{
  std::vector<RowEntry> rows[2];
  long long seq = 0;
  const bool cubic = c->grid.grid_type == 0;
  // sources first (source.c:602-753)
  for (size_t s = 0; s < c->src.size(); ++s) {
    const Ctx::Src &src = c->src[s];
    const int e = src.field < 3 ? 0 : 1;
    const int comp = src.field % 3;
    const int *l = src.flim;
    for (int i = l[0]; i <= l[1]; ++i)
      for (int j = l[2]; j <= l[3]; ++j)
        for (int k = l[4]; k <= l[5]; ++k) {
          if (i < 0 || j < 0 || k < 0 || i >= c->NX || j >= c->NY || k >= c->NZ)
            return set_error(c, VGPU_EINVAL, "source outside the grid");
          CorrEntry ce{};
          ce.kind = src.kind;
          ce.idx = (int)s;
          rows[e].push_back({c->lin(i, j, k) * 3 + comp, seq++, ce});
        }
  }
  // then plane waves, face by face
  const std::vector<double> *idh[3] = {&c->idhx, &c->idhy, &c->idhz};
  const std::vector<double> *ide[3] = {&c->idex, &c->idey, &c->idez};
  for (int e = 0; e < 2; ++e)
    for (size_t p = 0; p < c->pw.size(); ++p) {
      const vgpu_pw_desc &d = c->pw[p].d;
      for (int q = 0; q < 12; ++q) {
        const int *t = e == 0 ? TFSF_E[q] : TFSF_H[q];
        const int face = t[0], comp = t[1], inc = t[2], off = t[3];
        if (!d.active[face]) continue;
        const int axis = face >> 1;
        const int *l = d.flim[face][comp];
        int lo[3] = {l[0], l[2], l[4]}, hi[3] = {l[1], l[3], l[5]};
        lo[axis] = hi[axis] = (face & 1) ? l[2 * axis + 1] : l[2 * axis];
        for (int i = lo[0]; i <= hi[0]; ++i)
          for (int j = lo[1]; j <= hi[1]; ++j)
            for (int k = lo[2]; k <= hi[2]; ++k) {
              const int n3[3] = {i, j, k};
              int s3[3] = {i, j, k};
              s3[axis] += off;
              if (i < 0 || j < 0 || k < 0 || i >= c->NX || j >= c->NY || k >= c->NZ || s3[axis] < 0 ||
                  s3[axis] >= (axis == 0 ? c->NX : axis == 1 ? c->NY : c->NZ))
                return set_error(c, VGPU_EINVAL, "plane-wave face outside the grid");
              CorrEntry ce{};
              ce.idx = (int)p;
              ce.inc = inc;
              ce.sign = t[4];
              ce.il = e == 0 ? (*idh[axis])[n3[axis]] : (*ide[axis])[n3[axis]];
              double r[3], dd;
              if (cubic && d.use_aux) {
                field_index_location(r, inc, s3[0], s3[1], s3[2]);
                dd = d.kinc[0] * (r[0] - d.ijk0[0]) + d.kinc[1] * (r[1] - d.ijk0[1]) + d.kinc[2] * (r[2] - d.ijk0[2]);
                ce.kind = 2;
                int id;
                double dp;
                if (inc < 3) {
                  id = (int)floor(dd);
                  dp = dd - id;
                  if (2 + id < 0 || 2 + id + 1 > d.aux_nx) return set_error(c, VGPU_EINVAL, "aux grid too short");
                } else {
                  id = (int)floor(dd + 0.5);
                  dp = dd + 0.5 - id;
                  if (1 + id < 0 || 2 + id > d.aux_nx) return set_error(c, VGPU_EINVAL, "aux grid too short");
                }
                ce.auxid = id;
                ce.geo = dp;
              } else {
                for (int a = 0; a < 3; ++a) {
                  const int m = s3[a] - c->grid.gibox[2 * a];
                  if (m < 0 || m >= (int)(a == 0 ? c->xl : a == 1 ? c->yl : c->zl).size())
                    return set_error(c, VGPU_EINVAL, "plane-wave face outside the mesh lines");
                }
                field_physical_location(c, r, inc, s3[0], s3[1], s3[2]);
                dd = d.kinc[0] * (r[0] - d.r0[0]) + d.kinc[1] * (r[1] - d.r0[1]) + d.kinc[2] * (r[2] - d.r0[2]);
                ce.kind = 3;
                ce.geo = dd / d.phase_velocity;
              }
              rows[e].push_back({c->lin(i, j, k) * 3 + comp % 3, seq++, ce});
            }
      }
    }
  if (c->leap)
    for (const RowEntry &r : rows[0]) c->eTouched.push_back(r.key);
  for (int e = 0; e < 2; ++e) {
    std::stable_sort(rows[e].begin(), rows[e].end(), [](const RowEntry &a, const RowEntry &b) { return a.key < b.key; });
    std::vector<int> rstart;
    std::vector<long long> rkey;
    std::vector<CorrEnt> ents;
    for (size_t q = 0; q < rows[e].size(); ++q) {
      if (q == 0 || rows[e][q].key != rows[e][q - 1].key) {
        rkey.push_back(rows[e][q].key / 3 << 2 | rows[e][q].key % 3);
        rstart.push_back((int)ents.size());
      }
      const CorrEntry &ce = rows[e][q].e;
      if (ce.idx < 0 || ce.idx >= (1 << 23)) return set_error(c, VGPU_EINVAL, "too many sources or plane waves");
      CorrEnt d{};
      d.code = ce.kind | ce.inc << 2 | (ce.sign < 0 ? 1 : 0) << 5 | ce.idx << 8;
      d.auxid = ce.auxid;
      d.geo = ce.geo;
      d.il = ce.il;
      ents.push_back(d);
    }
    rstart.push_back((int)ents.size());
    if (ents.size() >= (size_t)1 << 31) return set_error(c, VGPU_EINVAL, "too many correction entries");
    c->corrRows[e] = (long long)rkey.size();
    if (c->corrRows[e]) {
      c->d_ckey[e] = upload_vec(c, rkey);
      c->d_cstart[e] = upload_vec(c, rstart);
      c->d_cent[e] = upload_vec(c, ents);
      c->d_ccoef[e] = dev_alloc(c, rkey.size() * c->esz);
      c->d_calpha[e] = dev_alloc(c, rkey.size() * c->esz);
      if (!c->d_ckey[e] || !c->d_cstart[e] || !c->d_cent[e] || !c->d_ccoef[e] || !c->d_calpha[e])
        return set_error(c, VGPU_ENOMEM, "correction rows");
    }
  }
  // waveform, source and plane-wave tables
  std::vector<WaveformDev> wf = c->wf;
  for (size_t w = 0; w < wf.size(); ++w)
    if (wf[w].type == 9) {
      void *t = upload_vec(c, c->wftab[w]);
      wf[w].table = (const double *)t;
    }
  c->d_wf = upload_vec(c, wf);
  std::vector<SourceDev> sd;
  for (auto &s : c->src) sd.push_back({s.kind, s.soft, s.waveform, s.size, s.delay});
  c->d_src = upload_vec(c, sd);
  std::vector<PwDev> pd;
  int auxmax = 1;
  for (auto &p : c->pw) {
    PwDev x{};
    for (int q = 0; q < 6; ++q) x.Finc[q] = p.d.Finc[q];
    x.waveform = p.d.waveform; x.delay = p.d.delay;
    x.use_aux = cubic && p.d.use_aux; x.nx = p.d.aux_nx; x.xb = p.d.aux_xb;
    x.beta = p.d.beta_eyi; x.gamma = p.d.gamma_hzi;
    pd.push_back(x);
    auxmax = std::max(auxmax, p.d.aux_nx + 1);
  }
  c->auxmax = auxmax;
  c->d_pw = upload_vec(c, pd);
  if (!c->pw.empty()) {
    c->d_aux = dev_alloc(c, c->pw.size() * 2 * (size_t)auxmax * c->esz);
    c->d_auxpml = dev_alloc(c, c->pw.size() * 30 * c->esz);
    std::vector<double> prof;
    for (auto &p : c->pw)
      for (int q = 0; q < 4; ++q)
        for (int i = 0; i < 10; ++i) prof.push_back(p.prof[q][i]);
    c->d_auxprof = upload_reals(c, prof);
  }
  return VGPU_OK;
}

static int build_observers(Ctx *c)
{
  if (!c->pobs.empty()) {
    std::vector<long long> cell, off;
    std::vector<unsigned long long> first, last;
    std::vector<int> nf;
    std::vector<double> om;
    long long o6 = 0;
    for (auto &o : c->pobs) {
      cell.push_back(c->lin(o.i, o.j, o.k));
      first.push_back(o.first);
      last.push_back(o.last);
      nf.push_back((int)o.omega.size());
      off.push_back(o6);
      o6 += 6 * (long long)o.omega.size();
      om.insert(om.end(), o.omega.begin(), o.omega.end());
    }
    c->d_ocell = upload_vec(c, cell); c->d_ofirst = upload_vec(c, first); c->d_olast = upload_vec(c, last);
    c->d_onf = upload_vec(c, nf); c->d_ooff = upload_vec(c, off); c->d_oomega = upload_vec(c, om);
    c->d_otd = dev_alloc(c, c->pobs.size() * (size_t)std::max<uint64_t>(c->totalSteps, 1) * 6 * sizeof(double));
    c->d_odre = dev_alloc(c, (size_t)std::max<long long>(o6, 1) * sizeof(double));
    c->d_odim = dev_alloc(c, (size_t)std::max<long long>(o6, 1) * sizeof(double));
    if (!c->d_otd || !c->d_odre || !c->d_odim) return set_error(c, VGPU_ENOMEM, "observer buffers");
  }
  if (!c->vobs.empty()) {
    std::vector<long long> cell;
    std::vector<int> obs;
    std::vector<unsigned long long> first, last;
    for (size_t v = 0; v < c->vobs.size(); ++v) {
      const Ctx::VObs &o = c->vobs[v];
      c->vobsNodeStart.push_back((long long)cell.size());
      for (int k = o.box[4]; k <= o.box[5]; k += o.step[2])
        for (int j = o.box[2]; j <= o.box[3]; j += o.step[1])
          for (int i = o.box[0]; i <= o.box[1]; i += o.step[0]) {
            cell.push_back(c->lin(i, j, k));
            obs.push_back((int)v);
          }
      first.push_back(o.first);
      last.push_back(o.last);
    }
    c->vobsNodeStart.push_back((long long)cell.size());
    c->vnodes = (long long)cell.size();
    c->d_vcell = upload_vec(c, cell); c->d_vobs = upload_vec(c, obs);
    c->d_vfirst = upload_vec(c, first); c->d_vlast = upload_vec(c, last);
    c->vring = 2 * std::max<uint64_t>(c->maxRun, 1);
    const size_t ringBytes = (size_t)c->vnodes * 6 * c->vring * sizeof(float);
    c->d_vstage = dev_alloc(c, ringBytes);
    if (!c->d_vstage) return set_error(c, VGPU_ENOMEM, "volume observer staging");
    if (hipHostMalloc((void **)&c->h_vring, ringBytes, hipHostMallocDefault) != hipSuccess || !c->h_vring) {
      c->h_vring = nullptr;
      (void)hipGetLastError();
      return set_error(c, VGPU_ENOMEM, "pinned volume observer ring (%zu bytes)", ringBytes);
    }
  }
  return VGPU_OK;
}

// H cells of the fused box that the fused sweep cannot finish: those with an
// E input that is not an ordinary update (outside its component's inner-grid
// limits, so a UPML, Mur or ghost value) or that is changed after the E
// sweep (a Debye site, a source or TF/SF row, a zeroed SIBC face edge).  Hx
// reads Ey and Ez at its cell and Ez at j+1, Ey at k+1; Hy reads Ez, Ex and
// Ez at i+1, Ex at k+1; Hz reads Ex, Ey and Ey at i+1, Ex at j+1
// (grid.c:800-888).  They go to k_h_fix as 64-cell row chunks with masks.
static int build_leap(Ctx *c)
{
  if (!c->leap) return VGPU_OK;
  const int *fb = c->fastE;
  std::vector<unsigned char> ord(c->ncell, 0);
  for (int C = 0; C < 3; ++C) {
    const int *l = c->grid.gfilim[C];
    for (int k = std::max(l[4], 0); k <= std::min(l[5], c->NZ - 1); ++k)
      for (int j = std::max(l[2], 0); j <= std::min(l[3], c->NY - 1); ++j)
        for (int i = std::max(l[0], 0); i <= std::min(l[1], c->NX - 1); ++i) ord[c->lin(i, j, k)] |= 1u << C;
  }
  for (long long key : c->eTouched) {
    const long long cell = key / 3;
    if (cell >= 0 && cell < (long long)c->ncell) ord[cell] &= ~(1u << (key % 3));
  }
  std::vector<long long>().swap(c->eTouched);
  std::vector<FixEnt> ents;
  const long long P = c->plane, SJ = c->PX;
  long long cells = 0;
  for (int k = fb[4]; k <= fb[5]; ++k)
    for (int j = fb[2]; j <= fb[3]; ++j) {
      FixEnt cur{-1, j, k, 0, 0ull};
      for (int i = fb[0]; i <= fb[1]; ++i) {
        const long long id = c->lin(i, j, k);
        const bool ok = (ord[id] & 7u) == 7u && (ord[id + 1] & 6u) == 6u && (ord[id + SJ] & 5u) == 5u &&
                        (ord[id + P] & 3u) == 3u;
        if (ok) continue;
        const int i0 = i / 64 * 64;
        if (cur.i0 != i0) {
          if (cur.mask) ents.push_back(cur);
          cur.i0 = i0;
          cur.mask = 0;
        }
        cur.mask |= 1ull << (i - i0);
        ++cells;
      }
      if (cur.mask) ents.push_back(cur);
    }
  c->nFix = (long long)ents.size();
  c->fixCells = cells;
  if (!ents.empty()) {
    c->d_fix = upload_vec(c, ents);
    if (!c->d_fix) return set_error(c, VGPU_ENOMEM, "fused sweep fix-up list");
  }
  return VGPU_OK;
}

int vgpu_finalize(vgpu_ctx *c, uint64_t total_steps, uint64_t max_steps_per_run)
{
  if (!c || c->finalized) return VGPU_ESTATE;
  if (!c->media_set) return set_error(c, VGPU_ESTATE, "vgpu_set_media was not called");
  const DeviceCalls guard(g_capture_mu);
  c->totalSteps = total_steps;
  c->maxRun = max_steps_per_run ? max_steps_per_run : total_steps;
  c->d_step = dev_alloc(c, 64);
  if (!c->d_step) return set_error(c, VGPU_ENOMEM, "step counter");
  int r;
  if ((r = build_debye(c)) || (r = build_class_tables(c)) || (r = build_sibc(c)) || (r = build_corrections(c)) ||
      (r = build_observers(c)) || (r = build_leap(c)))
    return r;
  if ((r = c->prec == VGPU_FP64 ? prepare_sweeps<double>(c) : prepare_sweeps<float>(c))) return r;
  if (!c->leap && !c->debyeFused && (r = c->prec == VGPU_FP64 ? upload_pack<double>(c) : upload_pack<float>(c)))
    return r;
  // set-up copies are synchronous and its kernels ran on the solve's stream
  if (hipStreamSynchronize(c->stream) != hipSuccess || hipGetLastError() != hipSuccess)
    return set_error(c, VGPU_EDEVICE, "device error during set-up");
  c->finalized = true;
  return VGPU_OK;
}

// Capture a graph of `nsteps` whole steps (plus the counter advance) on the
// context's stream.  On any failure the context keeps launching directly.
static void capture_graph(Ctx *c, uint64_t nsteps)
{
  std::unique_lock<std::shared_mutex> alone(g_capture_mu);
  c->graphTried = true;
  hipStream_t s = c->stream;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  const int r = c->prec == VGPU_FP64 ? run_steps<double>(c, nsteps, s, false) : run_steps<float>(c, nsteps, s, false);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (r != VGPU_OK || e != hipSuccess || !g) {
    if (g) hipGraphDestroy(g);
    (void)hipGetLastError();
    return;
  }
  hipGraphExec_t x = nullptr;
  if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess || !x) {
    hipGraphDestroy(g);
    (void)hipGetLastError();
    return;
  }
  c->graph = g;
  c->graphExec = x;
  c->graphSteps = nsteps;
  c->graphPar = c->par;
}

static int check_run(Ctx *c, uint64_t first, uint64_t nsteps)
{
  if (!c || !c->finalized) return VGPU_ESTATE;
  if (nsteps > c->maxRun && c->vnodes) return set_error(c, VGPU_EINVAL, "run longer than the observer staging");
  if (!c->pobs.empty() && first + nsteps > c->totalSteps) return set_error(c, VGPU_EINVAL, "run past total steps");
  return VGPU_OK;
}

// A run's start events on stream s.
static void open_run(Ctx *c, hipStream_t s)
{
  hipEventRecord(c->ev0, s);
  hipEventRecord(c->runs[c->runCount & 1].start, s);
}

// The step counter, on stream s, when a run does not continue the last one.
static void set_step(Ctx *c, uint64_t first, hipStream_t s)
{
  if (c->stepHost != first)
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, s, (unsigned long long *)c->d_step,
                       (unsigned long long)first);
}

// A run's close on stream s, after its kernels: its volume samples to its
// slot of the pinned ring (rows wrap at the end of the device ring), its end
// events and its record.
static void close_run(Ctx *c, uint64_t first, uint64_t nsteps, hipStream_t s)
{
  Ctx::RunRec &rr = c->runs[c->runCount & 1];
  if (c->vnodes && nsteps) {
    const size_t row = (size_t)c->vnodes * 6;
    float *dst = c->h_vring + (size_t)(c->runCount & 1) * c->maxRun * row;
    const uint64_t r0 = first % c->vring, n1 = std::min<uint64_t>(nsteps, c->vring - r0);
    hipMemcpyAsync(dst, (float *)c->d_vstage + r0 * row, n1 * row * sizeof(float), hipMemcpyDeviceToHost, s);
    if (n1 < nsteps)
      hipMemcpyAsync(dst + n1 * row, c->d_vstage, (nsteps - n1) * row * sizeof(float), hipMemcpyDeviceToHost, s);
  }
  hipEventRecord(rr.end, s);
  hipEventRecord(c->ev1, s);
  rr.first = first;
  rr.n = nsteps;
  rr.valid = true;
  ++c->runCount;
  if (c->leap) c->par ^= (int)(nsteps & 1);  // the current fields are in the set the last step wrote
  c->stepHost = first + nsteps;
  c->lastRunFirst = first;
  c->lastRunSteps = nsteps;
}

int vgpu_run(vgpu_ctx *c, uint64_t first, uint64_t nsteps, void *stream)
{
  int r = check_run(c, first, nsteps);
  if (r) return r;
  hipSetDevice(c->device);
  (void)hipGetLastError();  // launch errors reported below belong to this run only
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  open_run(c, s);
  set_step(c, first, s);
  // Runs of the staging length are replayed from a graph captured once,
  // when the solve is long enough to replay it at least twice (a graph of an
  // even number of steps starts and ends on the same field set, so it
  // replays wherever that set is current); VGPU_GRAPHS=0 launches every run
  // directly.  The capture holds g_capture_mu alone, so no other thread
  // synchronizes the device while it runs.
  const char *ge = getenv("VGPU_GRAPHS");
  const bool graphable = !(ge && ge[0] == '0') && s == c->stream && c->evE.empty() && nsteps >= 2 &&
                         nsteps % 2 == 0 && nsteps == c->maxRun && c->totalSteps >= 2 * nsteps;
  if (graphable && !c->graphExec && !c->graphTried) capture_graph(c, nsteps);
  if (graphable && c->graphExec && c->graphSteps == nsteps && c->graphPar == c->par) {
    if (hipGraphLaunch(c->graphExec, s) != hipSuccess)
      r = set_error(c, VGPU_EDEVICE, "graph launch failed: %s", hipGetErrorString(hipGetLastError()));
    ++c->graphRuns;
    c->lastRunTimed = false;
  } else {
    r = c->prec == VGPU_FP64 ? run_steps<double>(c, nsteps, s, true) : run_steps<float>(c, nsteps, s, true);
    c->lastRunTimed = true;
  }
  close_run(c, first, nsteps, s);
  return r;
}

int vgpu_run_group(vgpu_ctx *const *ctxs, int n, const uint64_t *first, uint64_t nsteps)
{
  if (!ctxs || !first || n < 1 || n > VGPU_GROUP_MAX) return VGPU_EINVAL;
  for (int i = 0; i < n; ++i) {
    Ctx *c = ctxs[i];
    int r = check_run(c, first[i], nsteps);
    if (r) return r;
    if (!c->d_pack) return set_error(c, VGPU_EINVAL, "group runs need the separate sweeps");
    if (c->device != ctxs[0]->device || c->prec != ctxs[0]->prec)
      return set_error(c, VGPU_EINVAL, "group members differ in device or precision");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == c) return set_error(c, VGPU_EINVAL, "context twice in one group");
  }
  hipSetDevice(ctxs[0]->device);
  (void)hipGetLastError();
  // The group runs on the first member's stream, after the work already
  // queued on every member's stream; each member's stream then waits for the
  // group's end before its own later work.
  hipStream_t s = ctxs[0]->stream;
  for (int i = 0; i < n; ++i) {
    Ctx *c = ctxs[i];
    open_run(c, c->stream);
    if (i) hipStreamWaitEvent(s, c->runs[c->runCount & 1].start, 0);
  }
  for (int i = 0; i < n; ++i) set_step(ctxs[i], first[i], s);
  Ctx *cs[VGPU_GROUP_MAX];
  for (int i = 0; i < n; ++i) cs[i] = ctxs[i];
  // The second run in a row of the same group and length is captured into a
  // graph (held by the first member), and later ones replay it; a group that
  // changes is launched directly.  Serials identify members, so a graph never
  // outlives the contexts it names in a way that could match again.
  Ctx *lead = cs[0];
  std::vector<uint64_t> key(n);
  for (int i = 0; i < n; ++i) key[i] = cs[i]->serial;
  const char *ge = getenv("VGPU_GRAPHS");
  const bool same = key == lead->groupKey && nsteps == lead->groupSteps;
  // replaced graphs whose last launch has finished are freed; a graph still
  // in flight is retired behind an event instead of being freed under it
  for (size_t q = 0; q < lead->retired.size();) {
    Ctx::RetiredGraph &g = lead->retired[q];
    if (hipEventQuery(g.done) != hipSuccess) {
      ++q;
      continue;
    }
    hipGraphExecDestroy(g.exec);
    hipGraphDestroy(g.graph);
    hipEventDestroy(g.done);
    g = lead->retired.back();
    lead->retired.pop_back();
  }
  (void)hipGetLastError();
  if (!same) {
    if (lead->groupExec) {
      Ctx::RetiredGraph g{lead->groupGraph, lead->groupExec, nullptr};
      if (hipEventCreateWithFlags(&g.done, hipEventDisableTiming) == hipSuccess &&
          hipEventRecord(g.done, s) == hipSuccess) {
        lead->retired.push_back(g);
      } else {  // no event: wait for the stream before freeing the graph
        if (g.done) hipEventDestroy(g.done);
        const DeviceCalls guard(g_capture_mu);
        hipStreamSynchronize(s);
        hipGraphExecDestroy(g.exec);
        hipGraphDestroy(g.graph);
      }
    }
    lead->groupExec = nullptr;
    lead->groupGraph = nullptr;
    lead->groupKey = key;
    lead->groupSteps = nsteps;
    lead->groupTried = false;
  } else if (!lead->groupTried && !(ge && ge[0] == '0')) {
    std::unique_lock<std::shared_mutex> alone(g_capture_mu);
    lead->groupTried = true;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      const int rc = lead->prec == VGPU_FP64 ? run_group_steps<double>(cs, n, nsteps, s)
                                             : run_group_steps<float>(cs, n, nsteps, s);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &g);
      hipGraphExec_t x = nullptr;
      if (rc == VGPU_OK && e == hipSuccess && g && hipGraphInstantiate(&x, g, nullptr, nullptr, 0) == hipSuccess &&
          x) {
        lead->groupGraph = g;
        lead->groupExec = x;
      } else if (g) {
        hipGraphDestroy(g);
      }
    }
    (void)hipGetLastError();
  }
  int r = VGPU_OK;
  if (lead->groupExec) {
    if (hipGraphLaunch(lead->groupExec, s) != hipSuccess)
      r = set_error(lead, VGPU_EDEVICE, "graph launch failed: %s", hipGetErrorString(hipGetLastError()));
    ++lead->graphRuns;
  } else {
    r = lead->prec == VGPU_FP64 ? run_group_steps<double>(cs, n, nsteps, s) : run_group_steps<float>(cs, n, nsteps, s);
  }
  // the first member closes last, so its end follows all of the group's work
  // on its stream
  for (int k = 1; k <= n; ++k) {
    const int i = k % n;
    Ctx *c = ctxs[i];
    c->lastRunTimed = false;
    hipEvent_t end = c->runs[c->runCount & 1].end;
    close_run(c, first[i], nsteps, s);
    if (i) hipStreamWaitEvent(c->stream, end, 0);
  }
  return r;
}

int vgpu_synchronize(vgpu_ctx *c)
{
  if (!c) return VGPU_EINVAL;
  if (hipEventSynchronize(c->ev1) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
    return set_error(c, VGPU_EDEVICE, "device error: %s", hipGetErrorString(hipGetLastError()));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->last_total_ms = ms;
  double se = 0, sh = 0;
  int ne = 0;
  // only the pairs the last run recorded: querying an unrecorded event fails
  // and would leave a stale error behind for the next launch check
  const size_t rec = c->lastRunTimed ? 2 * std::min<size_t>(c->evE.size() / 2, (size_t)c->lastRunSteps) : 0;
  for (size_t q = 0; q + 1 < rec; q += 2) {
    float a = 0, b = 0;
    if (hipEventElapsedTime(&a, c->evE[q], c->evE[q + 1]) == hipSuccess &&
        hipEventElapsedTime(&b, c->evH[q], c->evH[q + 1]) == hipSuccess) {
      se += a; sh += b; ++ne;
    }
  }
  c->last_e_ms = ne ? se / ne : 0;
  c->last_h_ms = ne ? sh / ne : 0;
  return VGPU_OK;
}

int vgpu_query(vgpu_ctx *c)
{
  if (!c || !c->finalized) return VGPU_EINVAL;
  const hipError_t e = hipEventQuery(c->ev1);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return set_error(c, VGPU_EDEVICE, "device error: %s", hipGetErrorString(e));
}

int vgpu_read_point_observer(vgpu_ctx *c, int id, double *td, uint64_t nsteps, double *dft_re, double *dft_im)
{
  if (!c || !c->finalized || id < 0 || id >= (int)c->pobs.size()) return VGPU_EINVAL;
  const DeviceCalls guard(g_capture_mu);
  wait_last_run(c);
  if (td) {
    const uint64_t n = std::min<uint64_t>(nsteps, c->totalSteps);
    hipMemcpy(td, (char *)c->d_otd + (size_t)id * c->totalSteps * 6 * sizeof(double), n * 6 * sizeof(double),
              hipMemcpyDeviceToHost);
  }
  long long off = 0;
  for (int o = 0; o < id; ++o) off += 6 * (long long)c->pobs[o].omega.size();
  const size_t nb = 6 * c->pobs[id].omega.size() * sizeof(double);
  if (nb && dft_re) hipMemcpy(dft_re, (char *)c->d_odre + off * sizeof(double), nb, hipMemcpyDeviceToHost);
  if (nb && dft_im) hipMemcpy(dft_im, (char *)c->d_odim + off * sizeof(double), nb, hipMemcpyDeviceToHost);
  return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_EDEVICE;
}

int vgpu_read_volume_observer(vgpu_ctx *c, int id, float *out, uint64_t nsteps, uint64_t *nodes)
{
  if (!c || !c->finalized || id < 0 || id >= (int)c->vobs.size()) return VGPU_EINVAL;
  const DeviceCalls guard(g_capture_mu);
  wait_last_run(c);
  const long long n0 = c->vobsNodeStart[id], n1 = c->vobsNodeStart[id + 1];
  if (nodes) *nodes = (uint64_t)(n1 - n0);
  if (!out) return VGPU_OK;
  const uint64_t ns = std::min<uint64_t>(nsteps, c->lastRunSteps);
  for (uint64_t s = 0; s < ns; ++s)
    hipMemcpy(out + s * (n1 - n0) * 6,
              (float *)c->d_vstage + (((c->lastRunFirst + s) % c->vring) * c->vnodes + n0) * 6,
              (size_t)(n1 - n0) * 6 * sizeof(float), hipMemcpyDeviceToHost);
  return VGPU_OK;
}

int vgpu_run_result(vgpu_ctx *c, uint64_t first, int wait, double *ms, const float **samples, uint64_t *nsteps)
{
  if (!c || !c->finalized) return VGPU_EINVAL;
  // records are kept for the last two runs only; the later one wins a tie
  int slot = -1;
  for (int back = 1; back >= 0 && c->runCount; --back) {
    const int q = (int)((c->runCount - 1 - (uint64_t)back) & 1);
    if ((uint64_t)back < c->runCount && c->runs[q].valid && c->runs[q].first == first) slot = q;
  }
  if (slot < 0)
    return set_error(c, VGPU_EINVAL, "no record of a run that began at step %llu", (unsigned long long)first);
  const Ctx::RunRec &rr = c->runs[slot];
  hipError_t e = hipEventQuery(rr.end);
  if (e == hipErrorNotReady) {
    if (!wait) return 0;
    const DeviceCalls guard(g_capture_mu);
    e = hipEventSynchronize(rr.end);
  }
  if (e != hipSuccess) return set_error(c, VGPU_EDEVICE, "device error: %s", hipGetErrorString(e));
  if (ms) {
    float t = 0;
    *ms = hipEventElapsedTime(&t, rr.start, rr.end) == hipSuccess ? t : 0.0;
  }
  if (samples) *samples = c->vnodes ? c->h_vring + (size_t)slot * c->maxRun * c->vnodes * 6 : nullptr;
  if (nsteps) *nsteps = rr.n;
  return 1;
}

int vgpu_download_field(vgpu_ctx *c, int component, double *host)
{
  if (!c || component < 0 || component > 5 || !host) return VGPU_EINVAL;
  const DeviceCalls guard(g_capture_mu);
  wait_last_run(c);
  std::vector<unsigned char> buf(c->ncell * c->esz);
  if (hipMemcpy(buf.data(), c->fld(c->par, component), buf.size(), hipMemcpyDeviceToHost) != hipSuccess)
    return set_error(c, VGPU_EDEVICE, "download failed");
  for (int i = 0; i < c->NX; ++i)
    for (int j = 0; j < c->NY; ++j)
      for (int k = 0; k < c->NZ; ++k) {
        const size_t d = ((size_t)k * c->NY + j) * c->PX + i;
        host[((size_t)i * c->NY + j) * c->NZ + k] =
            c->esz == 8 ? ((double *)buf.data())[d] : (double)((float *)buf.data())[d];
      }
  return VGPU_OK;
}

static int check_limits_entry(vgpu_ctx *c, const int folim[6][6], vgpu_limits_report *out, unsigned char *counts,
                              void *scratch, size_t scratchBytes)
{
  if (!c || !folim || !out) return VGPU_EINVAL;
  if (!c->finalized) return VGPU_ESTATE;
  const int N[3] = {c->NX, c->NY, c->NZ};
  for (int C = 0; C < 6; ++C)
    for (int a = 0; a < 3; ++a)
      if (folim[C][2 * a] < 0 || folim[C][2 * a + 1] >= N[a])
        return set_error(c, VGPU_EINVAL, "field limits outside the grid");
  if (c->NY > 65535 || c->NZ > 65535)  // the tally launches one block row per (j, k)
    return set_error(c, VGPU_EINVAL, "update-once check: more than 65535 cells along y or z");
  const DeviceCalls guard(g_capture_mu);
  hipSetDevice(c->device);
  (void)hipGetLastError();
  return c->prec == VGPU_FP64 ? check_limits<double>(c, folim, out, counts, scratch, scratchBytes)
                              : check_limits<float>(c, folim, out, counts, scratch, scratchBytes);
}

int vgpu_check_limits(vgpu_ctx *c, const int folim[6][6], vgpu_limits_report *out, unsigned char *counts)
{
  return check_limits_entry(c, folim, out, counts, nullptr, 0);
}

int vgpu_check_limits_noalloc(vgpu_ctx *c, const int folim[6][6], vgpu_limits_report *out, unsigned char *counts)
{
  if (!c) return VGPU_EINVAL;
  if (!c->finalized) return VGPU_ESTATE;
  void *p = scratch_of_device(c->device);
  if (!p) return set_error(c, VGPU_EDEVICE, "no-allocation scratch of device %d not set up", c->device);
  return check_limits_entry(c, folim, out, counts, p, VGPU_NOALLOC_SCRATCH_BYTES);
}

size_t vgpu_noalloc_scratch_bytes(void) { return VGPU_NOALLOC_SCRATCH_BYTES; }

size_t vgpu_footprint_bytes(const vgpu_ctx *c) { return c ? c->bytes : 0; }

int vgpu_get_stats(const vgpu_ctx *c, vgpu_stats *out)
{
  if (!c || !out || !c->finalized) return VGPU_EINVAL;
  const int *go = c->grid.gobox, *gi = c->grid.gibox;
  out->cells = (uint64_t)(go[1] - go[0]) * (uint64_t)(go[3] - go[2]) * (uint64_t)(go[5] - go[4]);
  out->mesh_cells = (uint64_t)(gi[1] - gi[0]) * (uint64_t)(gi[3] - gi[2]) * (uint64_t)(gi[5] - gi[4]);
  out->debye_sites = (uint64_t)c->debyeSites;
  out->debye_poles = c->debye.empty() ? 0 : c->debyeMaxPoles;
  out->debye_complex = c->debyeCplx;
  out->sibc_faces = (uint64_t)c->sibcFaces;
  out->correction_rows[0] = (uint64_t)c->corrRows[0];
  out->correction_rows[1] = (uint64_t)c->corrRows[1];
  out->bytes = c->bytes;
  for (int e = 0; e < 2; ++e) {
    const int *fb = e == 0 ? c->fastE : c->fastH;
    out->fast_cells[e] = (fb[0] <= fb[1] && fb[2] <= fb[3] && fb[4] <= fb[5])
                             ? (uint64_t)(fb[1] - fb[0] + 1) * (uint64_t)(fb[3] - fb[2] + 1) * (uint64_t)(fb[5] - fb[4] + 1)
                             : 0;
  }
  out->debye_fused_sites = c->debyeFused ? (uint64_t)c->debyeL0 : 0;
  out->indexed_media = c->indexed ? 1 : 0;
  out->debye_fused = c->debyeFused ? 1 : 0;
  out->sibc_state = c->sibcFactored ? 2 : c->sibcReal ? 1 : 0;
  out->leap = c->leap ? 1 : 0;
  out->fix_cells = (uint64_t)c->fixCells;
  out->fix_chunks = (uint64_t)c->nFix;
  out->graph_runs = c->graphRuns;
  return VGPU_OK;
}

int vgpu_last_run_timing(const vgpu_ctx *c, double *total_ms, double *volume_e_ms, double *volume_h_ms)
{
  if (!c) return VGPU_EINVAL;
  if (total_ms) *total_ms = c->last_total_ms;
  if (volume_e_ms) *volume_e_ms = c->last_e_ms;
  if (volume_h_ms) *volume_h_ms = c->last_h_ms;
  return VGPU_OK;
}

const char *vgpu_last_error(const vgpu_ctx *c) { return c ? c->err.c_str() : "null context"; }

void vgpu_destroy(vgpu_ctx *c)
{
  if (!c) return;
  const DeviceCalls guard(g_capture_mu);
  destroy_ctx(c);
}

}  // extern "C"

static void destroy_ctx(vgpu_ctx *c)
{
  hipSetDevice(c->device);
  if (c->stream) wait_last_run(c);
  for (void *p : c->allocs) hipFree(p);
  if (c->graphExec) hipGraphExecDestroy(c->graphExec);
  if (c->graph) hipGraphDestroy(c->graph);
  if (c->groupExec) hipGraphExecDestroy(c->groupExec);
  if (c->groupGraph) hipGraphDestroy(c->groupGraph);
  for (auto &g : c->retired) {  // the stream has drained (wait_last_run)
    hipGraphExecDestroy(g.exec);
    hipGraphDestroy(g.graph);
    hipEventDestroy(g.done);
  }
  for (auto e : c->evE) hipEventDestroy(e);
  for (auto e : c->evH) hipEventDestroy(e);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  for (auto &rr : c->runs) {
    if (rr.start) hipEventDestroy(rr.start);
    if (rr.end) hipEventDestroy(rr.end);
  }
  if (c->h_vring) hipHostFree(c->h_vring);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}
