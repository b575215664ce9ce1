// Kernel templates of the FDTD core, instantiated for double and for float,
// both compiled with -ffp-contract=off so every update performs the
// reference's operations in the reference's order.
//
// Volume sweeps: a fast kernel over the box where every component takes the
// ordinary update, and a shell kernel for the rest (see below).
#pragma once
#include "vgpu_internal.h"
#include <type_traits>

namespace vgpu {

#define VG_DEV __device__ __forceinline__
#define VG_HD __host__ __device__ __forceinline__
#define VG_GLOBAL __attribute__((address_space(1)))

// A pointer read from an argument block in memory is generic to the
// compiler, which then emits flat loads; each wait on a flat load drains the
// vector and LDS counters together.  Device arrays are global memory: saying
// so turns them into global loads, whose waits stay partial.
template <typename T> VG_DEV VG_GLOBAL T *gp(T *p) { return (VG_GLOBAL T *)p; }

// The 1-D edge-length and UPML profile arrays are never written by a kernel.
// Read through the constant address space, an element indexed by the plane
// number (uniform across a wavefront) is a scalar load into a scalar
// register instead of a vector load per lane.
#define VG_CONST __attribute__((address_space(4)))
template <typename T> VG_DEV const VG_CONST T *cp(const T *p) { return (const VG_CONST T *)p; }
template <typename T> VG_DEV const VG_CONST T *cp(const VG_GLOBAL T *p) { return (const VG_CONST T *)(const T *)p; }

// Coefficient class tables on the device are padded to 8 (E: alpha, beta
// per component) and 4 (H: gamma per component) entries per class, so a
// component's (alpha, beta) pair is one aligned 2-vector load and a class's
// three gammas one 4-vector load (two 2-vectors in FP64).
constexpr int ETAB_STRIDE = 8, HTAB_STRIDE = 4;
typedef float vf2 __attribute__((ext_vector_type(2)));
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef double vd2 __attribute__((ext_vector_type(2)));
template <typename T> struct Vec2 { using type = vd2; };
template <> struct Vec2<float> { using type = vf2; };

template <typename T> VG_DEV void load_pair(const T *p, T &a, T &b)
{
  const typename Vec2<T>::type v = *reinterpret_cast<const VG_GLOBAL typename Vec2<T>::type *>(gp(p));
  a = v.x;
  b = v.y;
}

template <typename T> VG_DEV void load_gammas(const T *p, T &g0, T &g1, T &g2)
{
  T u;
  load_pair(p, g0, g1);
  load_pair(p + 2, g2, u);
}
template <> VG_DEV void load_gammas<float>(const float *p, float &g0, float &g1, float &g2)
{
  const vf4 v = *reinterpret_cast<const VG_GLOBAL vf4 *>(gp(p));
  g0 = v.x;
  g1 = v.y;
  g2 = v.z;
}

VG_HD bool in_box(const int *b, int i, int j, int k)
{
  return i >= b[0] && i <= b[1] && j >= b[2] && j <= b[3] && k >= b[4] && k <= b[5];
}

template <typename T> VG_HD long long slab_index(const int *b, int i, int j, int k)
{
  long long ni = b[1] - b[0] + 1, nj = b[3] - b[2] + 1;
  return ((long long)(k - b[4]) * nj + (j - b[2])) * ni + (i - b[0]);
}

// Region of component C at (i,j,k) from the argument block in device memory:
// 6 = ordinary update (one of the NB boxes `in`), 0..5 = UPML slab (p is the
// cell's offset in it), -1 = not updated by the sweep.
template <int NB, typename A>
VG_HD int classify(const A &a, const int (*in)[6], int C, int i, int j, int k, long long &p)
{
  for (int b = 0; b < NB; ++b)
    if (in_box(in[b], i, j, k)) return 6;
  if (a.pmlmask)
    for (int r = 0; r < 6; ++r)
      if (((a.pmlmask >> r) & 1u) && in_box(a.pl[r][C], i, j, k)) {
        p = slab_index<int>(a.pl[r][C], i, j, k);
        return r;
      }
  return -1;
}

// Debye ADE correction of one site (debye.c:243-307); shared by the fused E
// sweep (layer 0) and k_debye (later layers).  g is the site, m its medium.
template <typename T>
VG_DEV T debye_site(T e, T beta, long long g, int m, T *lastE_, T *J_, long long stride, const double *dalpha_,
                    const double *dbeta_, const int *npoles_, int maxPoles, int cplx)
{
  VG_GLOBAL T *__restrict__ lastE = gp(lastE_);
  VG_GLOBAL T *__restrict__ J = gp(J_);
  const VG_GLOBAL double *__restrict__ dalpha = gp(dalpha_);
  const VG_GLOBAL double *__restrict__ dbeta = gp(dbeta_);
  const VG_GLOBAL int *__restrict__ npoles = gp(npoles_);
  const int np = npoles[m];
  const VG_GLOBAL double *da = dalpha + (long long)m * maxPoles * 2;
  const VG_GLOBAL double *db = dbeta + (long long)m * maxPoles * 2;
  if (!cplx) {
    T jsum = T(0);
    for (int p = 0; p < np; ++p) jsum = jsum + (T(1) + T(da[2 * p])) * J[p * stride + g];
    e = e - beta * jsum;
    const T de = e - lastE[g];
    for (int p = 0; p < np; ++p) {
      VG_GLOBAL T *Jp = &J[p * stride + g];
      *Jp = T(da[2 * p]) * *Jp + T(db[2 * p]) * de;
    }
  } else {
    T sr = T(0);
    for (int p = 0; p < np; ++p) {
      const T ar = T(1) + T(da[2 * p]), ai = T(da[2 * p + 1]);
      const T jr = J[(p * stride + g) * 2], ji = J[(p * stride + g) * 2 + 1];
      sr = sr + (ar * jr - ai * ji);
    }
    e = e - beta * sr;
    const T de = e - lastE[g];
    for (int p = 0; p < np; ++p) {
      VG_GLOBAL T *Jp = &J[(p * stride + g) * 2];
      const T jr = Jp[0], ji = Jp[1];
      const T dar = T(da[2 * p]), dai = T(da[2 * p + 1]), dbr = T(db[2 * p]), dbi = T(db[2 * p + 1]);
      Jp[0] = (dar * jr - dai * ji) + dbr * de;
      Jp[1] = (dar * ji + dai * jr) + dbi * de;
    }
  }
  lastE[g] = e;
  return e;
}

// ===========================================================================
// Volume sweeps.  Each half-step is two kernels over disjoint cells:
//  - the fast kernel over the fast box, where every component takes the
//    ordinary update (inner grid).  A 64x4 thread tile covers 64 x-cells by
//    4 y-rows and marches along z over a chunk of planes, carrying the plane
//    below (E) or above (H) in registers; its arguments are a few pointers
//    and integers passed by value, so they stay in scalar registers.
//  - the shell kernel over everything else (UPML slabs, Mur normal H, cells
//    of one component's inner grid only, ghost layers), planned on the host
//    as pieces of x runs (ShellPiece), reading its pointers from the argument
//    block in device memory.
// ===========================================================================

// XCD-aware block order.  Workgroups are dealt to the eight XCDs round-robin
// by launch index, and each XCD has its own L2, so consecutive launch indices
// would scatter neighbouring rows and planes over different L2s.  The grid is
// padded to a multiple of eight and launch index b runs logical block
// (b % 8) * (n / 8) + b / 8: every XCD sweeps one contiguous range of blocks
// (a z range of the domain), where neighbouring rows and planes share its L2.
VG_DEV unsigned xcd_block(unsigned b, unsigned n)
{
  const unsigned per = n / 8;
  return (b % 8) * per + b / 8;
}
VG_DEV unsigned xcd_block() { return xcd_block(blockIdx.x, gridDim.x); }

// Layer-0 Debye correction inside the fast E kernel (debye.c:243-307).  The
// chunk's first site comes from dbase (uniform over the wavefront, so rows
// without Debye sites skip this entirely); a site's number is that plus the
// count of lower lanes whose class carries a Debye medium for C, which is
// the site list's (component, cell) order.
template <typename T, int C>
VG_DEV void debye_fused(const FastE<T> &f, int bx, int j, int k, bool vx, unsigned cls, T be, T &e)
{
  const EArgs<T> &a = *f.ap;
  const long long row = ((long long)(C * f.NZ + k) * f.NY + j) * f.nchx + bx;
  const VG_GLOBAL int *dbase = gp(a.dbase);
  const int g0 = dbase[row];
  if (dbase[row + 1] == g0) return;
  const int ds = vx ? gp(a.edslot)[cls * 3 + C] : -1;
  const unsigned long long m = __ballot(ds >= 0);
  if (ds < 0) return;
  const long long g = g0 + __popcll(m & ((1ull << __lane_id()) - 1ull));
  e = debye_site<T>(e, be, g, ds, a.dlast, a.dJ, a.dstride, a.dalpha, a.dbeta, a.dnp, a.dmaxP, a.dcplx);
}

// One E component of a fast-box cell: its coefficients and old value, then
// the ordinary update (and, with FUSE, the first Debye layer).  All three
// components are computed before any is stored, so the stores do not hold
// back the next component's loads.
template <typename T, bool IDX, bool FUSE, int C>
VG_DEV T e_fast(const FastE<T> &f, int bx, int j, int k, bool vx, long long id, unsigned cls, T curl)
{
  T al, be;
  if (IDX) {
    load_pair(f.etab + cls * ETAB_STRIDE + 2 * C, al, be);
  } else {
    al = f.alpha[C][id];
    be = f.beta[C][id];
  }
  T e = al * f.Ein[C][id] + be * curl;
  if (FUSE) debye_fused<T, C>(f, bx, j, k, vx, cls, be, e);
  return e;
}

// Update-once check (the reference's CHECK_LIMITS build, grid.h:181-191):
// a sweep kernel instantiated with CHK counts its updates of each field cell
// instead of storing them, and leaves every field and auxiliary untouched.
VG_DEV void visit(unsigned *v, long long id) { atomicAdd(v + id, 1u); }

// Fast E tile: inner grid (grid.c:709-797) and, with FUSE, the first Debye
// layer.  L is the logical block (in XCD order).
template <typename T, bool IDX, bool FUSE, bool CHK>
VG_DEV void e_fast_tile(const FastE<T> &f, unsigned L)
{
  if (L >= (unsigned)(f.nchx * f.gy * f.gz)) return;
  const int bx = (int)(L % f.nchx), by = (int)((L / f.nchx) % f.gy), bz = (int)(L / (f.nchx * f.gy));
  const int i = f.box[0] + bx * 64 + (int)threadIdx.x;
  const int j = f.box[2] + by * (int)blockDim.y + (int)threadIdx.y;
  if (j > f.box[3]) return;  // whole wavefront
  const bool vx = i <= f.box[1];
  const int ic = vx ? i : f.box[1];
  const int kb = f.box[4] + bz * f.kchunk;
  const int ke = min(kb + f.kchunk, f.box[5] + 1);
  if (kb >= ke) return;
  const long long P = f.plane, SJ = f.PX;
  const T *__restrict__ Hx = f.H[0];
  const T *__restrict__ Hy = f.H[1];
  const T *__restrict__ Hz = f.H[2];
  const T idhx = f.idh[0][ic], idhy = f.idh[1][j];
  long long id = ((long long)kb * f.NY + j) * SJ + ic;
  // The fast box starts at index 1 or more on every axis, so the i-1, j-1
  // and k-1 neighbours are inside the allocation.
  T hxm = Hx[id - P], hym = Hy[id - P];
  for (int k = kb; k < ke; ++k, id += P) {
    const T hx = Hx[id], hy = Hy[id], hz = Hz[id];
    const T hzy = Hz[id - SJ], hzx = Hz[id - 1], hyx = Hy[id - 1], hxy = Hx[id - SJ];
    const unsigned cls = IDX ? (unsigned)f.ecls[id] : 0u;
    const T idhz = cp(f.idh[2])[k];
    const T e0 = e_fast<T, IDX, FUSE, 0>(f, bx, j, k, vx, id, cls, idhy * (hz - hzy) + idhz * (hym - hy));
    const T e1 = e_fast<T, IDX, FUSE, 1>(f, bx, j, k, vx, id, cls, idhz * (hx - hxm) + idhx * (hzx - hz));
    const T e2 = e_fast<T, IDX, FUSE, 2>(f, bx, j, k, vx, id, cls, idhx * (hy - hyx) + idhy * (hxy - hx));
    if (vx) {
      if (CHK) {
        visit(f.vis[0], id);
        visit(f.vis[1], id);
        visit(f.vis[2], id);
      } else {
        f.E[0][id] = e0;
        f.E[1][id] = e1;
        f.E[2][id] = e2;
      }
    }
    hxm = hx;
    hym = hy;
  }
}

template <typename T, bool IDX, bool FUSE, bool CHK = false>
__global__ __launch_bounds__(512) void k_e_fast(FastE<T> f)
{
  e_fast_tile<T, IDX, FUSE, CHK>(f, xcd_block());
}

// Fast H tile: inner grid (grid.c:800-888).
template <typename T, bool IDX, bool CHK> VG_DEV void h_fast_tile(const FastH<T> &f, unsigned L)
{
  if (L >= (unsigned)(f.nchx * f.gy * f.gz)) return;
  const int bx = (int)(L % f.nchx), by = (int)((L / f.nchx) % f.gy), bz = (int)(L / (f.nchx * f.gy));
  const int i = f.box[0] + bx * 64 + (int)threadIdx.x;
  const int j = f.box[2] + by * (int)blockDim.y + (int)threadIdx.y;
  if (j > f.box[3]) return;
  const bool vx = i <= f.box[1];
  const int ic = vx ? i : f.box[1];
  const int kb = f.box[4] + bz * f.kchunk;
  const int ke = min(kb + f.kchunk, f.box[5] + 1);
  if (kb >= ke) return;
  const long long P = f.plane, SJ = f.PX;
  const T *__restrict__ Ex = f.E[0];
  const T *__restrict__ Ey = f.E[1];
  const T *__restrict__ Ez = f.E[2];
  T *__restrict__ H0 = f.H[0];
  T *__restrict__ H1 = f.H[1];
  T *__restrict__ H2 = f.H[2];
  const T *__restrict__ I0 = f.Hin[0];
  const T *__restrict__ I1 = f.Hin[1];
  const T *__restrict__ I2 = f.Hin[2];
  const T idex = f.ide[0][ic], idey = f.ide[1][j];
  long long id = ((long long)kb * f.NY + j) * SJ + ic;
  // The fast box ends at index N-2 or less, so i+1, j+1 and k+1 are inside.
  T ex = Ex[id], ey = Ey[id];
  for (int k = kb; k < ke; ++k, id += P) {
    const T exp1 = Ex[id + P], eyp1 = Ey[id + P];
    const T ez = Ez[id], ezy = Ez[id + SJ], ezx = Ez[id + 1], exy = Ex[id + SJ], eyx = Ey[id + 1];
    const unsigned cls = IDX ? (unsigned)f.hcls[id] : 0u;
    const T idez = cp(f.ide[2])[k];
    T g0, g1, g2;
    if (IDX) {
      load_gammas(f.htab + cls * HTAB_STRIDE, g0, g1, g2);
    } else {
      g0 = f.gamma[0][id];
      g1 = f.gamma[1][id];
      g2 = f.gamma[2][id];
    }
    const T h0 = I0[id] + g0 * (idez * (eyp1 - ey) + idey * (ez - ezy));
    const T h1 = I1[id] + g1 * (idex * (ezx - ez) + idez * (ex - exp1));
    const T h2 = I2[id] + g2 * (idey * (exy - ex) + idex * (ey - eyx));
    if (vx) {
      if (CHK) {
        visit(f.vis[0], id);
        visit(f.vis[1], id);
        visit(f.vis[2], id);
      } else {
        H0[id] = h0;
        H1[id] = h1;
        H2[id] = h2;
      }
    }
    ex = exp1;
    ey = eyp1;
  }
}

template <typename T, bool IDX, bool CHK = false> __global__ __launch_bounds__(512) void k_h_fast(FastH<T> f)
{
  h_fast_tile<T, IDX, CHK>(f, xcd_block());
}

// ===========================================================================
// Fused leapfrog sweep (grid.c:709-797 then 800-888, in one pass).  A wave
// owns a tile of 63 x-cells by R rows of the fused box and marches over kz
// planes.  At plane k it computes new E of its tile's 64 x-cells on R + 1
// rows (its own cells, the first cell of the next tile and the first row of
// the next tile, which its H needs), stores E of the cells it owns, and then
// updates H of plane k-1 from the new E of planes k-1 and k.  Every input is
// an old field, never written during the step, so a neighbour's new E is
// recomputed here bit for bit as its owner computes it, without waiting for
// it.  The E plane above the chunk (plane ke) is computed once more as the
// first plane of the next chunk.  One E and one H update per cell use the
// operands and operation order of the two separate sweeps, so the results
// are the same bits.  H cells whose E inputs are changed after the E sweep
// are updated again by k_h_fix.
// ===========================================================================
template <typename T, bool IDX, int C>
VG_DEV void leap_ecoef(const LeapArgs<T> &a, unsigned cls, int id, T &al, T &be)
{
  if (IDX) {
    load_pair(a.etab + cls * ETAB_STRIDE + 2 * C, al, be);
  } else {
    al = a.alpha[C][id];
    be = a.beta[C][id];
  }
}

template <typename T, bool IDX>
VG_DEV void leap_gammas(const LeapArgs<T> &a, unsigned cls, int id, T &g0, T &g1, T &g2)
{
  if (IDX) {
    load_gammas(a.htab + cls * HTAB_STRIDE, g0, g1, g2);
  } else {
    g0 = a.gamma[0][id];
    g1 = a.gamma[1][id];
    g2 = a.gamma[2][id];
  }
}

template <bool B> using Flag = std::integral_constant<bool, B>;

template <typename T, bool IDX, int R, bool CHK = false>
__global__ __launch_bounds__(256) void k_leap(LeapArgs<T> a)
{
  const unsigned L = xcd_block();
  if (L >= (unsigned)(a.ntx * a.ngy * a.ntz)) return;
  const int tx = (int)(L % a.ntx), ty = (int)((L / a.ntx) % a.ngy), tz = (int)(L / (a.ntx * a.ngy));
  const int lane = (int)threadIdx.x;
  // the wavefront's rows are uniform: kept in scalar registers
  const int wy = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const int jb = a.box[2] + (ty * (int)blockDim.y + wy) * R;
  if (jb > a.box[3]) return;
  const int i = a.box[0] + 63 * tx + lane;
  // E is needed up to one cell past the box (the last owned cell's x
  // neighbour); lanes beyond read that cell again
  const int ic = min(i, a.box[1] + 1);
  const bool ownx = lane < 63 && i <= a.box[1];
  const int kb = a.box[4] + tz * a.kz;
  const int ke = min(kb + a.kz, a.box[5] + 1);
  const int jmax = a.box[3] + 1;
  // 32-bit cell offsets (the host keeps the allocation below 2^31 cells)
  const int P = (int)a.plane, SJ = a.PX;
  // Cells this lane does not own are computed but stored to the lane's slot
  // of the discard area past the end of every field array, so the plane's
  // code has no branches and its loads of all rows can be issued together.
  const int dslot = a.discard + ((((int)L * 4 + wy) & 1023) << 6 | lane);
  const VG_GLOBAL T *__restrict__ Ex0 = gp(a.Ein[0]);
  const VG_GLOBAL T *__restrict__ Ey0 = gp(a.Ein[1]);
  const VG_GLOBAL T *__restrict__ Ez0 = gp(a.Ein[2]);
  const VG_GLOBAL T *__restrict__ Hx = gp(a.Hin[0]);
  const VG_GLOBAL T *__restrict__ Hy = gp(a.Hin[1]);
  const VG_GLOBAL T *__restrict__ Hz = gp(a.Hin[2]);
  VG_GLOBAL T *__restrict__ Ex1 = gp(a.E[0]);
  VG_GLOBAL T *__restrict__ Ey1 = gp(a.E[1]);
  VG_GLOBAL T *__restrict__ Ez1 = gp(a.E[2]);
  VG_GLOBAL T *__restrict__ Hx1 = gp(a.H[0]);
  VG_GLOBAL T *__restrict__ Hy1 = gp(a.H[1]);
  VG_GLOBAL T *__restrict__ Hz1 = gp(a.H[2]);
  // Rows jb .. jb+R (row jb+R: the next tile's first row, E only); rows past
  // the box's next row read that row again.  The box ends at N-2 at most, so
  // every index below stays inside the allocation.
  int rowo[R + 1];
  bool own[R + 1];
#pragma unroll
  for (int r = 0; r <= R; ++r) {
    rowo[r] = min(jb + r, jmax) * SJ + ic;
    own[r] = ownx && r < R && jb + r <= a.box[3];
  }
  const T idhx = a.idh[0][ic], idex = a.ide[0][ic];
  // carried from plane k-1: old H, new E, the H class
  T hxm[R + 1], hym[R + 1], hzm[R + 1], ex[R + 1], ey[R + 1], ez[R + 1];
  unsigned hcm[R + 1];
#pragma unroll
  for (int r = 0; r <= R; ++r) {
    const int id = (kb - 1) * P + rowo[r];
    hxm[r] = Hx[id];
    hym[r] = Hy[id];
    hzm[r] = T(0);
    ex[r] = ey[r] = ez[r] = T(0);
    hcm[r] = 0u;
  }
  // One plane: new E of plane k on all rows (stored when SE), then H of
  // plane k-1 (when DH).
  auto plane = [&](int k, auto se, auto dh) {
    constexpr bool SE = decltype(se)::value, DH = decltype(dh)::value;
    const int pk = k * P;
    const T idhz = cp(a.idh[2])[k];
    const T idez = cp(a.ide[2])[DH ? k - 1 : k];
    T hx[R + 1], hy[R + 1], hz[R + 1], hzx[R + 1], hyx[R + 1], e0[R + 1], e1[R + 1], e2[R + 1];
    unsigned ecl[R + 1], hcl[R + 1];
    const T hzy0 = Hz[pk + rowo[0] - SJ], hxy0 = Hx[pk + rowo[0] - SJ];
#pragma unroll
    for (int r = 0; r <= R; ++r) {
      const int id = pk + rowo[r];
      hx[r] = Hx[id];
      hy[r] = Hy[id];
      hz[r] = Hz[id];
      hzx[r] = Hz[id - 1];
      hyx[r] = Hy[id - 1];
      e0[r] = Ex0[id];
      e1[r] = Ey0[id];
      e2[r] = Ez0[id];
      ecl[r] = IDX ? (unsigned)a.ecls[id] : 0u;
      hcl[r] = IDX ? (unsigned)a.hcls[id] : 0u;
    }
    // new E of plane k (grid.c:709-797)
    T n0[R + 1], n1[R + 1], n2[R + 1];
#pragma unroll
    for (int r = 0; r <= R; ++r) {
      const int id = pk + rowo[r];
      const T idhy = cp(a.idh[1])[min(jb + r, jmax)];
      const T hzy = r == 0 ? hzy0 : hz[r == 0 ? 0 : r - 1];
      const T hxy = r == 0 ? hxy0 : hx[r == 0 ? 0 : r - 1];
      T al, be;
      leap_ecoef<T, IDX, 0>(a, ecl[r], id, al, be);
      n0[r] = al * e0[r] + be * (idhy * (hz[r] - hzy) + idhz * (hym[r] - hy[r]));
      leap_ecoef<T, IDX, 1>(a, ecl[r], id, al, be);
      n1[r] = al * e1[r] + be * (idhz * (hx[r] - hxm[r]) + idhx * (hzx[r] - hz[r]));
      leap_ecoef<T, IDX, 2>(a, ecl[r], id, al, be);
      n2[r] = al * e2[r] + be * (idhx * (hy[r] - hyx[r]) + idhy * (hxy - hx[r]));
    }
    if (DH) {
      // H of plane k-1 (grid.c:800-888): row r+1 of the carried state is
      // the new E of plane k-1 one row up, the next lane one cell along x
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const T ezx = __shfl_down(ez[r], 1), eyx = __shfl_down(ey[r], 1);
        const int idm = pk - P + rowo[r];
        const T idey = cp(a.ide[1])[min(jb + r, jmax)];
        T g0, g1, g2;
        leap_gammas<T, IDX>(a, hcm[r], idm, g0, g1, g2);
        const T h0 = hxm[r] + g0 * (idez * (n1[r] - ey[r]) + idey * (ez[r] - ez[r + 1]));
        const T h1 = hym[r] + g1 * (idex * (ezx - ez[r]) + idez * (ex[r] - n0[r]));
        const T h2 = hzm[r] + g2 * (idey * (ex[r + 1] - ex[r]) + idex * (ey[r] - eyx));
        if (CHK) {
          if (own[r]) {
            visit(a.vis[3], idm);
            visit(a.vis[4], idm);
            visit(a.vis[5], idm);
          }
        } else {
          const int d = own[r] ? idm : dslot;
          Hx1[d] = h0;
          Hy1[d] = h1;
          Hz1[d] = h2;
        }
      }
    }
    if (SE) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int id = pk + rowo[r];
        if (CHK) {
          if (own[r]) {
            visit(a.vis[0], id);
            visit(a.vis[1], id);
            visit(a.vis[2], id);
          }
        } else {
          const int d = own[r] ? id : dslot;
          Ex1[d] = n0[r];
          Ey1[d] = n1[r];
          Ez1[d] = n2[r];
        }
      }
    }
#pragma unroll
    for (int r = 0; r <= R; ++r) {
      ex[r] = n0[r];
      ey[r] = n1[r];
      ez[r] = n2[r];
      hxm[r] = hx[r];
      hym[r] = hy[r];
      hzm[r] = hz[r];
      hcm[r] = hcl[r];
    }
  };
  plane(kb, Flag<true>(), Flag<false>());
  for (int k = kb + 1; k < ke; ++k) plane(k, Flag<true>(), Flag<true>());
  plane(ke, Flag<false>(), Flag<true>());
}

// H cells of the fused box updated again once E is final (after Debye,
// SIBC, sources, TF/SF and ghosts): the ordinary H update from the old H.
// One wavefront per entry; four entries per block.
template <typename T, bool IDX> __global__ __launch_bounds__(256) void k_h_fix(FixArgs<T> a)
{
  const long long e = (long long)blockIdx.x * blockDim.y + threadIdx.y;
  if (e >= a.n) return;
  const FixEnt en = a.ent[e];
  const int lane = (int)threadIdx.x;
  if (!((en.mask >> lane) & 1ull)) return;
  const int i = en.i0 + lane, j = en.j, k = en.k;
  const long long SJ = a.PX, P = a.plane;
  const long long id = ((long long)k * a.NY + j) * SJ + i;
  const T *__restrict__ Ex = a.E[0];
  const T *__restrict__ Ey = a.E[1];
  const T *__restrict__ Ez = a.E[2];
  const T ex = Ex[id], exy = Ex[id + SJ], exz = Ex[id + P];
  const T ey = Ey[id], eyx = Ey[id + 1], eyz = Ey[id + P];
  const T ez = Ez[id], ezx = Ez[id + 1], ezy = Ez[id + SJ];
  const T idex = a.ide[0][i], idey = a.ide[1][j], idez = a.ide[2][k];
  T g0, g1, g2;
  if (IDX) {
    load_gammas(a.htab + (unsigned)a.hcls[id] * HTAB_STRIDE, g0, g1, g2);
  } else {
    g0 = a.gamma[0][id];
    g1 = a.gamma[1][id];
    g2 = a.gamma[2][id];
  }
  const T h0 = a.Hin[0][id] + g0 * (idez * (eyz - ey) + idey * (ez - ezy));
  const T h1 = a.Hin[1][id] + g1 * (idex * (ezx - ez) + idez * (ex - exz));
  const T h2 = a.Hin[2][id] + g2 * (idey * (exy - ex) + idex * (ey - eyx));
  a.H[0][id] = h0;
  a.H[1][id] = h1;
  a.H[2][id] = h2;
}

// Shell sweeps.  A block is 256 cells of one shell piece's x-y plane
// marching over a few planes (ShellPiece): the rows of a piece share one
// sequence of x runs per component, so a lane finds its regions once, among a
// few uniform runs; the march carries the neighbouring plane's fields in
// registers, and each plane issues all its loads before its first store.  The
// launch order of the blocks (ShellWork) gives each XCD one contiguous eighth
// of every piece.
VG_DEV bool shell_column(const ShellPiece *pcs, const ShellWork *work, unsigned b, const ShellPiece *&pc, int &di,
                         int &dj, int &dk0, int &dk1)
{
  const ShellWork w = work[b];
  if (w.piece < 0) return false;
  pc = pcs + w.piece;
  const int zc = w.block / pc->nxyb, xyb = w.block - zc * pc->nxyb;
  const int l = xyb * 256 + (int)threadIdx.x;
  if (l >= pc->nxy) return false;
  dj = l / pc->nip;
  di = l - dj * pc->nip - pc->xo;
  if (di < 0 || di >= pc->ni) return false;
  dk0 = zc * pc->kc;
  dk1 = min(dk0 + pc->kc, pc->nk);
  return true;
}

// Region of component C at x = i of a piece: reg 6 ordinary, 0..5 UPML slab
// (p its slab index at plane dk of the piece, psz the step of p per plane,
// a/b the slab's auxiliaries); reg -1 not updated.
struct ColRegion {
  int reg;
  long long p, psz;
  void *a, *b;
};
// A lane's region in the kernels: reg as above, and for a UPML slab the
// 32-bit element offset q of the lane's current cell in the auxiliaries'
// pools (the slab's offset in the pool plus its slab index), advanced by psz
// per plane.  Three 32-bit values per component instead of a 64-bit index,
// stride and two 64-bit slab addresses keep the shell kernels' register
// count, and so their occupancy, lower.
struct LaneRegion {
  int reg, q, psz;
  template <typename T> VG_DEV LaneRegion(const ColRegion &r, const T *pool)
      : reg(r.reg), q(r.reg >= 0 && r.reg < 6 ? (int)(static_cast<const T *>(r.a) - pool + r.p) : 0),
        psz((int)r.psz)
  {
  }
};

template <int C> VG_HD ColRegion shell_region(const ShellPiece *pc, int i, int dj, int dk)
{
  ColRegion r{-1, 0, 0, nullptr, nullptr};
  for (int q = 0; q < pc->nrun[C]; ++q) {
    const ShellRun &rn = pc->run[C][q];
    if (i >= rn.xa && i <= rn.xb) {
      r.reg = rn.reg;
      r.p = rn.p0 + (i - rn.xa) + dj * rn.psy + dk * rn.psz;
      r.psz = rn.psz;
      r.a = rn.a;
      r.b = rn.b;
    }
  }
  return r;
}

// The E shell's arrays as global pointers, read from the argument block
// once, before the plane loop.
template <typename T> struct EShellPtrs {
  VG_GLOBAL T *E[3];
  VG_GLOBAL T *PP, *Pa;  // UPML auxiliary pools
  VG_GLOBAL const T *Ein[3], *H[3], *alpha[3], *beta[3], *ad[3], *bd[3], *ah[3], *ibh[3], *idh[3], *etab;
  VG_GLOBAL const uint16_t *ecls;
  // sep false: the step reads and writes one field set (no fused sweep), so
  // the old E is read through E itself; one pointer fewer per component
  // keeps FP32 k_e_shell at 78 VGPRs, six waves per SIMD, instead of 86
  VG_DEV explicit EShellPtrs(const EArgs<T> &a, bool sep)
  {
    PP = gp(a.PPpool);
    Pa = gp(a.Ppool);
    for (int d = 0; d < 3; ++d) {
      E[d] = gp(a.E[d]);
      Ein[d] = sep ? gp(a.Ein[d]) : E[d];
      H[d] = gp(a.H[d]);
      alpha[d] = gp(a.alpha[d]);
      beta[d] = gp(a.beta[d]);
      ad[d] = gp(a.ad[d]);
      bd[d] = gp(a.bd[d]);
      ah[d] = gp(a.ah[d]);
      ibh[d] = gp(a.ibh[d]);
      idh[d] = gp(a.idh[d]);
    }
    etab = gp(a.etab);
    ecls = gp(a.ecls);
  }
};

// One E component of a shell cell (grid.c:709-795 ordinary update;
// pml.c:859-1023 UPML split update with the axis roles of component C: P
// along axis C+1, E along C+2): loads, then the stores.  The UPML profile
// coefficients indexed by x or y are fixed over a thread's z march and load
// once (init); those indexed by z load per plane.
template <typename T, bool IDX, int C> struct EShellCell {
  static constexpr int A1 = (C + 1) % 3, A2 = (C + 2) % 3;
  T e, al, be, opp, op, cf[6];
  VG_DEV void coef(const EShellPtrs<T> &g, const int n3[3], bool zaxis)
  {
    if ((A1 == 2) == zaxis) {
      cf[0] = A1 == 2 ? cp(g.ad[A1])[n3[A1]] : g.ad[A1][n3[A1]];
      cf[1] = A1 == 2 ? cp(g.bd[A1])[n3[A1]] : g.bd[A1][n3[A1]];
    }
    if ((A2 == 2) == zaxis) {
      cf[2] = A2 == 2 ? cp(g.ad[A2])[n3[A2]] : g.ad[A2][n3[A2]];
      cf[3] = A2 == 2 ? cp(g.bd[A2])[n3[A2]] : g.bd[A2][n3[A2]];
    }
    if ((C == 2) == zaxis) {
      cf[4] = C == 2 ? cp(g.ibh[C])[n3[C]] : g.ibh[C][n3[C]];
      cf[5] = C == 2 ? cp(g.ah[C])[n3[C]] : g.ah[C][n3[C]];
    }
  }
  VG_DEV void init(const EShellPtrs<T> &g, const LaneRegion &rg, const int n3[3])
  {
    if (rg.reg >= 0 && rg.reg < 6) coef(g, n3, false);
  }
  VG_DEV void load(const EShellPtrs<T> &g, const LaneRegion &rg, const int n3[3], long long id, unsigned cls)
  {
    if (rg.reg < 0) return;
    e = g.Ein[C][id];
    if (IDX) {
      load_pair((const T *)(g.etab + cls * ETAB_STRIDE + 2 * C), al, be);
    } else {
      al = g.alpha[C][id];
      be = g.beta[C][id];
    }
    if (rg.reg < 6) {
      opp = g.PP[rg.q];
      op = g.Pa[rg.q];
      coef(g, n3, true);
    }
  }
  VG_DEV void store(const EShellPtrs<T> &g, const LaneRegion &rg, long long id, T curl) const
  {
    if (rg.reg == 6) {
      g.E[C][id] = al * e + be * curl;
    } else if (rg.reg >= 0) {
      const T ppn = al * opp + be * curl;
      g.PP[rg.q] = ppn;
      const T pn = cf[0] * op + cf[1] * (ppn - opp);
      g.Pa[rg.q] = pn;
      g.E[C][id] = cf[2] * e + cf[3] * cf[4] * (pn - cf[5] * op);
    }
  }
};

template <typename T, bool IDX, bool CHK, bool SEP>
VG_DEV void e_shell_column(const ShellPiece *pcs, const ShellWork *work, const EArgs<T> *ap, unsigned b)
{
  const ShellPiece *pc;
  int di, dj, dk0, dk1;
  if (!shell_column(pcs, work, b, pc, di, dj, dk0, dk1)) return;
  const EArgs<T> &a = *ap;
  const EShellPtrs<T> g(a, SEP);
  const int NY = a.NY;
  const long long P = a.plane, SJ = a.PX;
  const int i = pc->box[0] + di, j = pc->box[2] + dj;
  LaneRegion rg[3] = {LaneRegion(shell_region<0>(pc, i, dj, dk0), a.PPpool),
                      LaneRegion(shell_region<1>(pc, i, dj, dk0), a.PPpool),
                      LaneRegion(shell_region<2>(pc, i, dj, dk0), a.PPpool)};
  if (rg[0].reg < 0 && rg[1].reg < 0 && rg[2].reg < 0) return;
  const VG_GLOBAL T *Hx = g.H[0], *Hy = g.H[1], *Hz = g.H[2];
  // (the sweep never updates index 0 of any axis; the guards keep every read
  // inside the allocation regardless)
  const bool xm = i > 0, ym = j > 0;
  const T idh0 = g.idh[0][i], idh1 = g.idh[1][j];
  int k = pc->box[4] + dk0;
  long long id = ((long long)k * NY + j) * SJ + i;
  T hxm = k > 0 ? Hx[id - P] : T(0), hym = k > 0 ? Hy[id - P] : T(0);  // plane k-1, carried
  EShellCell<T, IDX, 0> c0;
  EShellCell<T, IDX, 1> c1;
  EShellCell<T, IDX, 2> c2;
  {
    const int n3[3] = {i, j, k};
    c0.init(g, rg[0], n3);
    c1.init(g, rg[1], n3);
    c2.init(g, rg[2], n3);
  }
  for (int dk = dk0; dk < dk1; ++dk, ++k, id += P) {
    const T hx = Hx[id], hy = Hy[id], hz = Hz[id];
    const T hzy = ym ? Hz[id - SJ] : T(0), hzx = xm ? Hz[id - 1] : T(0);
    const T hyx = xm ? Hy[id - 1] : T(0), hxy = ym ? Hx[id - SJ] : T(0);
    const T idh2 = cp(g.idh[2])[k];
    const unsigned cls = IDX ? (unsigned)g.ecls[id] : 0u;
    const int n3[3] = {i, j, k};
    c0.load(g, rg[0], n3, id, cls);
    c1.load(g, rg[1], n3, id, cls);
    c2.load(g, rg[2], n3, id, cls);
    if (CHK) {
      for (int C = 0; C < 3; ++C)
        if (rg[C].reg >= 0) visit(a.vis[C], id);
    } else {
      c0.store(g, rg[0], id, idh1 * (hz - hzy) + idh2 * (hym - hy));
      c1.store(g, rg[1], id, idh2 * (hx - hxm) + idh0 * (hzx - hz));
      c2.store(g, rg[2], id, idh0 * (hy - hyx) + idh1 * (hxy - hx));
    }
    hxm = hx;
    hym = hy;
    rg[0].q += rg[0].psz;
    rg[1].q += rg[1].psz;
    rg[2].q += rg[2].psz;
  }
}

template <typename T, bool IDX, bool CHK = false, bool SEP = true>
__global__ __launch_bounds__(256) void k_e_shell(const ShellPiece *pcs, const ShellWork *work, const EArgs<T> *ap)
{
  e_shell_column<T, IDX, CHK, SEP>(pcs, work, ap, blockIdx.x);
}

// The H shell's arrays as global pointers.
template <typename T> struct HShellPtrs {
  VG_GLOBAL T *H[3];
  VG_GLOBAL T *B;  // UPML auxiliary pool
  VG_GLOBAL const T *Hin[3], *E[3], *gamma[3], *ad[3], *ah[3], *bh[3], *ibd[3], *ide[3], *htab;
  VG_GLOBAL const uint16_t *hcls;
  VG_DEV explicit HShellPtrs(const HArgs<T> &a)
  {
    B = gp(a.Bpool);
    for (int d = 0; d < 3; ++d) {
      H[d] = gp(a.H[d]);
      Hin[d] = gp(a.Hin[d]);
      E[d] = gp(a.E[d]);
      gamma[d] = gp(a.gamma[d]);
      ad[d] = gp(a.ad[d]);
      ah[d] = gp(a.ah[d]);
      bh[d] = gp(a.bh[d]);
      ibd[d] = gp(a.ibd[d]);
      ide[d] = gp(a.ide[d]);
    }
    htab = gp(a.htab);
    hcls = gp(a.hcls);
  }
};

// One H component of a shell cell: ordinary update (inner grid or Mur normal
// H, grid.c:797-888) or UPML split update (pml.c:1026-1176): B along axis
// C+1, H along C+2.  Profile coefficients indexed by x or y load once.
template <typename T, bool IDX, int C> struct HShellCell {
  static constexpr int A1 = (C + 1) % 3, A2 = (C + 2) % 3;
  T h, g, ob, cf[6];
  VG_DEV void coef(const HShellPtrs<T> &q, const int n3[3], bool zaxis)
  {
    if ((A1 == 2) == zaxis) {
      cf[0] = A1 == 2 ? cp(q.ah[A1])[n3[A1]] : q.ah[A1][n3[A1]];
      cf[1] = A1 == 2 ? cp(q.bh[A1])[n3[A1]] : q.bh[A1][n3[A1]];
    }
    if ((A2 == 2) == zaxis) {
      cf[2] = A2 == 2 ? cp(q.ah[A2])[n3[A2]] : q.ah[A2][n3[A2]];
      cf[3] = A2 == 2 ? cp(q.bh[A2])[n3[A2]] : q.bh[A2][n3[A2]];
    }
    if ((C == 2) == zaxis) {
      cf[4] = C == 2 ? cp(q.ibd[C])[n3[C]] : q.ibd[C][n3[C]];
      cf[5] = C == 2 ? cp(q.ad[C])[n3[C]] : q.ad[C][n3[C]];
    }
  }
  VG_DEV void init(const HShellPtrs<T> &q, const LaneRegion &rg, const int n3[3])
  {
    if (rg.reg >= 0 && rg.reg < 6) coef(q, n3, false);
  }
  VG_DEV void load(const HShellPtrs<T> &q, const LaneRegion &rg, const int n3[3], long long id, unsigned cls)
  {
    if (rg.reg < 0) return;
    h = q.Hin[C][id];
    g = IDX ? q.htab[cls * HTAB_STRIDE + C] : q.gamma[C][id];
    if (rg.reg < 6) {
      ob = q.B[rg.q];
      coef(q, n3, true);
    }
  }
  VG_DEV void store(const HShellPtrs<T> &q, const LaneRegion &rg, long long id, T curl) const
  {
    if (rg.reg == 6) {
      q.H[C][id] = h + g * curl;
    } else if (rg.reg >= 0) {
      const T bn = cf[0] * ob + g * cf[1] * curl;
      q.B[rg.q] = bn;
      q.H[C][id] = cf[2] * h + cf[3] * cf[4] * (bn - cf[5] * ob);
    }
  }
};

template <typename T, bool IDX, bool CHK>
VG_DEV void h_shell_column(const ShellPiece *pcs, const ShellWork *work, const HArgs<T> *ap, unsigned b)
{
  const ShellPiece *pc;
  int di, dj, dk0, dk1;
  if (!shell_column(pcs, work, b, pc, di, dj, dk0, dk1)) return;
  const HArgs<T> &a = *ap;
  const HShellPtrs<T> q(a);
  const int NX = a.NX, NY = a.NY, NZ = a.NZ;
  const long long P = a.plane, SJ = a.PX;
  const int i = pc->box[0] + di, j = pc->box[2] + dj;
  LaneRegion rg[3] = {LaneRegion(shell_region<0>(pc, i, dj, dk0), a.Bpool),
                      LaneRegion(shell_region<1>(pc, i, dj, dk0), a.Bpool),
                      LaneRegion(shell_region<2>(pc, i, dj, dk0), a.Bpool)};
  if (rg[0].reg < 0 && rg[1].reg < 0 && rg[2].reg < 0) return;
  const VG_GLOBAL T *Ex = q.E[0], *Ey = q.E[1], *Ez = q.E[2];
  const bool xin = i + 1 < NX, yin = j + 1 < NY;
  const T ide0 = q.ide[0][i], ide1 = q.ide[1][j];
  int k = pc->box[4] + dk0;
  long long id = ((long long)k * NY + j) * SJ + i;
  T exc = Ex[id], eyc = Ey[id];  // this plane's Ex, Ey: carried from the previous plane's k+1 loads
  HShellCell<T, IDX, 0> c0;
  HShellCell<T, IDX, 1> c1;
  HShellCell<T, IDX, 2> c2;
  {
    const int n3[3] = {i, j, k};
    c0.init(q, rg[0], n3);
    c1.init(q, rg[1], n3);
    c2.init(q, rg[2], n3);
  }
  for (int dk = dk0; dk < dk1; ++dk, ++k, id += P) {
    const bool zin = k + 1 < NZ;
    const T exp1 = zin ? Ex[id + P] : T(0), eyp1 = zin ? Ey[id + P] : T(0);
    const T ez = Ez[id], ezy = yin ? Ez[id + SJ] : T(0), ezx = xin ? Ez[id + 1] : T(0);
    const T eyx = xin ? Ey[id + 1] : T(0), exy = yin ? Ex[id + SJ] : T(0);
    const T ide2 = cp(q.ide[2])[k];
    const unsigned cls = IDX ? (unsigned)q.hcls[id] : 0u;
    const int n3[3] = {i, j, k};
    c0.load(q, rg[0], n3, id, cls);
    c1.load(q, rg[1], n3, id, cls);
    c2.load(q, rg[2], n3, id, cls);
    if (CHK) {
      for (int C = 0; C < 3; ++C)
        if (rg[C].reg >= 0) visit(a.vis[C], id);
    } else {
      c0.store(q, rg[0], id, ide2 * (eyp1 - eyc) + ide1 * (ez - ezy));
      c1.store(q, rg[1], id, ide0 * (ezx - ez) + ide2 * (exc - exp1));
      c2.store(q, rg[2], id, ide1 * (exy - exc) + ide0 * (eyc - eyx));
    }
    exc = exp1;
    eyc = eyp1;
    rg[0].q += rg[0].psz;
    rg[1].q += rg[1].psz;
    rg[2].q += rg[2].psz;
  }
}

template <typename T, bool IDX, bool CHK = false>
__global__ __launch_bounds__(256) void k_h_shell(const ShellPiece *pcs, const ShellWork *work, const HArgs<T> *ap)
{
  h_shell_column<T, IDX, CHK>(pcs, work, ap, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Mur tangential E (mur.c:135-395), before the E sweep.  blockIdx.z selects
// the (face, component) slot; x/y of the launch span the face.
// ---------------------------------------------------------------------------
template <typename T, bool CHK = false> __global__ void k_mur_e(MurArgs<T> a)
{
  const int slot = blockIdx.z;
  const int c = a.comp[slot];
  if (c < 0) return;
  const int face = slot >> 1, axis = face >> 1;
  const int *b = a.box[slot];
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y * blockDim.y + threadIdx.y;
  // (u,v) run over the two axes other than the normal, lowest axis first.
  const int ax1 = axis == 0 ? 1 : 0, ax2 = axis == 2 ? 1 : 2;
  int ijk[3];
  ijk[axis] = (face & 1) ? b[2 * axis + 1] : b[2 * axis];
  ijk[ax1] = b[2 * ax1] + u;
  ijk[ax2] = b[2 * ax2] + v;
  if (ijk[ax1] > b[2 * ax1 + 1] || ijk[ax2] > b[2 * ax2 + 1]) return;
  const int step = (face & 1) ? -1 : 1;
  int n3[3] = {ijk[0], ijk[1], ijk[2]};
  n3[axis] += step;
  const long long SJ = a.PX, P = a.plane;
  const long long id = ((long long)ijk[2] * a.NY + ijk[1]) * SJ + ijk[0];
  const long long nb = ((long long)n3[2] * a.NY + n3[1]) * SJ + n3[0];
  const int i = n3[0], j = n3[1], k = n3[2];
  const T *Hx = a.H[0], *Hy = a.H[1], *Hz = a.H[2];
  T curl;
  if (c == 0) curl = a.idh[1][j] * (Hz[nb] - Hz[nb - SJ]) + a.idh[2][k] * (Hy[nb - P] - Hy[nb]);
  else if (c == 1) curl = a.idh[2][k] * (Hx[nb] - Hx[nb - P]) + a.idh[0][i] * (Hz[nb - 1] - Hz[nb]);
  else curl = a.idh[0][i] * (Hy[nb] - Hy[nb - 1]) + a.idh[1][j] * (Hx[nb - SJ] - Hx[nb]);
  const T *Eo = a.Ein[c];
  T al, be;
  if (a.ecls) {
    load_pair(a.etab + (unsigned)a.ecls[nb] * ETAB_STRIDE + 2 * c, al, be);
  } else {
    al = a.alpha[c][nb];
    be = a.beta[c][nb];
  }
  const T tmp = al * Eo[nb] + be * curl;
  if (CHK) visit(a.vis[c], id);
  else a.E[c][id] = Eo[nb] + a.zeta[face] * (tmp - Eo[id]);
}

// Fill n unsigned values with v (grid-stride).
static __global__ void k_fill_u32(unsigned *p, size_t n, unsigned v)
{
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) p[q] = v;
}

// Tally of one component's visit counts against its outer field limits lim
// (grid.c:951-1025 checks the same limits): cells inside that no sweep
// updated, cells updated more than once, cells outside that were updated;
// out[3] keeps the lowest (k*NY + j)*NX + i of an offending cell, and counts
// (when given) the visits per cell, x fastest, saturated at 255.
struct Lim6 { int v[6]; };
static __global__ void k_visit_tally(const unsigned *vis, int NX, int NY, int NZ, long long PX, Lim6 lim,
                                     unsigned long long *out, unsigned char *counts)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= NX) return;
  const unsigned v = vis[((long long)k * NY + j) * PX + i];
  const long long lin = ((long long)k * NY + j) * NX + i;
  if (counts) counts[lin] = (unsigned char)(v > 255u ? 255u : v);
  const bool in = i >= lim.v[0] && i <= lim.v[1] && j >= lim.v[2] && j <= lim.v[3] && k >= lim.v[4] && k <= lim.v[5];
  const bool missed = in && v == 0, repeated = v > 1, outside = !in && v > 0;
  if (missed) atomicAdd(out + 0, 1ull);
  if (repeated) atomicAdd(out + 1, 1ull);
  if (outside) atomicAdd(out + 2, 1ull);
  if (missed || repeated || outside) atomicMin(out + 3, (unsigned long long)lin);
}

// ---------------------------------------------------------------------------
// Debye ADE correction of one layer of sites (debye.c:243-307)
// ---------------------------------------------------------------------------
template <typename T> VG_DEV void debye_layer_site(const DebyeArgs<T> &a, long long base, long long n, long long s)
{
  if (s >= n) return;
  const long long g = base + s;
  const long long key = a.key[g];
  const int c = (int)(key & 3), m = (int)(key >> 42);
  const long long id = (key >> 2) & ((1LL << 40) - 1);
  T *E = a.E[c];
  E[id] = debye_site<T>(E[id], a.sbeta[g], g, m, a.lastE, a.J, a.stride, a.dalpha, a.dbeta, a.npoles, a.maxPoles,
                        a.cplx);
}

template <typename T> __global__ void k_debye(DebyeArgs<T> a)
{
  debye_layer_site(a, a.base, a.n, blockIdx.x * (long long)blockDim.x + threadIdx.x);
}

// Each Debye site's beta, gathered once: from the coefficient classes when
// there are classes, else from the per-cell array.
template <typename T>
__global__ void k_debye_gather(long long n, const long long *key, const T *const *beta, const uint16_t *ecls,
                               const T *etab, T *sbeta)
{
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int c = (int)(key[g] & 3);
  const long long id = (key[g] >> 2) & ((1LL << 40) - 1);
  sbeta[g] = etab ? etab[(unsigned)ecls[id] * ETAB_STRIDE + 2 * c + 1] : beta[c][id];
}

// ---------------------------------------------------------------------------
// SIBC: E pass per face (sibc.c:443-586) with FP64 recursive convolution
// (filter.c:309-327).
// ---------------------------------------------------------------------------
template <typename T> __global__ __launch_bounds__(128) void k_sibc_e(SibcEArgs<T> a)
{
  const SibcBlock w = a.blocks[blockIdx.x];  // one surface per block: s is uniform
  const long long f = w.f0 + threadIdx.x;
  if (f >= w.f1) return;
  const int i = a.ijk[3 * f], j = a.ijk[3 * f + 1], k = a.ijk[3 * f + 2];
  const int s = w.surf;
  // the surface's matrices and filter tables are the same for every face of
  // the block: read through the constant address space they are scalar loads
  const VG_CONST double *qtab = cp(a.q), *frat = cp(a.fratio), *asym = cp(a.asymp);
  const VG_CONST int *npol = cp(a.npoles), *fbas = cp(a.fbase);
  const int nrm = cp(a.normal)[s];
  const long long SJ = a.PX, P = a.plane;
  const long long id = ((long long)k * a.NY + j) * SJ + i;
  const unsigned char *A = a.adjA + 4 * f, *B = a.adjB + 4 * f;
  T Htan[4];
  const T *Hx = a.H[0], *Hy = a.H[1], *Hz = a.H[2];
  T *Ex = a.E[0], *Ey = a.E[1], *Ez = a.E[2];
  if (nrm == 0) {
    Ey[id] = T(0); Ey[id + P] = T(0); Ez[id] = T(0); Ez[id + SJ] = T(0);
    Htan[0] = T(0.5 * (1 + A[0] + A[1]) * (double)(Hy[id - 1] + Hy[id - 1 + SJ]));
    Htan[1] = T(0.5 * (1 + B[0] + B[1]) * (double)(Hy[id] + Hy[id + SJ]));
    Htan[2] = T(0.5 * (1 + A[2] + A[3]) * (double)(Hz[id - 1] + Hz[id - 1 + P]));
    Htan[3] = T(0.5 * (1 + B[2] + B[3]) * (double)(Hz[id] + Hz[id + P]));
  } else if (nrm == 1) {
    Ez[id] = T(0); Ez[id + 1] = T(0); Ex[id] = T(0); Ex[id + P] = T(0);
    Htan[0] = T(0.5 * (1 + A[0] + A[1]) * (double)(Hz[id - SJ] + Hz[id - SJ + P]));
    Htan[1] = T(0.5 * (1 + B[0] + B[1]) * (double)(Hz[id] + Hz[id + P]));
    Htan[2] = T(0.5 * (1 + A[2] + A[3]) * (double)(Hx[id - SJ] + Hx[id - SJ + 1]));
    Htan[3] = T(0.5 * (1 + B[2] + B[3]) * (double)(Hx[id] + Hx[id + 1]));
  } else {
    Ex[id] = T(0); Ex[id + SJ] = T(0); Ey[id] = T(0); Ey[id + 1] = T(0);
    Htan[0] = T(0.5 * (1 + A[0] + A[1]) * (double)(Hx[id - P] + Hx[id - P + 1]));
    Htan[1] = T(0.5 * (1 + B[0] + B[1]) * (double)(Hx[id] + Hx[id + 1]));
    Htan[2] = T(0.5 * (1 + A[2] + A[3]) * (double)(Hy[id - P] + Hy[id - P + SJ]));
    Htan[3] = T(0.5 * (1 + B[2] + B[3]) * (double)(Hy[id] + Hy[id + SJ]));
  }
  const VG_CONST double *Am = cp(a.A) + 16 * s, *Bm = cp(a.B) + 16 * s;
  T Hin[4], Eout[4];
  for (int p = 0; p < 4; ++p) {
    Hin[p] = T(0);
    for (int q = 0; q < 4; ++q) Hin[p] = Hin[p] + T(Am[4 * p + q]) * Htan[q];
  }
  const int set = cp(a.filtset)[s];
  const long long NF = a.nfaces;
  // every filter fed by input q saw the same input last step: one stored
  // previous input per (input, face), read before and written after all
  double xo4[4];
  bool fed[4] = {false, false, false, false};
  for (int q = 0; q < 4; ++q) {
    xo4[q] = a.old[q * NF + f];
    for (int p = 0; p < 4; ++p) fed[q] = fed[q] || npol[16 * set + 4 * p + q] > 0;
  }
  if (a.factored) {
    // one state per (input, pole), scaled by each filter's residue ratio
    double y16[16];
    for (int flt = 0; flt < 16; ++flt) y16[flt] = asym[16 * set + flt] * (double)Hin[flt & 3];
    for (int q = 0; q < 4; ++q) {
      const int b = fbas[4 * set + q];
      if (b < 0) continue;
      const double x = (double)Hin[q], xo = xo4[q];
      const int np = npol[16 * set + b];
      const VG_CONST double *qq = qtab + ((long long)(16 * set + b) * a.maxPoles) * 6;
      VG_GLOBAL double *zq = gp(a.zeta) + (long long)q * a.maxPoles * NF + f;
      // states in groups of eight: all loads of a group issued before its math
      for (int k0 = 0; k0 < np; k0 += 8) {
        double zs[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (k0 + u < np) zs[u] = zq[(long long)(k0 + u) * NF];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int kp = k0 + u;
          if (kp >= np) break;
          const double xi = (qq[6 * kp] * zs[u] + qq[6 * kp + 2] * xo) + qq[6 * kp + 4] * x;
          zq[(long long)kp * NF] = xi;
          for (int p = 0; p < 4; ++p) {
            const int flt = 4 * p + q;
            if (npol[16 * set + flt] > 0) y16[flt] += frat[((long long)(16 * set + flt)) * a.maxPoles + kp] * xi;
          }
        }
      }
    }
    for (int p = 0; p < 4; ++p) {
      Eout[p] = T(0);
      for (int q = 0; q < 4; ++q) {
        const int flt = 4 * p + q;
        if (npol[16 * set + flt] == 0 && asym[16 * set + flt] == 0.0) continue;
        Eout[p] = T((double)Eout[p] + y16[flt]);
      }
    }
  } else
  for (int p = 0; p < 4; ++p) {
    Eout[p] = T(0);
    for (int q = 0; q < 4; ++q) {
      const int flt = 4 * p + q;
      const int np = npol[16 * set + flt];
      const double asy = asym[16 * set + flt];
      if (np == 0 && asy == 0.0) continue;
      const double x = (double)Hin[q];
      double y = asy * x;
      if (np > 0) {
        const double xo = xo4[q];
        const VG_CONST double *qq = qtab + ((long long)(16 * set + flt) * a.maxPoles) * 6;
        if (a.real) {
          // real coefficients: the complex step below with zero imaginary
          // parts, which leave the same real part bit for bit
          VG_GLOBAL double *zf = gp(a.zeta) + (long long)flt * a.maxPoles * NF + f;
          for (int k0 = 0; k0 < np; k0 += 8) {
            double zs[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
              if (k0 + u < np) zs[u] = zf[(long long)(k0 + u) * NF];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int kp = k0 + u;
              if (kp >= np) break;
              const double nr = (qq[6 * kp] * zs[u] + qq[6 * kp + 2] * xo) + qq[6 * kp + 4] * x;
              zf[(long long)kp * NF] = nr;
              y = y + nr;
            }
          }
        } else
        for (int kp = 0; kp < np; ++kp) {
          double *z = &a.zeta[(((long long)flt * a.maxPoles + kp) * NF + f) * 2];
          const double q0r = qq[6 * kp], q0i = qq[6 * kp + 1], q1r = qq[6 * kp + 2], q1i = qq[6 * kp + 3];
          const double q2r = qq[6 * kp + 4], q2i = qq[6 * kp + 5];
          const double zr = z[0], zi = z[1];
          const double nr = ((q0r * zr - q0i * zi) + q1r * xo) + q2r * x;
          const double ni = ((q0r * zi + q0i * zr) + q1i * xo) + q2i * x;
          z[0] = nr;
          z[1] = ni;
          y = y + nr;
        }
      }
      Eout[p] = T((double)Eout[p] + y);
    }
  }
  for (int q = 0; q < 4; ++q)
    if (fed[q]) a.old[q * NF + f] = (double)Hin[q];
  T *Et = a.Etan + 4 * f;
  for (int p = 0; p < 4; ++p) {
    T v = T(0);
    for (int q = 0; q < 4; ++q) v = v + T(Bm[4 * p + q]) * Eout[q];
    Et[p] = v;
  }
}

// SIBC: gathered H corrections (sibc.c:616-776).
template <typename T> __global__ void k_sibc_h(SibcHArgs<T> a)
{
  const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (r >= a.nrows) return;
  const long long key = a.rowkey[r];
  T *H = a.H[key & 3];
  const long long id = key >> 3;
  if (key & 4) {
    H[id] = T(0);
    return;
  }
  T h = H[id];
  const T g = a.rowg[r];
  for (int e = a.rowstart[r]; e < a.rowstart[r + 1]; ++e) {
    const SibcHEnt en = a.ent[e];
    const int b = en.bits, ax = (b >> 9) & 3, ea = (b >> 1) & 3, eb = (b >> 3) & 3;
    const T *ide = ax == 0 ? a.ide[0] : ax == 1 ? a.ide[1] : a.ide[2];
    const T w = (b >> 6) & 1 ? T(1) : T(0.5);
    const T *Et = a.Etan + 4 * (long long)en.face;
    T c;
    if (!(b & 1)) c = (g * w) * (ide[en.ia] * Et[ea]);
    else c = ((g * T(0.5)) * w) * (T((b >> 7) & 1) * (ide[en.ia] * Et[ea]) + T((b >> 8) & 1) * (ide[en.ib] * Et[eb]));
    h = h + ((b >> 5) & 1 ? T(-1) : T(1)) * c;
  }
  H[id] = h;
}

// ---------------------------------------------------------------------------
// Waveforms (waveform.c:413-638) on the device, in double.
// ---------------------------------------------------------------------------
VG_DEV double wf_compact(double pi, double time, double width, double span)
{
  if (time <= 0.0) return 0.0;
  if (time < span * width)
    return 1.0 / 32.0 * (10.0 - 15.0 * cos(pi / width * time) + 6.0 * cos(2.0 * pi / width * time) -
                         cos(3.0 * pi / width * time));
  return span < 1.5 ? 1.0 : 0.0;
}

VG_DEV double wf_value(const WaveformDev &w, double t, double delay, double pi)
{
  const double tt = t - delay - w.delay;
  const double width = w.width;
  switch (w.type) {
  case 0:
  case 1: return w.size * exp(-0.5 * (pow(tt / width, 2.0)));
  case 2: return w.size * (-tt / width * exp(-0.5 * (pow(tt / width, 2.0))));
  case 3: return w.size * ((1.0 - pow(tt / width, 2)) * exp(-0.5 * (pow(tt / width, 2.0))));
  case 4: return w.size * (exp(-0.5 * (pow(tt / width, 2.0))) * sin(2.0 * pi * w.frequency * tt));
  case 5: return w.size * wf_compact(pi, tt, width, 2.0);
  case 6:
    if (tt <= 0.0 || tt >= 2.0 * width) return 0.0;
    return w.size * (1.0 / 32.0 * (15.0 * sin(pi / width * tt) - 12.0 * sin(2.0 * pi / width * tt) +
                                   3.0 * sin(3.0 * pi / width * tt)));
  case 7: return w.size * (wf_compact(pi, tt, width, 2.0) * sin(2.0 * pi * w.frequency * tt));
  case 8: return w.size * (wf_compact(pi, tt, width, 1.0) * sin(2.0 * pi * w.frequency * tt));
  case 9: {
    const double *T3 = w.table;
    const unsigned long n = w.n;
    if (n < 2 || tt < T3[0] || tt > T3[3 * (n - 1)]) return 0.0;
    unsigned long lo = 0, hi = n - 1;
    while (hi - lo > 1) {
      const unsigned long m = (hi + lo) >> 1;
      if (T3[3 * m] > tt) hi = m;
      else lo = m;
    }
    const double h = T3[3 * hi] - T3[3 * lo];
    const double aa = (T3[3 * hi] - tt) / h, bb = (tt - T3[3 * lo]) / h;
    return w.size * (aa * T3[3 * lo + 1] + bb * T3[3 * hi + 1] +
                     ((aa * aa * aa - aa) * T3[3 * lo + 2] + (bb * bb * bb - bb) * T3[3 * hi + 2]) * (h * h) / 6.0);
  }
  default: return 0.0;
  }
}

// ---------------------------------------------------------------------------
// 1-D auxiliary grids of the TF/SF plane waves (planewave.c:1315-1372).
// One block per plane wave; the blocks ride in other launches (k_head for E,
// the H corrections' k_corr for H), whose own work they do not touch.
// ---------------------------------------------------------------------------
template <typename T> VG_DEV void aux_e_block(const AuxArgs<T> &a, int p, double pi)
{
  const PwDev &pw = a.pw[p];
  if (!pw.use_aux) return;
  T *Eyi = a.aux + (long long)p * 2 * a.auxmax, *Hzi = Eyi + a.auxmax;
  T *PPy = a.pml + p * 30, *Py = PPy + 10;
  const T *adx = a.prof + p * 40, *bdx = adx + 10;
  const T beta = T(pw.beta);
  for (int i = 1 + threadIdx.x; i < pw.nx; i += blockDim.x) {
    if (i < pw.xb) {
      Eyi[i] = Eyi[i] + beta * (Hzi[i - 1] - Hzi[i]);
    } else {
      const int L = i - pw.xb;
      const T oldPP = PPy[L];
      PPy[L] = PPy[L] + beta * (Hzi[i - 1] - Hzi[i]);
      const T oldP = Py[L];
      Py[L] = Py[L] + (PPy[L] - oldPP);
      Eyi[i] = adx[L] * Eyi[i] + bdx[L] * (Py[L] - oldP);
    }
  }
  const double t = (double)(*a.step + (unsigned long long)a.soff) * a.dt;
  if (threadIdx.x == 0) Eyi[0] = T(wf_value(a.wf[pw.waveform], t, pw.delay, pi));
}

template <typename T> VG_DEV void aux_h_block(const AuxArgs<T> &a, int p)
{
  const PwDev &pw = a.pw[p];
  if (!pw.use_aux) return;
  T *Eyi = a.aux + (long long)p * 2 * a.auxmax, *Hzi = Eyi + a.auxmax;
  T *Bz = a.pml + p * 30 + 20;
  const T *ahx = a.prof + p * 40 + 20, *bhx = ahx + 10;
  const T gamma = T(pw.gamma);
  for (int i = threadIdx.x; i < pw.nx; i += blockDim.x) {
    if (i < pw.xb) {
      Hzi[i] = Hzi[i] + gamma * (Eyi[i] - Eyi[i + 1]);
    } else {
      const int L = i - pw.xb;
      const T oldB = Bz[L];
      Bz[L] = ahx[L] * Bz[L] + gamma * bhx[L] * (-Eyi[i + 1] + Eyi[i]);
      Hzi[i] = Hzi[i] + (Bz[L] - oldB);
    }
  }
}

// ---------------------------------------------------------------------------
// Sources and TF/SF corrections, one thread per touched cell, entries in
// reference order (source.c:602-753, planewave.c:576-1007).
// ---------------------------------------------------------------------------
// Row coefficients, gathered once: the row cell's beta (E) or gamma (H), and
// its alpha (E).
// Row keys hold the cell above bit `shift` and the component in bits 0-1.
template <typename T>
__global__ void k_corr_gather(long long nrows, const long long *key, int shift, const T *const *coef,
                              const T *const *alpha, T *rowcoef, T *rowalpha)
{
  const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int c = (int)(key[r] & 3);
  const long long id = key[r] >> shift;
  rowcoef[r] = coef[c][id];
  if (rowalpha) rowalpha[r] = alpha ? alpha[c][id] : T(1);
}

template <typename T> VG_DEV void corr_rows(const CorrArgs<T> &a, double pi, long long r, int soff)
{
  if (r >= a.nrows) return;
  const double t = ((double)(*a.step + (unsigned long long)soff) + a.half) * a.dt;
  const long long key = a.rowkey[r];
  const long long id = key >> 2;
  T *F = a.F[key & 3];
  T f = F[id];
  const T coef = a.rowcoef[r];
  const T alpha = a.rowalpha[r];
  for (int e = a.rowstart[r]; e < a.rowstart[r + 1]; ++e) {
    const CorrEnt en = a.ent[e];
    const int kind = en.code & 3, inc = (en.code >> 2) & 7, idx = en.code >> 8;
    if (kind <= 1) {
      const SourceDev &s = a.src[idx];
      const T source = T(s.size * wf_value(a.wf[s.waveform], t, s.delay, pi));
      if (kind == 0) f = T(s.soft) * f + source;
      else f = T(s.soft) * f - coef * source;
    } else {
      const PwDev &pw = a.pw[idx];
      T incv;
      if (kind == 2) {
        const T dp = T(en.geo);
        const T *Eyi = a.aux + (long long)idx * 2 * a.auxmax, *Hzi = Eyi + a.auxmax;
        T value;
        if (inc < 3) value = (T(1) - dp) * Eyi[2 + en.auxid] + dp * Eyi[2 + en.auxid + 1];
        else value = T(a.eta0) * ((T(1) - dp) * Hzi[1 + en.auxid] + dp * Hzi[2 + en.auxid]);
        incv = T(pw.Finc[inc]) * value;
      } else {
        incv = T(pw.Finc[inc]) * T(wf_value(a.wf[pw.waveform], t - en.geo, pw.delay, pi));
      }
      const T term = coef * (T(en.il) * incv);
      const T sign = (en.code >> 5) & 1 ? T(-1) : T(1);
      if (a.electric) f = alpha * f + sign * term;
      else f = f + sign * term;
    }
  }
  F[id] = f;
}

template <typename T> __global__ void k_corr(CorrArgs<T> a, double pi)
{
  corr_rows(a, pi, blockIdx.x * (long long)blockDim.x + threadIdx.x, a.soff);
}

// H corrections with the aux grids' H update in the blocks from corrBlocks
// on: it writes aux H, which the H corrections do not read (they read aux E).
template <typename T>
VG_DEV void corr_aux_block(const CorrArgs<T> &a, double pi, const AuxArgs<T> &xa, int corrBlocks, int b, int soff)
{
  if (b >= corrBlocks) aux_h_block(xa, b - corrBlocks);
  else corr_rows(a, pi, b * (long long)blockDim.x + threadIdx.x, soff);
}

template <typename T> __global__ void k_corr_aux(CorrArgs<T> a, double pi, AuxArgs<T> xa, int corrBlocks)
{
  corr_aux_block(a, pi, xa, corrBlocks, (int)blockIdx.x, a.soff);
}

// ---------------------------------------------------------------------------
// Observers: point samples + FP64 running DFT (observer.c:760-899) and volume
// samples for impulse.dat (observer.c:1164-1198).  One thread per
// (observer, frequency); thread f==0 also stores the time sample.
// ---------------------------------------------------------------------------
template <typename T> VG_DEV void observe_block(const ObsArgs<T> &a, int o)
{
  const unsigned long long n = *a.step + (unsigned long long)a.soff;
  if (n < a.first[o] || n > a.last[o]) return;
  const double t = (double)n * a.dt;
  const long long id = a.cell[o];
  double v[6];
  for (int c = 0; c < 6; ++c) v[c] = (double)a.F[c][id];
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f == 0) {
    double *td = a.td + ((long long)o * a.tdsteps + (n - a.first[o])) * 6;
    for (int c = 0; c < 6; ++c) td[c] = v[c];
  }
  const int nf = a.nfreq[o];
  if (f >= nf) return;
  const double ot = a.omega[a.dftoff[o] / 6 + f] * t;
  double si, co;
  sincos(ot, &si, &co);
  double *re = a.dre + a.dftoff[o], *im = a.dim + a.dftoff[o];
  for (int c = 0; c < 6; ++c) {
    re[c * nf + f] += v[c] * co;
    im[c * nf + f] -= v[c] * si;
  }
}

// Head of a step: the observers sample the fields (blocks y < nobs) and each
// plane wave's aux grid takes its E update (blocks y >= nobs, x = 0).  The
// aux E update reads the aux H of the previous step and is read only by the
// H corrections, so it may run before the E sweep instead of after it
// (planewave.c:576-590 runs it just before the E corrections).
template <typename T> VG_DEV void head_block(const ObsArgs<T> &oa, const AuxArgs<T> &xa, double pi, int y)
{
  if (y < oa.nobs) observe_block(oa, y);
  else if (blockIdx.x == 0) aux_e_block(xa, y - oa.nobs, pi);
}

template <typename T> __global__ void k_head(ObsArgs<T> oa, AuxArgs<T> xa, double pi)
{
  head_block(oa, xa, pi, (int)blockIdx.y);
}

template <typename T> VG_DEV void observe_volume_node(const VolObsArgs<T> &a, long long nd)
{
  if (nd >= a.nnodes) return;
  const int o = a.nodeobs[nd];
  const unsigned long long n = *a.step + (unsigned long long)a.soff;
  if (n < a.first[o] || n > a.last[o]) return;
  const long long id = a.nodecell[nd];
  float *out = a.stage + ((long long)(n % a.ring) * a.nnodes + nd) * 6;
  for (int c = 0; c < 6; ++c) out[c] = (float)a.F[c][id];
}

template <typename T> __global__ void k_observe_volume(VolObsArgs<T> a)
{
  observe_volume_node(a, blockIdx.x * (long long)blockDim.x + threadIdx.x);
}

// Ghost plane copy for one face (surface.c:569-812).
template <typename T> __global__ void k_ghost(GhostArgs<T> a)
{
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y * blockDim.y + threadIdx.y;
  const int ax1 = a.axis == 0 ? 1 : 0, ax2 = a.axis == 2 ? 1 : 2;
  int d3[3], s3[3];
  d3[ax1] = s3[ax1] = a.lo[ax1] + u;
  d3[ax2] = s3[ax2] = a.lo[ax2] + v;
  if (d3[ax1] > a.hi[ax1] || d3[ax2] > a.hi[ax2]) return;
  d3[a.axis] = a.dst;
  s3[a.axis] = a.src;
  const long long di = ((long long)d3[2] * a.NY + d3[1]) * a.PX + d3[0];
  const long long si = ((long long)s3[2] * a.NY + s3[1]) * a.PX + s3[0];
  a.F[di] = a.sign * a.F[si];
}

// ---------------------------------------------------------------------------
// Packed launches: one launch of a step kernel over several solves of a group
// run.  A block finds its solve among the map's block ranges (a scan over at
// most PACK_MAX uniform entries), reads that solve's argument block through
// the constant address space (scalar loads) and runs the same code its own
// launch would, with its block number in that launch; block ranges of the
// fast kernels are multiples of eight, so each solve keeps its XCD order.
// ---------------------------------------------------------------------------
VG_DEV int pack_member(const PackMap &m, unsigned b)
{
  int d = 0;
  while (d + 1 < m.n && b >= m.start[d + 1]) ++d;
  return d;
}

template <typename T> VG_DEV const VG_CONST PackArgs<T> *pack_args(const PackMap &m, int d)
{
  return cp(static_cast<const PackArgs<T> *>(m.args[d]));
}

// An argument block copied out of the constant address space, word by word.
template <typename S> VG_DEV S ld_args(const VG_CONST S *p)
{
  static_assert(sizeof(S) % 4 == 0, "argument blocks are whole words");
  S out;
  const VG_CONST unsigned *w = reinterpret_cast<const VG_CONST unsigned *>(p);
  unsigned *o = reinterpret_cast<unsigned *>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(S) / 4); ++i) o[i] = w[i];
  return out;
}

template <typename T, bool IDX> __global__ __launch_bounds__(512) void k_e_fast_pk(PackMap m)
{
  const int d = pack_member(m, blockIdx.x);
  const FastE<T> f = ld_args(&pack_args<T>(m, d)->fe);
  e_fast_tile<T, IDX, false, false>(f, xcd_block(blockIdx.x - m.start[d], m.start[d + 1] - m.start[d]));
}

template <typename T, bool IDX> __global__ __launch_bounds__(512) void k_h_fast_pk(PackMap m)
{
  const int d = pack_member(m, blockIdx.x);
  const FastH<T> f = ld_args(&pack_args<T>(m, d)->fh);
  h_fast_tile<T, IDX, false>(f, xcd_block(blockIdx.x - m.start[d], m.start[d + 1] - m.start[d]));
}

template <typename T, bool IDX> __global__ __launch_bounds__(256) void k_e_shell_pk(PackMap m)
{
  const int d = pack_member(m, blockIdx.x);
  const VG_CONST PackArgs<T> *a = pack_args<T>(m, d);
  e_shell_column<T, IDX, false, false>(a->epc, a->ewk, a->ea, blockIdx.x - m.start[d]);
}

template <typename T, bool IDX> __global__ __launch_bounds__(256) void k_h_shell_pk(PackMap m)
{
  const int d = pack_member(m, blockIdx.x);
  const VG_CONST PackArgs<T> *a = pack_args<T>(m, d);
  h_shell_column<T, IDX, false>(a->hpc, a->hwk, a->ha, blockIdx.x - m.start[d]);
}

// The Debye and correction kernels pick a component's array by an index read
// from memory; from a copy in registers that would be an indexed read of
// scratch, so they read their argument blocks in place.
template <typename T> VG_DEV const PackArgs<T> &pack_ref(const PackMap &m, int d)
{
  return *static_cast<const PackArgs<T> *>(m.args[d]);
}

template <typename T> __global__ void k_debye_pk(PackMap m)
{
  const int d = pack_member(m, blockIdx.x);
  debye_layer_site(pack_ref<T>(m, d).da, m.base[d], m.cnt[d],
                   (blockIdx.x - m.start[d]) * (long long)blockDim.x + threadIdx.x);
}

template <typename T> __global__ void k_corr_e_pk(PackMap m, double pi)
{
  const int d = pack_member(m, blockIdx.x);
  corr_rows(pack_ref<T>(m, d).ce, pi, (blockIdx.x - m.start[d]) * (long long)blockDim.x + threadIdx.x, m.soff);
}

// H corrections, and the aux grids' H update in each solve's blocks from
// cnt[d] on (none when the solve has no aux grid).
template <typename T> __global__ void k_corr_h_pk(PackMap m, double pi)
{
  const int d = pack_member(m, blockIdx.x);
  const PackArgs<T> &p = pack_ref<T>(m, d);
  corr_aux_block(p.ch, pi, p.xa, (int)m.cnt[d], (int)(blockIdx.x - m.start[d]), m.soff);
}

// Head of a step: block rows are the solves' own k_head rows; x blocks past
// a solve's own launch width return.
template <typename T> __global__ void k_head_pk(PackMap m, double pi)
{
  const int d = pack_member(m, blockIdx.y);
  const VG_CONST PackArgs<T> *p = pack_args<T>(m, d);
  if ((int)blockIdx.x >= p->headGx) return;
  ObsArgs<T> oa = ld_args(&p->oa);
  AuxArgs<T> xa = ld_args(&p->xa);
  oa.soff = xa.soff = m.soff;
  head_block(oa, xa, pi, (int)(blockIdx.y - m.start[d]));
}

template <typename T> __global__ void k_observe_volume_pk(PackMap m)
{
  const int d = pack_member(m, blockIdx.x);
  VolObsArgs<T> a = ld_args(&pack_args<T>(m, d)->va);
  a.soff = m.soff;
  observe_volume_node(a, (blockIdx.x - m.start[d]) * (long long)blockDim.x + threadIdx.x);
}

// Device step counter: set before a run that does not continue the previous
// one, advanced after every run (a one-lane vector store).
static __global__ void k_set_step(unsigned long long *step, unsigned long long v)
{
  if (threadIdx.x == 0) step[threadIdx.x] = v;
}

static __global__ void k_advance_step(unsigned long long *step, unsigned long long by)
{
  if (threadIdx.x == 0) step[threadIdx.x] = step[threadIdx.x] + by;
}

}  // namespace vgpu
